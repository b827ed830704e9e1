// Host-side hat-basis tables for the HIP library (small, built once per level).
//
// Closed-form element integrals of products of multilinear hat functions and
// their partial derivatives (reference: Include/Hat.inl:121-194,
// Include/Hat.scalars.inl:530-565), and the per-boundary-case rows of the
// scalar stiffness stencil (Include/Hat.scalars.inl:259-308).  Arithmetic is
// carried out in the same order as the reference so the tables equal the
// reference's bit for bit.
#pragma once
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

namespace epr
{
namespace hat
{
// 2^Dim corner index with coordinate 0 as the most significant bit.
template< int Dim > inline int CornerBit( int c , int d ){ return ( c>>( Dim-1-d ) ) & 1; }

inline double Integral1D( int p , int q )
{
	auto Prod = []( int first , int last ){ double f = 1 ; for( int i=first ; i<=last ; i++ ) f *= i ; return f; };
	if( p>q ) return Prod( 1 , q ) / Prod( p+1 , p+q+1 );
	else      return Prod( 1 , p ) / Prod( q+1 , p+q+1 );
}

// coefficient * prod_d x_d^p_d (1-x_d)^q_d (dx_d)^dd_d on one element
template< int Dim >
struct Mono
{
	double c = 0;
	int p[Dim] = {} , q[Dim] = {} , dd[Dim] = {};
	static Mono Hat( int corner )
	{
		Mono m ; m.c = 1;
		for( int d=0 ; d<Dim ; d++ ){ if( CornerBit< Dim >( corner , d ) ) m.p[d] = 1 ; else m.q[d] = 1; }
		return m;
	}
	Mono D( int d ) const
	{
		Mono m = *this;
		if( p[d] ){ m.c *= p[d] ; m.p[d]-- ; m.dd[d] = 1; }
		else if( q[d] ){ m.c *= -q[d] ; m.q[d]-- ; m.dd[d] = 1; }
		return m;
	}
	Mono operator*( const Mono &o ) const
	{
		Mono m = *this;
		for( int d=0 ; d<Dim ; d++ ) m.p[d] += o.p[d] , m.q[d] += o.q[d] , m.dd[d] += o.dd[d];
		m.c *= o.c;
		return m;
	}
	double integral( unsigned int res ) const
	{
		double I = c;
		for( int d=0 ; d<Dim ; d++ ) I *= Integral1D( p[d] , q[d] ) * std::pow( (double)res , (double)dd[d] ) / (double)res;
		return I;
	}
};

// stiffness[i1*C+i2] = int grad phi_i1 . grad phi_i2 over one element at resolution res
template< int Dim >
std::vector< double > StiffnessTable( unsigned int res )
{
	const int C = 1<<Dim;
	std::vector< double > t( C*C );
	for( int i1=0 ; i1<C ; i1++ ) for( int i2=0 ; i2<C ; i2++ )
	{
		Mono< Dim > a = Mono< Dim >::Hat( i1 ) , b = Mono< Dim >::Hat( i2 );
		double s = 0;
		for( int d=0 ; d<Dim ; d++ ) s += ( a.D( d )*b.D( d ) ).integral( res );
		t[i1*C+i2] = s;
	}
	return t;
}

// fs[((i1*C+i2)*Dim + a)*Dim + b] = int d_b phi_i1 * d_a phi_i2 (the matrix m(a,b) of Hat.scalars.inl:556-562)
template< int Dim >
std::vector< double > FullStiffnessTable( unsigned int res )
{
	const int C = 1<<Dim;
	std::vector< double > t( C*C*Dim*Dim );
	for( int i1=0 ; i1<C ; i1++ ) for( int i2=0 ; i2<C ; i2++ )
	{
		Mono< Dim > a = Mono< Dim >::Hat( i1 ) , b = Mono< Dim >::Hat( i2 );
		for( int d1=0 ; d1<Dim ; d1++ ) for( int d2=0 ; d2<Dim ; d2++ )
			t[ ( (i1*C+i2)*Dim + d2 )*Dim + d1 ] = 0. + ( a.D( d1 )*b.D( d2 ) ).integral( res );
	}
	return t;
}

// Boundary-case index of a function for radius-0 full stencils (Hat.scalars.inl:295-308).
template< int Dim > inline int CaseIndex( const int *F , int res )
{
	int idx = 0;
	for( int d=0 ; d<Dim ; d++ )
	{
		if     ( F[d]<=0   ) idx = idx*3 + F[d];
		else if( F[d]>=res ) idx = idx*3 + 2 + F[d] - res;
		else                 idx = idx*3 + 1;
	}
	return idx;
}

// Dense per-case rows of the scalar stiffness stencil: K[case*W + w] for the
// neighbour at 3^Dim window index w (first coordinate most significant).  Rows
// are built from the case's representative function exactly as the reference
// does, which at resolution 1 leaves every case but the origin's empty.
template< int Dim >
std::vector< double > StiffnessCaseRows( int res )
{
	int W = 1 ; for( int d=0 ; d<Dim ; d++ ) W *= 3;
	const int C = 1<<Dim;
	std::vector< double > st = StiffnessTable< Dim >( res );
	std::vector< double > rows( (size_t)W*W , 0. );
	int hi = std::min( 3 , res+1 );
	int f1c[Dim] = {};
	while( true )
	{
		int f1[Dim];
		for( int d=0 ; d<Dim ; d++ ) f1[d] = f1c[d]<=0 ? f1c[d] : ( f1c[d]==1 ? res/2 : res-2+f1c[d] );
		// row[ f2-f1 ] accumulated over elements e (lexicographic) then f2 (lexicographic)
		std::map< std::vector< int > , double > row;
		int elo[Dim] , ehi[Dim] , e[Dim];
		for( int d=0 ; d<Dim ; d++ ) elo[d] = std::max( 0 , f1[d]-1 ) , ehi[d] = std::min( res , f1[d]+1 ) , e[d] = elo[d];
		bool eok = true;
		for( int d=0 ; d<Dim ; d++ ) if( elo[d]>=ehi[d] ) eok = false;
		while( eok )
		{
			int flo[Dim] , fhi[Dim] , f2[Dim];
			for( int d=0 ; d<Dim ; d++ ) flo[d] = std::max( 0 , e[d] ) , fhi[d] = std::min( res+1 , e[d]+2 ) , f2[d] = flo[d];
			while( true )
			{
				int r1 = 0 , r2 = 0;
				std::vector< int > off( Dim );
				for( int d=0 ; d<Dim ; d++ ) r1 = r1*2 + ( f1[d]-e[d] ) , r2 = r2*2 + ( f2[d]-e[d] ) , off[d] = f2[d]-f1[d];
				row[ off ] += st[ r1*C+r2 ];
				int d = Dim-1;
				for( ; d>=0 ; d-- ){ if( ++f2[d]<fhi[d] ) break ; f2[d] = flo[d]; }
				if( d<0 ) break;
			}
			int d = Dim-1;
			for( ; d>=0 ; d-- ){ if( ++e[d]<ehi[d] ) break ; e[d] = elo[d]; }
			if( d<0 ) break;
		}
		int cse = CaseIndex< Dim >( f1 , res );
		for( int w=0 ; w<W ; w++ ) rows[ (size_t)cse*W + w ] = 0;
		for( auto &kv : row )
		{
			int w = 0;
			for( int d=0 ; d<Dim ; d++ ) w = w*3 + kv.first[d]+1;
			rows[ (size_t)cse*W + w ] = kv.second;
		}
		int d = Dim-1;
		for( ; d>=0 ; d-- ){ if( ++f1c[d]<hi ) break ; f1c[d] = 0; }
		if( d<0 ) break;
	}
	return rows;
}
} // namespace hat
} // namespace epr
