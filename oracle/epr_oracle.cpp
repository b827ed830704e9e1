// CPU oracle for the ExteriorPoissonRecon solve path (dense grid).
//
// TEST INFRASTRUCTURE ONLY: a from-scratch FP64 restatement of the reference
// algorithm that the tests use as the checker for the HIP library.  The
// product never links or calls this file.
//
// Design: every level d of the hierarchy is a dense (2^d+1)^Dim grid of hat
// functions.  Sparse operators (R, B) are stored as 3^Dim-neighbour columns.
// Wherever the reference's summation order is observable in the result, the
// loops below visit terms in the reference's order so that the oracle agrees
// with the reference bit for bit on the serial path; the citations give the
// reference line each loop follows.  Exceptions (equal to rounding only):
//   * the Dim x Dim symmetric eigenvalues of the noise estimate use cyclic
//     Jacobi instead of Eigen's SelfAdjointEigenSolver (GridSamples.inl:94);
//   * the final energies use plain sequential dot products instead of Eigen's
//     vectorised dot (SystemEnergy.h:493-495).

#include "epr_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace
{
std::string g_error;

template< int Dim > struct Ix
{
	int v[Dim];
	Ix(){ for( int d=0 ; d<Dim ; d++ ) v[d] = 0; }
	int &operator[]( int d ){ return v[d]; }
	int  operator[]( int d ) const { return v[d]; }
	// lexicographic, coordinate 0 most significant (Misha/RegularGrid.inl:123-130)
	bool operator<( const Ix &o ) const { for( int d=0 ; d<Dim ; d++ ){ if( v[d]<o.v[d] ) return true ; if( v[d]>o.v[d] ) return false; } return false; }
	bool operator==( const Ix &o ) const { for( int d=0 ; d<Dim ; d++ ) if( v[d]!=o.v[d] ) return false; return true; }
	bool operator!=( const Ix &o ) const { return !( *this==o ); }
	bool operator<=( const Ix &o ) const { return *this<o || *this==o; }
	Ix operator+( const Ix &o ) const { Ix r ; for( int d=0 ; d<Dim ; d++ ) r.v[d] = v[d]+o.v[d] ; return r; }
	Ix operator-( const Ix &o ) const { Ix r ; for( int d=0 ; d<Dim ; d++ ) r.v[d] = v[d]-o.v[d] ; return r; }
};

// Lexicographic visit of the box [lo,hi) with coordinate 0 outermost
// (Misha/RegularGrid.inl:198-228).
template< int Dim , typename F >
void ForBox( const Ix< Dim > &lo , const Ix< Dim > &hi , F f )
{
	for( int d=0 ; d<Dim ; d++ ) if( lo[d]>=hi[d] ) return;
	Ix< Dim > I = lo;
	while( true )
	{
		f( I );
		int d = Dim-1;
		for( ; d>=0 ; d-- ){ if( ++I[d]<hi[d] ) break ; I[d] = lo[d]; }
		if( d<0 ) return;
	}
}

template< int Dim > constexpr int Pow3(){ int p=1 ; for( int d=0 ; d<Dim ; d++ ) p *= 3 ; return p; }
template< int Dim > constexpr int CDim(){ return ( Dim*(Dim-1) )/2; }

// function index: coordinate 0 most significant (Hat.scalars.inl:475-481)
template< int Dim > size_t FIndex( const Ix< Dim > &I , int res ){ size_t i=0 ; for( int d=0 ; d<Dim ; d++ ) i = i*(res+1) + I[d] ; return i; }
template< int Dim > Ix< Dim > FIndex( size_t i , int res ){ Ix< Dim > I ; for( int d=Dim-1 ; d>=0 ; d-- ){ I[d] = (int)( i%(res+1) ) ; i /= res+1; } return I; }
// element index: coordinate 0 fastest (Hat.scalars.inl:455-473)
template< int Dim > size_t EIndex( const Ix< Dim > &I , int res ){ size_t i=0 ; for( int d=Dim-1 ; d>=0 ; d-- ) i = i*res + I[d] ; return i; }
template< int Dim > Ix< Dim > EIndex( size_t i , int res ){ Ix< Dim > I ; for( int d=0 ; d<Dim ; d++ ){ I[d] = (int)( i%res ) ; i /= res; } return I; }
// 3^Dim window index of an offset in {-1,0,1}^Dim, first coordinate most significant (Misha/Window.h:79-83)
template< int Dim > int W3( const Ix< Dim > &off ){ int w=0 ; for( int d=0 ; d<Dim ; d++ ) w = w*3 + off[d]+1 ; return w; }
template< int Dim > Ix< Dim > W3( int w ){ Ix< Dim > o ; for( int d=Dim-1 ; d>=0 ; d-- ){ o[d] = w%3-1 ; w /= 3; } return o; }
// 2^Dim corner index relative to an element, first coordinate most significant (Hat.inl:462-477)
template< int Dim > int Rel( const Ix< Dim > &e , const Ix< Dim > &f ){ int i=0 ; for( int d=0 ; d<Dim ; d++ ) i = i*2 + ( f[d]-e[d] ) ; return i; }

////////////////////////////////////////////////////////////////
// Element polynomials (Hat.h:110-137, Hat.inl:121-210, 240-248)
////////////////////////////////////////////////////////////////
double BasisIntegral( int p , int q ) // Hat.inl:121-131
{
	auto Prod = []( int first , int last ){ double f=1 ; for( int i=first ; i<=last ; i++ ) f *= i ; return f; };
	if( p>q ) return Prod(1,q)/Prod(p+1,p+q+1);
	else      return Prod(1,p)/Prod(q+1,p+q+1);
}

template< int Dim >
struct EFn
{
	double c;
	int p[Dim] , q[Dim] , dd[Dim];
	EFn( void ) : c(0){ for( int d=0 ; d<Dim ; d++ ) p[d] = q[d] = dd[d] = 0; }
	// the hat function of corner 'corner' (relative bits, coordinate 0 most significant) on the origin element (Hat.inl:136-147)
	static EFn Hat( int corner )
	{
		EFn f ; f.c = 1;
		for( int d=0 ; d<Dim ; d++ ){ int b = ( corner>>(Dim-1-d) )&1 ; if( b==0 ) f.q[d] = 1 ; else f.p[d] = 1; }
		return f;
	}
	EFn D( int d ) const // Hat.inl:149-169
	{
		EFn f = *this;
		if( p[d] ){ f.c *= p[d] ; f.p[d]-- ; f.dd[d] = 1; }
		else if( q[d] ){ f.c *= -q[d] ; f.q[d]-- ; f.dd[d] = 1; }
		return f;
	}
	EFn operator*( const EFn &g ) const { EFn f = *this ; for( int d=0 ; d<Dim ; d++ ) f.p[d] += g.p[d] , f.q[d] += g.q[d] , f.dd[d] += g.dd[d] ; f.c *= g.c ; return f; }
	EFn scaled( double s ) const { EFn f = *this ; f.c *= s ; return f; }
	double integral( unsigned int res ) const // Hat.inl:188-194
	{
		double I = c;
		for( int d=0 ; d<Dim ; d++ ) I *= BasisIntegral( p[d] , q[d] ) * pow( (double)res , (double)dd[d] ) / (double)res;
		return I;
	}
};

// Skew element product of two gradient vectors (Hat.inl:243-248): _c[0][idx] = v1[j]*v2[i]/2 , _c[1][idx] = -v1[i]*v2[j]/2
template< int Dim >
void SkewProduct( const EFn< Dim > *v1 , const EFn< Dim > *v2 , EFn< Dim > c0[] , EFn< Dim > c1[] )
{
	for( int i=0 , idx=0 ; i<Dim ; i++ ) for( int j=0 ; j<i ; j++ , idx++ )
	{
		c0[idx] = ( v1[j]*v2[i] ).scaled( 1./2 );
		c1[idx] = ( v1[i].scaled( -1. )*v2[j] ).scaled( 1./2 );
	}
}

template< int Dim >
struct ElementTables
{
	static const int C = 1<<Dim;
	std::vector< double > wedgeMass;   // [i1][i2][j1][j2] (Hat.products.inl:299-313, Hat.inl:358-370)
	std::vector< double > stiffness;   // [i1][i2]         (Hat.scalars.inl:530-547)
	std::vector< double > fullStiff;   // [i1][i2][c][r]   (Hat.scalars.inl:550-565), m(d2,d1) = int d_{d1}phi_i1 d_{d2}phi_i2

	explicit ElementTables( unsigned int res )
	{
		const int CD = CDim< Dim >();
		EFn< Dim > grad[C][Dim];
		for( int c=0 ; c<C ; c++ ){ EFn< Dim > h = EFn< Dim >::Hat( c ) ; for( int d=0 ; d<Dim ; d++ ) grad[c][d] = h.D( d ); }
		stiffness.assign( C*C , 0 );
		fullStiff.assign( C*C*Dim*Dim , 0 );
		for( int i1=0 ; i1<C ; i1++ ) for( int i2=0 ; i2<C ; i2++ )
		{
			double s = 0;
			for( int d=0 ; d<Dim ; d++ ) s += ( grad[i1][d]*grad[i2][d] ).integral( res );
			stiffness[i1*C+i2] = s;
			for( int d1=0 ; d1<Dim ; d1++ ) for( int d2=0 ; d2<Dim ; d2++ )
				fullStiff[ ( (i1*C+i2)*Dim + d2 )*Dim + d1 ] = 0. + ( grad[i1][d1]*grad[i2][d2] ).integral( res );
		}
		std::vector< EFn< Dim > > P0( C*C*CD ) , P1( C*C*CD );
		for( int a=0 ; a<C ; a++ ) for( int b=0 ; b<C ; b++ ) SkewProduct< Dim >( grad[a] , grad[b] , &P0[(a*C+b)*CD] , &P1[(a*C+b)*CD] );
		wedgeMass.assign( (size_t)C*C*C*C , 0 );
		for( int i1=0 ; i1<C ; i1++ ) for( int i2=0 ; i2<C ; i2++ ) for( int j1=0 ; j1<C ; j1++ ) for( int j2=0 ; j2<C ; j2++ )
		{
			const EFn< Dim > *p1[2] = { &P0[(i1*C+i2)*CD] , &P1[(i1*C+i2)*CD] } , *p2[2] = { &P0[(j1*C+j2)*CD] , &P1[(j1*C+j2)*CD] };
			double integral = 0;
			for( int d=0 ; d<CD ; d++ ) for( int ii=0 ; ii<2 ; ii++ ) for( int jj=0 ; jj<2 ; jj++ )
				integral += ( p1[ii][d]*p2[jj][d] ).integral( res ) * 2;
			wedgeMass[ ( ( (size_t)i1*C+i2 )*C+j1 )*C+j2 ] = integral;
		}
	}
};

// The case index of a function for Radius-0 full stencils (Hat.scalars.inl:295-308)
template< int Dim > int StencilIndex( const Ix< Dim > &f , int res )
{
	int idx = 0;
	for( int d=0 ; d<Dim ; d++ )
	{
		if     ( f[d]<=0   ) idx = idx*3 + f[d];
		else if( f[d]>=res ) idx = idx*3 + 2 + f[d] - res;
		else                 idx = idx*3 + 1;
	}
	return idx;
}

// Scalar full integration stencil (Hat.scalars.inl:259-293): per case, entries (window, value)
template< int Dim >
struct ScalarFullStencil
{
	std::vector< std::vector< std::pair< int , double > > > rows;
	ScalarFullStencil( const std::vector< double > &stencil , int res )
	{
		const int C = 1<<Dim;
		rows.resize( Pow3< Dim >() );
		Ix< Dim > lo , hi;
		for( int d=0 ; d<Dim ; d++ ) hi[d] = std::min( 3 , res+1 );
		ForBox< Dim >( lo , hi , [&]( Ix< Dim > f1 )
		{
			for( int d=0 ; d<Dim ; d++ ){ if( f1[d]<=0 ); else if( f1[d]==1 ) f1[d] = res/2 ; else f1[d] = res-2+f1[d]; }
			std::map< Ix< Dim > , double > row;
			Ix< Dim > elo , ehi;
			for( int d=0 ; d<Dim ; d++ ) elo[d] = std::max( 0 , f1[d]-1 ) , ehi[d] = std::min( res , f1[d]+1 );
			ForBox< Dim >( elo , ehi , [&]( const Ix< Dim > &e )
			{
				Ix< Dim > flo , fhi;
				for( int d=0 ; d<Dim ; d++ ) flo[d] = std::max( 0 , e[d] ) , fhi[d] = std::min( res+1 , e[d]+2 );
				ForBox< Dim >( flo , fhi , [&]( const Ix< Dim > &f2 ){ row[ f2-f1 ] += stencil[ Rel( e , f1 )*C + Rel( e , f2 ) ]; } );
			} );
			std::vector< std::pair< int , double > > r;
			for( auto &kv : row ) r.emplace_back( W3< Dim >( kv.first ) , kv.second );
			rows[ StencilIndex< Dim >( f1 , res ) ] = r;
		} );
	}
};

// Wedge full integration stencil (Hat.products.inl:32-107)
template< int Dim >
struct WedgeFullStencil
{
	struct Entry { int g1 , g2 ; double value; };
	struct Row { int f2 ; long long df2 ; std::vector< Entry > entries; };
	std::vector< std::vector< Row > > rows;
	WedgeFullStencil( const std::vector< double > &mass , int res )
	{
		const int C = 1<<Dim;
		rows.resize( Pow3< Dim >() );
		Ix< Dim > lo , hi;
		for( int d=0 ; d<Dim ; d++ ) hi[d] = 3;
		ForBox< Dim >( lo , hi , [&]( Ix< Dim > f1 )
		{
			for( int d=0 ; d<Dim ; d++ ){ if( f1[d]==1 ) f1[d] = res/2 ; else if( f1[d]==2 ) f1[d] = res; }
			std::vector< Ix< Dim > > f2s;
			{
				Ix< Dim > flo , fhi;
				for( int d=0 ; d<Dim ; d++ ) flo[d] = std::max( 0 , f1[d]-1 ) , fhi[d] = std::min( res , f1[d]+1 )+1;
				ForBox< Dim >( flo , fhi , [&]( const Ix< Dim > &f2 ){ if( f2!=f1 ) f2s.push_back( f2 ); } );
			}
			std::vector< Row > _rows;
			for( const Ix< Dim > &f2 : f2s )
			{
				Row row;
				row.f2 = W3< Dim >( f2-f1 );
				row.df2 = (long long)FIndex< Dim >( f2 , res ) - (long long)FIndex< Dim >( f1 , res );
				std::map< std::pair< Ix< Dim > , Ix< Dim > > , double > m;
				Ix< Dim > elo , ehi;
				for( int d=0 ; d<Dim ; d++ ) elo[d] = std::max( 0 , std::max( f1[d] , f2[d] )-1 ) , ehi[d] = std::min( res , std::min( f1[d] , f2[d] )+1 );
				ForBox< Dim >( elo , ehi , [&]( const Ix< Dim > &e )
				{
					Ix< Dim > glo = e , ghi;
					for( int d=0 ; d<Dim ; d++ ) ghi[d] = e[d]+2;
					ForBox< Dim >( glo , ghi , [&]( const Ix< Dim > &g1 )
					{
						ForBox< Dim >( glo , ghi , [&]( const Ix< Dim > &g2 )
						{
							if( g1<g2 ) m[ std::make_pair( g1 , g2 ) ] += mass[ ( ( (size_t)Rel( e , f1 )*C + Rel( e , f2 ) )*C + Rel( e , g1 ) )*C + Rel( e , g2 ) ];
						} );
					} );
				} );
				for( auto &kv : m ) row.entries.push_back( Entry{ W3< Dim >( kv.first.first-f1 ) , W3< Dim >( kv.first.second-f1 ) , kv.second } );
				_rows.push_back( row );
			}
			rows[ StencilIndex< Dim >( f1 , res ) ] = _rows;
		} );
		// df2>0 before df2<0, with the reference's (unstable) std::sort (Hat.products.inl:99-106)
		auto Sort = []( const Row &r1 , const Row &r2 ){ return ( r1.df2==0 && r2.df2!=0 ) || ( r1.df2>0 && r2.df2<0 ); };
		for( auto &r : rows ) std::sort( r.begin() , r.end() , Sort );
	}
};

//////////////////////////////////////
// B-splines (GridSamples.inl:32-48)
//////////////////////////////////////
template< int Degree > void BSplineEval( double x , double *values )
{
	if constexpr( Degree==0 ) values[0] = 1.;
	else
	{
		const double Scale = 1./Degree;
		BSplineEval< Degree-1 >( x , values+1 );
		values[0] = values[1] * (1.-x) * Scale;
		for( int i=1 ; i<Degree ; i++ )
		{
			double x1 = (x-i+Degree) , x2 = (-x+i+1);
			values[i] = ( values[i]*x1 + values[i+1]*x2 ) * Scale;
		}
		values[Degree] *= x * Scale;
	}
}
void BSplineEvalRT( int degree , double x , double *values ) // GridSamples.inl:50-65
{
	if( degree==0 ) values[0] = 1.;
	else
	{
		const double Scale = 1./degree;
		BSplineEvalRT( degree-1 , x , values+1 );
		values[0] = values[1] * (1.-x) * Scale;
		for( int i=1 ; i<degree ; i++ )
		{
			double x1 = (x-i+degree) , x2 = (-x+i+1);
			values[i] = ( values[i]*x1 + values[i+1]*x2 ) * Scale;
		}
		values[degree] *= x * Scale;
	}
}

/////////////////////////////////////////////////////////////////////////////
// Symmetric eigenvalues (cyclic Jacobi), ascending.  Restates the role of
// Eigen::SelfAdjointEigenSolver in DensityAndNoiseInfo::noise (GridSamples.inl:94).
/////////////////////////////////////////////////////////////////////////////
template< int Dim > void SymEigenvalues( double A[Dim][Dim] , double ev[Dim] )
{
	for( int sweep=0 ; sweep<64 ; sweep++ )
	{
		double off = 0;
		for( int i=0 ; i<Dim ; i++ ) for( int j=i+1 ; j<Dim ; j++ ) off += A[i][j]*A[i][j];
		if( off==0 ) break;
		for( int p=0 ; p<Dim ; p++ ) for( int q=p+1 ; q<Dim ; q++ )
		{
			if( A[p][q]==0 ) continue;
			double theta = ( A[q][q]-A[p][p] ) / ( 2*A[p][q] );
			double t = ( theta>=0 ? 1. : -1. ) / ( std::fabs( theta ) + std::sqrt( theta*theta+1 ) );
			double c = 1./std::sqrt( t*t+1 ) , s = t*c;
			for( int k=0 ; k<Dim ; k++ )
			{
				double akp = A[k][p] , akq = A[k][q];
				A[k][p] = c*akp - s*akq , A[k][q] = s*akp + c*akq;
			}
			for( int k=0 ; k<Dim ; k++ )
			{
				double apk = A[p][k] , aqk = A[q][k];
				A[p][k] = c*apk - s*aqk , A[q][k] = s*apk + c*aqk;
			}
		}
	}
	for( int i=0 ; i<Dim ; i++ ) ev[i] = A[i][i];
	std::sort( ev , ev+Dim );
}

/////////////////////////////////////////////////////////////////////////////
// Density and noise statistics (GridSamples.h:78-94, GridSamples.inl:70-105)
/////////////////////////////////////////////////////////////////////////////
template< int Dim >
struct DNI
{
	double w , pos[Dim] , cov[Dim][Dim];  // cov[c][r] as MishaK::SquareMatrix coords
	DNI( void ) : w(0) { for( int i=0 ; i<Dim ; i++ ){ pos[i] = 0 ; for( int j=0 ; j<Dim ; j++ ) cov[i][j] = 0; } }
	DNI( const double *p , double weight ) : w(weight)
	{
		for( int i=0 ; i<Dim ; i++ ) pos[i] = p[i]*weight;
		for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) cov[i][j] = p[i]*p[j]*weight;
	}
	void add( const DNI &o ){ w += o.w ; for( int i=0 ; i<Dim ; i++ ) pos[i] += o.pos[i] ; for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) cov[i][j] += o.cov[i][j]; }
	DNI scaled( double s ) const { DNI o = *this ; o.w *= s ; for( int i=0 ; i<Dim ; i++ ) o.pos[i] *= s ; for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) o.cov[i][j] *= s ; return o; }
	double measure( void ) const { return 1./w; }
	double noise( int coDim ) const
	{
		double C[Dim][Dim] , ev[Dim];
		for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) C[i][j] = cov[i][j]/w - pos[i]/w * pos[j]/w;
		SymEigenvalues< Dim >( C , ev );
		double sum = 0 , lowSum = 0;
		for( int d=0 ; d<Dim ; d++ ){ sum += ev[d] ; if( d<coDim ) lowSum += ev[d]; }
		return ( lowSum * Dim ) / ( sum * coDim );
	}
};

template< int Dim , int CD >
struct Skew
{
	double p[CD];
	Skew( void ){ for( int i=0 ; i<CD ; i++ ) p[i] = 0; }
	void add( const Skew &o ){ for( int i=0 ; i<CD ; i++ ) p[i] += o.p[i]; }
	Skew scaled( double s ) const { Skew o = *this ; for( int i=0 ; i<CD ; i++ ) o.p[i] *= s ; return o; }
	// full matrix m(i,j)=p (i>j), m(j,i)=-p (Hat.inl:47-54)
	void full( double m[Dim][Dim] ) const
	{
		for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) m[i][j] = 0;
		for( int i=0 , idx=0 ; i<Dim ; i++ ) for( int j=0 ; j<i ; j++ , idx++ ) m[i][j] = p[idx] , m[j][i] = -p[idx];
	}
	double squareNorm( void ) const // Matrix::InnerProduct of the full matrix (Misha/Geometry.inl:88-95)
	{
		double m[Dim][Dim] ; full( m );
		double dot = 0;
		for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) dot += m[i][j]*m[i][j];
		return dot;
	}
};

// Per-depth dense corner grids of a tree splatter.  Tree depth t in [1,maxDepth+1]
// holds the corners [0,2^(t-1)]^Dim (GridSamples.inl:458-491).  Node existence
// (which decides where a point's leaf is, GridSamples.inl:600-651) is tracked
// as "has children" flags per depth (Misha/RegularTree.inl:762-800).
template< int Dim , typename Data >
struct DenseSplat
{
	int maxDepth;                                   // finest grid depth D; tree depths 1..D+1
	std::vector< std::vector< Data > > data;        // data[t] over (2^(t-1)+1)^Dim corners
	std::vector< std::vector< unsigned char > > kids; // kids[t] over offsets [0,2^t)^Dim (has children)
	explicit DenseSplat( int D ) : maxDepth(D) , data( D+2 ) , kids( D+2 )
	{
		for( int t=1 ; t<=D+1 ; t++ )
		{
			size_t n1 = ( (size_t)1<<(t-1) ) + 1 , n=1 ; for( int d=0 ; d<Dim ; d++ ) n *= n1;
			data[t].assign( n , Data() );
			if( t>D ) continue;
			size_t k = 1 ; for( int d=0 ; d<Dim ; d++ ) k *= (size_t)1<<t;
			kids[t].assign( k , 0 );
		}
	}
	size_t corner( int t , const Ix< Dim > &o ) const { int r = 1<<(t-1) ; size_t i=0 ; for( int d=0 ; d<Dim ; d++ ) i = i*(r+1) + o[d] ; return i; }
	size_t node( int t , const Ix< Dim > &o ) const { size_t i=0 ; for( int d=0 ; d<Dim ; d++ ) i = i*( (size_t)1<<t ) + o[d] ; return i; }

	// Node creation for one inserted leaf cell E (finest, depth D+1) with a
	// neighbour key of radii [K,K+1]: at every tree depth t in [2,D+1], the
	// parents (C_t+i)>>1, i in [-K,K+1], receive children.
	void touch( const Ix< Dim > &E , int K )
	{
		for( int t=2 ; t<=maxDepth+1 ; t++ )
		{
			Ix< Dim > C , plo , phi;
			for( int d=0 ; d<Dim ; d++ )
			{
				C[d] = E[d]>>( maxDepth+1-t );
				plo[d] = std::max( 0 , ( C[d]-K )>>1 );
				phi[d] = std::min( (1<<(t-1))-1 , ( C[d]+K+1 )>>1 ) + 1;
			}
			ForBox< Dim >( plo , phi , [&]( const Ix< Dim > &P ){ kids[t-1][ node( t-1 , P ) ] = 1; } );
		}
	}
	// depth of the leaf containing p (getLeafNode from the space root at depth 1)
	int leafDepth( const double *p ) const
	{
		int t = 1;
		while( t<=maxDepth )
		{
			Ix< Dim > o;
			for( int d=0 ; d<Dim ; d++ ) o[d] = (int)floor( p[d] * ( 1<<(t-1) ) );
			if( !kids[t][ node( t , o ) ] ) break;
			t++;
		}
		return t;
	}
};

/////////////////////////////////////////////////////////////////////////////
// The reconstruction state
/////////////////////////////////////////////////////////////////////////////
struct Base
{
	virtual ~Base(){}
	virtual int dim( void ) const = 0;
	virtual size_t numFunctions( int l ) const = 0;
	virtual size_t numElements( int l ) const = 0;
	virtual size_t get( const std::string &name , double *out , size_t cap ) const = 0;
	virtual void matvec( int l , char which , const double *x , double *y ) const = 0;
	virtual void energies( int l , const double *x , const double *y , double out[2] ) const = 0;
	virtual double wedge( int l , const double *x , const double *y ) const = 0;
	virtual void qfit( int l , const double *x , const double *y , double *q ) const = 0;
	virtual void gs( int l , double *x , double *y , int iters ) const = 0;
	virtual void init( double *x , double *y ) const = 0;
	virtual void solve( double *hx , double *hy , double *fx , double *fy , double E[2] ) const = 0;
	virtual int writeTree( const char *path , const double *hx , const double *hy , unsigned long long hash ) const = 0;
};

template< int Dim >
struct Recon : public Base
{
	static const int CD = CDim< Dim >();
	static const int NW = Pow3< Dim >();
	static const int NC = 1<<Dim;
	static const int CoDim = 2;
	using SkewT = Skew< Dim , CD >;

	orc_params prm;
	int D;
	size_t n;
	std::vector< double > pos;          // n*Dim unit-cube positions
	std::vector< SkewT > skew;          // packed frames
	double unitCubeToWorld[Dim+1][Dim+1];
	std::vector< std::pair< Ix< Dim > , std::vector< size_t > > > cells;
	std::vector< double > measureV , noiseV , depthV , S;

	struct Level
	{
		int res;
		size_t N , Ne;
		std::vector< double > R , B;     // N*NW, column f: M(f+off(w), f)
		std::unique_ptr< ScalarFullStencil< Dim > > K;
		std::unique_ptr< WedgeFullStencil< Dim > > W;
		std::vector< std::vector< unsigned int > > colours;
		// adapted hierarchy (Hat.adaptive.inl:349-729), empty for the dense grid:
		// function flags (0 not indexed, 1 active, 2 padding), element flags, and
		// the rank of every indexed function in the hierarchy's numbering
		std::vector< unsigned char > fmask , emask;
		std::vector< long long > rank;
		std::vector< size_t > order;
		bool indexed( size_t f ) const { return fmask.empty() || fmask[f]!=0; }
		bool active( size_t f ) const { return fmask.empty() || fmask[f]==1; }
		bool elementIndexed( size_t e ) const { return emask.empty() || emask[e]!=0; }
	};
	// set before build() for an adapted reconstruction
	struct AdaptedInput { const unsigned char *fmask , *emask ; const long long *forder ; const size_t *nforder; };
	const AdaptedInput *adaptedIn = nullptr;
	bool adapted( void ) const { return !levels.empty() && !levels[0].fmask.empty(); }
	std::vector< Level > levels;
	double c_;

	int dim( void ) const { return Dim; }
	size_t numFunctions( int l ) const { return levels[l].N; }
	size_t numElements( int l ) const { return levels[l].Ne; }

	static size_t NumF( int res ){ size_t s=1 ; for( int d=0 ; d<Dim ; d++ ) s *= res+1 ; return s; }
	static size_t NumE( int res ){ size_t s=1 ; for( int d=0 ; d<Dim ; d++ ) s *= res ; return s; }

	// ---- ToUnitCube (ExteriorPoisson.h:100-123) with Matrix products (Misha/Geometry.inl:97-107, 410-427)
	static void MatMul( const double A[Dim+1][Dim+1] , const double B[Dim+1][Dim+1] , double out[Dim+1][Dim+1] )
	{
		for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ){ out[i][j] = 0 ; for( int k=0 ; k<=Dim ; k++ ) out[i][j] += B[i][k]*A[k][j]; }
	}

	Recon( const orc_params &p , size_t _n , const double *wpos , const double *wskew , const AdaptedInput *ad=nullptr ) : prm(p) , D(p.depth) , n(_n) , adaptedIn(ad)
	{
		if( !n ) throw std::string( "Expected some samples: 0" );
		if( prm.tRadius<0 || prm.tRadius>4 ) throw std::string( "Kernel radius must be in range [0,4]" );
		if( prm.dRadius<0 || prm.dRadius>4 ) throw std::string( "Only radii [0,4] supported" );
		// frames: SkewSymmetricMatrix( m ) with m(i,j)=coords[i][j]=skew_i_j (Hat.inl:39-45)
		skew.resize( n );
		for( size_t s=0 ; s<n ; s++ ) for( int i=0 , idx=0 ; i<Dim ; i++ ) for( int j=0 ; j<i ; j++ , idx++ )
			skew[s].p[idx] = ( wskew[ s*Dim*Dim + i*Dim + j ] - wskew[ s*Dim*Dim + j*Dim + i ] ) / 2.;

		// ToUnitCube
		double bb0[Dim] , bb1[Dim];
		for( int d=0 ; d<Dim ; d++ ) bb0[d] = bb1[d] = wpos[d];
		for( size_t i=1 ; i<n ; i++ ) for( int d=0 ; d<Dim ; d++ ){ bb0[d] = std::min( bb0[d] , wpos[i*Dim+d] ) ; bb1[d] = std::max( bb1[d] , wpos[i*Dim+d] ); }
		double _scale = bb1[0]-bb0[0];
		for( int d=1 ; d<Dim ; d++ ) _scale = std::max( _scale , bb1[d]-bb0[d] );
		double t1[Dim+1][Dim+1] , s[Dim+1][Dim+1] , t2[Dim+1][Dim+1] , ts[Dim+1][Dim+1] , M[Dim+1][Dim+1];
		for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ) t1[i][j] = s[i][j] = t2[i][j] = ( i==j ? 1. : 0. );
		for( int d=0 ; d<Dim ; d++ ) t2[Dim][d] = -( bb0[d] + bb1[d] ) / 2.;
		for( int d=0 ; d<Dim ; d++ ) s[d][d] = 1. / ( _scale * prm.scale );
		for( int d=0 ; d<Dim ; d++ ) t1[Dim][d] = 0.5;
		MatMul( t1 , s , ts );
		MatMul( ts , t2 , M );
		pos.resize( n*Dim );
		for( size_t i=0 ; i<n ; i++ )
		{
			double scl = 1;
			for( int a=0 ; a<Dim ; a++ )
			{
				double o = 0;
				for( int b=0 ; b<Dim ; b++ ) o += wpos[i*Dim+b]*M[b][a];
				o += M[Dim][a];
				scl += M[a][Dim];
				pos[i*Dim+a] = o;
			}
			for( int a=0 ; a<Dim ; a++ ) pos[i*Dim+a] /= scl;
		}
		// unitCubeToWorld = inverse( worldToUnitCube ) for a scale+translate transform
		for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ) unitCubeToWorld[i][j] = ( i==j ? 1. : 0. );
		for( int d=0 ; d<Dim ; d++ ){ unitCubeToWorld[d][d] = 1./M[d][d] ; unitCubeToWorld[Dim][d] = -M[Dim][d]/M[d][d]; }
		// --jitter (ExteriorPoissonRecon.cpp:174-180): translation( Dim , d ) = ( Random()*2-1 )*jitter
		// drawn from the unseeded rand() sequence (the seed is set only later, :220)
		if( prm.jitter>0 )
		{
			double T[Dim+1][Dim+1] , Ti[Dim+1][Dim+1] , U[Dim+1][Dim+1];
			for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ) T[i][j] = ( i==j ? 1. : 0. );
			srand( 1 );
			for( int d=0 ; d<Dim ; d++ ) T[Dim][d] = ( Random() * 2. - 1. ) * prm.jitter;
			for( size_t i=0 ; i<n ; i++ )
			{
				double q[Dim] , scl = 1;
				for( int a=0 ; a<Dim ; a++ )
				{
					double o = 0;
					for( int b=0 ; b<Dim ; b++ ) o += pos[i*Dim+b]*T[b][a];
					o += T[Dim][a];
					scl += T[a][Dim];
					q[a] = o;
				}
				for( int a=0 ; a<Dim ; a++ ) pos[i*Dim+a] = q[a] / scl;
			}
			// translation.inverse(): the identity with the negated translation row
			for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ) Ti[i][j] = ( i==Dim && j<Dim ) ? -T[i][j] : T[i][j];
			MatMul( unitCubeToWorld , Ti , U );
			memcpy( unitCubeToWorld , U , sizeof(U) );
		}

		orderSamples();
		build();
	}

	// ---- OrderedSampler (OrderedSamples.inl:32-85)
	void orderSamples( void )
	{
		int res = 1<<D;
		struct IS { size_t idx ; Ix< Dim > I ; bool operator<( const IS &s ) const { return I<s.I; } };
		std::vector< IS > v( n );
		for( size_t i=0 ; i<n ; i++ )
		{
			v[i].idx = i;
			for( int d=0 ; d<Dim ; d++ ) v[i].I[d] = std::max< int >( 0 , std::min< int >( (int)floor( pos[i*Dim+d] * res ) , res-1 ) );
		}
		std::sort( v.begin() , v.end() );
		size_t start = 0;
		while( start<v.size() )
		{
			size_t end = start;
			while( end<v.size() && v[end].I==v[start].I ) end++;
			std::vector< size_t > idx;
			for( size_t i=start ; i<end ; i++ ) idx.push_back( v[i].idx );
			cells.emplace_back( v[start].I , idx );
			start = end;
		}
	}

	// ---- splat at one tree depth (GridSamples.inl:458-491)
	template< int K , typename Data >
	static void SplatAt( DenseSplat< Dim , Data > &tree , int t , const double *p , const Data &data )
	{
		const int Deg = 2*K+1;
		int res = 1<<(t-1);
		double bv[Dim][Deg+1];
		Ix< Dim > E;
		for( int d=0 ; d<Dim ; d++ ){ E[d] = (int)floor( p[d]*res ) ; BSplineEval< Deg >( p[d]*res - E[d] , bv[d] ); }
		Ix< Dim > lo , hi;
		for( int d=0 ; d<Dim ; d++ ) hi[d] = Deg+1;
		ForBox< Dim >( lo , hi , [&]( const Ix< Dim > &i )
		{
			double scl = 1.;
			Ix< Dim > o;
			for( int d=0 ; d<Dim ; d++ ){ scl = scl * bv[d][ i[d] ] ; o[d] = E[d] - K + i[d]; if( o[d]<0 || o[d]>res ) return; }
			tree.data[t][ tree.corner( t , o ) ].add( data.scaled( scl ) );
		} );
	}

	// ---- multilinear evaluation at one tree depth (GridSamples.inl:774-809)
	template< typename Data >
	static Data EvalAt( const DenseSplat< Dim , Data > &tree , int t , const double *p )
	{
		int res = 1<<(t-1);
		double v[Dim][2];
		Ix< Dim > e;
		for( int d=0 ; d<Dim ; d++ ){ e[d] = (int)floor( p[d]*res ) ; BSplineEval< 1 >( p[d]*res - e[d] , v[d] ); }
		Data out;
		Ix< Dim > lo , hi;
		for( int d=0 ; d<Dim ; d++ ) hi[d] = 2;
		ForBox< Dim >( lo , hi , [&]( const Ix< Dim > &f )
		{
			Ix< Dim > o = e + f;
			for( int d=0 ; d<Dim ; d++ ) if( o[d]>res ) return;
			double scale = 1.;
			for( int d=0 ; d<Dim ; d++ ) scale *= v[d][ f[d] ];
			out.add( tree.data[t][ tree.corner( t , o ) ].scaled( scale ) );
		} );
		return out;
	}

	// ---- TreeEstimator queries (GridSamples.inl:811-896)
	std::unique_ptr< DenseSplat< Dim , DNI< Dim > > > est;
	double targetSPC;
	double samplesPerCell( int t , const double *p ) const { return EvalAt( *est , t , p ).w / (double)( 1<<( (t-1)*(Dim-CoDim) ) ); }
	double estDepth( int maxT , const double *p ) const
	{
		int d1 = maxT , d2 = maxT+1;
		while( d2>1 && samplesPerCell( d1 , p )<targetSPC ) d1-- , d2--;
		if( d2==1 || d2==maxT+1 )
		{
			int d = ( d2==1 ) ? 1 : maxT;
			double spc = samplesPerCell( d , p );
			return d + log( spc/targetSPC ) / log(2.);
		}
		double s1 = samplesPerCell( d1 , p ) , s2 = samplesPerCell( d2 , p );
		if( !s2 ) return d1;
		return d1 + log( targetSPC/s1 ) / log( s2/s1 );
	}
	DNI< Dim > estInfo( const double *p ) const
	{
		int maxT = D+1;
		double depth = estDepth( maxT , p );
		if( depth>=maxT ) return EvalAt( *est , maxT , p );
		else if( depth<0 ) return EvalAt( *est , 1 , p );
		int d1 = (int)floor( depth ) , d2 = d1+1;
		if( d1<1 ) throw std::string( "estimator depth below the coarsest level" );
		double d = depth-d1;
		DNI< Dim > a = EvalAt( *est , d1 , p ).scaled( 1.-d ) , b = EvalAt( *est , d2 , p ).scaled( d );
		a.add( b );
		return a;
	}

	template< int K >
	void buildEstimator( void )
	{
		est.reset( new DenseSplat< Dim , DNI< Dim > >( D ) );
		targetSPC = pow( 2. , Dim-CoDim );
		double scale;
		{
			std::vector< double > v( 2*K+2 );
			BSplineEvalRT( 2*K+1 , 0.5 , &v[0] );
			scale = pow( 1./v[K+1] , CoDim );
		}
		std::vector< double > scales( D+1 );
		for( int d=0 ; d<=D ; d++ ) scales[d] = scale / pow( pow( 0.5 , d ) , Dim-CoDim );
		// TreeSplatter::addSamples with merged cells (GridSamples.inl:361-391, 436-441)
		for( auto &cell : cells )
		{
			est->touch( cell.first , K );
			double p[Dim] = {};
			DNI< Dim > data;
			for( size_t j : cell.second ){ for( int d=0 ; d<Dim ; d++ ) p[d] += pos[j*Dim+d] ; data.add( DNI< Dim >( &pos[j*Dim] , 1. ) ); }
			double inv = 1. / static_cast< double >( cell.second.size() );
			for( int d=0 ; d<Dim ; d++ ) p[d] *= inv;
			for( int t=1 ; t<=D+1 ; t++ ) SplatAt< K >( *est , t , p , data.scaled( scales[t-1] ) );
		}
	}

	std::unique_ptr< DenseSplat< Dim , SkewT > > splat;

	template< int K >
	void buildSkewSplat( void )
	{
		splat.reset( new DenseSplat< Dim , SkewT >( D ) );
		std::vector< std::pair< SkewT , double > > samples( n );
		// ExteriorPoisson.h:181-213
		for( auto &cell : cells ) for( size_t j : cell.second )
		{
			SkewT s = skew[j];
			s = s.scaled( 1./sqrt( s.squareNorm() ) );
			s = s.scaled( measureV[j] );
			s = s.scaled( (double)( 1<<(D*CoDim) ) );
			samples[j] = std::make_pair( s , depthV[j] );
		}
		auto DW = []( int d ){ return (double)( 1<<(d*Dim) ); };
		// TreeSplatter::addSamples, un-merged, fractional depth (GridSamples.inl:393-456)
		for( auto &cell : cells )
		{
			splat->touch( cell.first , K );
			for( size_t j : cell.second )
			{
				const double *p = &pos[j*Dim];
				const SkewT &data = samples[j].first;
				double depth = samples[j].second + 1.;
				if( depth>=D+1 ) SplatAt< K >( *splat , D+1 , p , data.scaled( DW( D ) ) );
				else if( depth<=1 ) SplatAt< K >( *splat , 1 , p , data.scaled( DW( 0 ) ) );
				else
				{
					int d1 = (int)floor( depth ) , d2 = d1+1;
					double dd = depth-d1;
					SplatAt< K >( *splat , d1 , p , data.scaled( 1.-dd ).scaled( DW( d1-1 ) ) );
					SplatAt< K >( *splat , d2 , p , data.scaled( dd ).scaled( DW( d2-1 ) ) );
				}
			}
		}
		// ExteriorPoisson.h:260-264
		double scale = pow( pow( 0.5 , D ) , 1.5 * CoDim );
		for( int t=1 ; t<=D+1 ; t++ ) for( auto &s : splat->data[t] ) s = s.scaled( scale );
	}

	// S(E) at the centre of element e (ExteriorPoisson.h:294-307, GridSamples.inl:600-651)
	SkewT fieldAt( const double *p ) const
	{
		int L = splat->leafDepth( p );
		SkewT s;
		for( int t=1 ; t<=L ; t++ )
		{
			int res = 1<<(t-1);
			double v[Dim][2];
			Ix< Dim > e;
			for( int d=0 ; d<Dim ; d++ ){ e[d] = (int)floor( p[d]*res ) ; BSplineEval< 1 >( p[d]*res - e[d] , v[d] ); }
			Ix< Dim > lo , hi;
			for( int d=0 ; d<Dim ; d++ ) hi[d] = 2;
			ForBox< Dim >( lo , hi , [&]( const Ix< Dim > &f )
			{
				double scl = 1.;
				Ix< Dim > o = e + f;
				for( int d=0 ; d<Dim ; d++ ){ scl = scl * v[d][ f[d] ] ; if( o[d]>res ) return; }
				s.add( splat->data[t][ splat->corner( t , o ) ].scaled( scl ) );
			} );
		}
		return s;
	}

	void build( void )
	{
		switch( prm.dRadius ){ case 0: buildEstimator<0>() ; break ; case 1: buildEstimator<1>() ; break ; case 2: buildEstimator<2>() ; break ; case 3: buildEstimator<3>() ; break ; default: buildEstimator<4>(); }
		measureV.resize( n ) , noiseV.resize( n ) , depthV.resize( n );
		for( size_t i=0 ; i<n ; i++ )
		{
			const double *p = &pos[i*Dim];
			DNI< Dim > info = estInfo( p );
			measureV[i] = info.measure();
			noiseV[i] = info.noise( CoDim );
			depthV[i] = estDepth( D+1 , p ) - 1.;
		}
		switch( prm.tRadius ){ case 0: buildSkewSplat<0>() ; break ; case 1: buildSkewSplat<1>() ; break ; case 2: buildSkewSplat<2>() ; break ; case 3: buildSkewSplat<3>() ; break ; default: buildSkewSplat<4>(); }

		levels.resize( D+1 );
		for( int l=0 ; l<=D ; l++ )
		{
			Level &L = levels[l];
			L.res = 1<<l , L.N = NumF( L.res ) , L.Ne = NumE( L.res );
			ElementTables< Dim > tab( L.res );
			L.K.reset( new ScalarFullStencil< Dim >( tab.stiffness , L.res ) );
			L.W.reset( new WedgeFullStencil< Dim >( tab.wedgeMass , L.res ) );
			if( adaptedIn )
			{
				size_t fo = 0 , eo = 0 , oo = 0;
				for( int k=0 ; k<l ; k++ ){ fo += NumF( 1<<k ) ; eo += NumE( 1<<k ) ; oo += adaptedIn->nforder[k]; }
				L.fmask.assign( adaptedIn->fmask+fo , adaptedIn->fmask+fo+L.N );
				L.emask.assign( adaptedIn->emask+eo , adaptedIn->emask+eo+L.Ne );
				L.rank.assign( L.N , -1 );
				L.order.resize( adaptedIn->nforder[l] );
				for( size_t i=0 ; i<adaptedIn->nforder[l] ; i++ ){ L.order[i] = (size_t)adaptedIn->forder[oo+i] ; L.rank[ L.order[i] ] = (long long)i; }
			}
			// colour lists over the active functions (ExteriorPoisson.h:152-171)
			L.colours.resize( NC );
			for( size_t f=0 ; f<L.N ; f++ ) if( L.active( f ) )
			{
				Ix< Dim > F = FIndex< Dim >( f , L.res );
				unsigned int mc = 0;
				for( int d=0 ; d<Dim ; d++ ) if( F[d]&1 ) mc |= 1u<<d;
				L.colours[mc].push_back( (unsigned int)f );
			}
		}
		Level &F = levels[D];
		int res = F.res;
		// per-element field
		S.assign( F.Ne*CD , 0 );
		std::vector< SkewT > Se( F.Ne );
		for( size_t e=0 ; e<F.Ne ; e++ )
		{
			Ix< Dim > I = EIndex< Dim >( e , res );
			double p[Dim];
			for( int d=0 ; d<Dim ; d++ ) p[d] = ( I[d] + 0.5 ) / res;
			Se[e] = fieldAt( p );
			for( int c=0 ; c<CD ; c++ ) S[e*CD+c] = Se[e].p[c];
		}
		// c = sum ||S||^2 over the indexed elements / #elements (SystemEnergy.h:455-461)
		c_ = 0;
		for( size_t e=0 ; e<F.Ne ; e++ ) if( F.elementIndexed( e ) ) c_ += Se[e].squareNorm();
		c_ /= F.Ne;
		buildR( F );
		buildB( F , Se );
		// Galerkin chain (SystemEnergy.h:464-477, ExteriorPoisson.h:311)
		for( int l=D-1 ; l>=0 ; l-- ) galerkin( levels[l] , levels[l+1] );
	}

	// ---- R = deltaMass * sWeight * mean(measure) (ExteriorPoisson.h:267-287, Hat.scalars.inl:976-1063)
	void buildR( Level &F )
	{
		F.R.assign( F.N*NW , 0 );
		if( !( prm.sWeight>0 ) ) return;
		int res = F.res;
		double measure = 0;
		size_t sampleNum = 0;
		for( auto &cell : cells )
		{
			sampleNum += cell.second.size();
			double m = 0;
			for( size_t j : cell.second ) m += measureV[j];
			measure += m;
		}
		measure /= sampleNum;
		// triplets in serial order; duplicates collapse in insertion order (Eigen setFromTriplets)
		std::vector< double > E( F.N*NW , 0 );   // column idx2, row idx1 (upper part F1<=F2)
		std::vector< unsigned char > seen( F.N*NW , 0 );
		for( auto &cell : cells )
		{
			const Ix< Dim > &Ec = cell.first;
			std::vector< double > w( cell.second.size() );
			std::vector< std::array< double , NC > > fv( cell.second.size() );
			for( size_t j=0 ; j<cell.second.size() ; j++ )
			{
				double nz = noiseV[ cell.second[j] ];
				w[j] = exp( -nz*nz/(2.*prm.nSigma*prm.nSigma) );
				double lp[Dim] , vals[2][Dim];
				for( int d=0 ; d<Dim ; d++ ){ lp[d] = pos[ cell.second[j]*Dim+d ] * res ; lp[d] -= Ec[d]; }
				for( int i=0 ; i<2 ; i++ ) for( int d=0 ; d<Dim ; d++ ) vals[i][d] = i==0 ? (1.-lp[d]) : lp[d];
				for( int c=0 ; c<NC ; c++ )
				{
					double value = 1.;
					for( int d=0 ; d<Dim ; d++ ) value *= ( ( c>>(Dim-1-d) )&1 ) ? vals[1][d] : vals[0][d];
					fv[j][c] = value;
				}
			}
			for( int i1=0 ; i1<NC ; i1++ )
			{
				Ix< Dim > F1 = Ec; for( int d=0 ; d<Dim ; d++ ) F1[d] += ( i1>>(Dim-1-d) )&1;
				for( int i2=0 ; i2<NC ; i2++ )
				{
					Ix< Dim > F2 = Ec; for( int d=0 ; d<Dim ; d++ ) F2[d] += ( i2>>(Dim-1-d) )&1;
					if( !( F1<=F2 ) ) continue;
					size_t idx2 = FIndex< Dim >( F2 , res );
					double value = 0;
					for( size_t j=0 ; j<cell.second.size() ; j++ ) value += fv[j][i1] * fv[j][i2] * w[j];
					if( F1==F2 ) value /= 2.;
					// entry (row idx1, col idx2)
					size_t k = idx2*NW + W3< Dim >( F1-F2 );
					if( seen[k] ) E[k] = E[k] + value; else E[k] = value , seen[k] = 1;
				}
			}
		}
		// R = (E + E^T) * sWeight * measure
		for( size_t f=0 ; f<F.N ; f++ )
		{
			Ix< Dim > Fi = FIndex< Dim >( f , res );
			for( int w=0 ; w<NW ; w++ )
			{
				Ix< Dim > o = W3< Dim >( w ) , G = Fi + o;
				bool in = true;
				for( int d=0 ; d<Dim ; d++ ) if( G[d]<0 || G[d]>res ) in = false;
				if( !in ) continue;
				size_t g = FIndex< Dim >( G , res );
				Ix< Dim > mo; for( int d=0 ; d<Dim ; d++ ) mo[d] = -o[d];
				// R(g,f) = E(g,f) + E(f,g)
				double a = E[ f*NW+w ] , b = E[ g*NW+W3< Dim >( mo ) ];
				bool sa = seen[ f*NW+w ] , sb = seen[ g*NW+W3< Dim >( mo ) ];
				double v;
				if( sa && sb ) v = a+b; else if( sa ) v = a; else if( sb ) v = b; else continue;
				F.R[ f*NW+w ] = v * prm.sWeight * measure;
			}
		}
	}

	// ---- B = stiffness(indexer, S) (SystemEnergy.h:453, Hat.scalars.inl:1584-1690)
	void buildB( Level &F , const std::vector< SkewT > &Se )
	{
		int res = F.res;
		F.B.assign( F.N*NW , 0 );
		ElementTables< Dim > tab( res );
		for( size_t f1=0 ; f1<F.N ; f1++ )
		{
			if( !F.indexed( f1 ) ) continue;
			Ix< Dim > F1 = FIndex< Dim >( f1 , res );
			for( int w=0 ; w<NW ; w++ )
			{
				Ix< Dim > F2 = F1 + W3< Dim >( w );
				bool in = true;
				for( int d=0 ; d<Dim ; d++ ) if( F2[d]<0 || F2[d]>res ) in = false;
				if( !in || !F.indexed( FIndex< Dim >( F2 , res ) ) ) continue;
				double dot = 0;
				Ix< Dim > elo , ehi;
				for( int d=0 ; d<Dim ; d++ ) elo[d] = std::max( 0 , std::max( F1[d] , F2[d] )-1 ) , ehi[d] = std::min( res , std::min( F1[d] , F2[d] )+1 );
				ForBox< Dim >( elo , ehi , [&]( const Ix< Dim > &E )
				{
					// adapted: elements outside the hierarchy contribute a zero frame (Hat.scalars.inl:1640-1650)
					if( !F.elementIndexed( EIndex< Dim >( E , res ) ) ) return;
					double m[Dim][Dim];
					Se[ EIndex< Dim >( E , res ) ].full( m );
					int i1 = Rel( E , F1 ) , i2 = Rel( E , F2 );
					const double *st = &tab.fullStiff[ (size_t)( i1*NC+i2 )*Dim*Dim ];
					double d_ = 0;
					for( int a=0 ; a<Dim ; a++ ) for( int b=0 ; b<Dim ; b++ ) d_ += st[a*Dim+b] * m[a][b];
					dot += d_;
				} );
				F.B[ f1*NW+w ] = dot;   // B(f2,f1)
			}
		}
	}

	// ---- prolongation x_f = sum_c P(f,c) x_c (Hat.scalars.inl:373-404, 1694-1744): y[f] accumulated over coarse c in increasing order
	//      adapted: P has rows for the indexed fine functions and columns for the
	//      indexed coarse ones; the sparse product accumulates over the columns in
	//      the hierarchy's numbering
	void prolong( int l , const double *xc , double *xf ) const
	{
		const Level &C = levels[l] , &F = levels[l+1];
		for( size_t f=0 ; f<F.N ; f++ ) xf[f] = 0;
		for( size_t ci=0 ; ci<( adapted() ? C.order.size() : C.N ) ; ci++ )
		{
			const size_t c = adapted() ? C.order[ci] : ci;
			Ix< Dim > Cc = FIndex< Dim >( c , C.res );
			for( int w=0 ; w<NW ; w++ )
			{
				Ix< Dim > o = W3< Dim >( w ) , G;
				double val = 1;
				bool in = true;
				for( int d=0 ; d<Dim ; d++ ){ G[d] = 2*Cc[d]+o[d] ; if( G[d]<0 || G[d]>F.res ) in = false ; if( o[d] ) val *= 0.5; }
				if( !in || !F.indexed( FIndex< Dim >( G , F.res ) ) ) continue;
				xf[ FIndex< Dim >( G , F.res ) ] += val * xc[c];
			}
		}
	}

	// ---- Galerkin M_c = (P^T M) P with Eigen's conservative product order (SystemEnergy.h:472-473)
	void galerkin( Level &C , const Level &F )
	{
		C.R.assign( C.N*NW , 0 );
		C.B.assign( C.N*NW , 0 );
		auto P = [&]( const Ix< Dim > &fine , const Ix< Dim > &coarse , double &v )
		{
			v = 1;
			for( int d=0 ; d<Dim ; d++ ){ int o = fine[d]-2*coarse[d] ; if( o<-1 || o>1 ) return false ; if( o ) v *= 0.5; }
			return true;
		};
		auto Do = [&]( const std::vector< double > &Mf , std::vector< double > &Mc )
		{
			// T = P^T Mf: T(i,k) = sum over stored rows m of Mf(:,k) (increasing) of P(m,i)*Mf(m,k)
			// Mc(i,j) = sum over k in P(:,j) (increasing) of T(i,k)*P(k,j)
			// inner indices in increasing order: dense index, or the hierarchy's numbering
			auto Ordered = [&]( const Level &L , const Ix< Dim > &base , int scale , int *ws )
			{
				int n = 0;
				long long key[NW];
				for( int w=0 ; w<NW ; w++ )
				{
					Ix< Dim > G; bool in = true;
					for( int d=0 ; d<Dim ; d++ ){ G[d] = scale*base[d] + W3< Dim >( w )[d] ; if( G[d]<0 || G[d]>L.res ) in = false; }
					if( !in ) continue;
					size_t g = FIndex< Dim >( G , L.res );
					if( !L.indexed( g ) ) continue;
					key[n] = adapted() ? L.rank[g] : (long long)g;
					ws[n++] = w;
				}
				for( int a=1 ; a<n ; a++ ) for( int b=a ; b>0 && key[b-1]>key[b] ; b-- ){ std::swap( key[b-1] , key[b] ) ; std::swap( ws[b-1] , ws[b] ); }
				return n;
			};
			for( size_t j=0 ; j<C.N ; j++ )
			{
				if( !C.indexed( j ) ) continue;
				Ix< Dim > J = FIndex< Dim >( j , C.res );
				int wks[NW];
				const int nk = Ordered( F , J , 2 , wks );
				int wmss[NW][NW] , nms[NW];
				for( int ik=0 ; ik<nk ; ik++ )
				{
					Ix< Dim > Kf; for( int d=0 ; d<Dim ; d++ ) Kf[d] = 2*J[d] + W3< Dim >( wks[ik] )[d];
					nms[ik] = Ordered( F , Kf , 1 , wmss[ik] );
				}
				for( int wi=0 ; wi<NW ; wi++ )
				{
					Ix< Dim > I = J + W3< Dim >( wi );
					bool in = true;
					for( int d=0 ; d<Dim ; d++ ) if( I[d]<0 || I[d]>C.res ) in = false;
					if( !in || !C.indexed( FIndex< Dim >( I , C.res ) ) ) continue;
					double acc = 0; bool any = false;
					for( int ik=0 ; ik<nk ; ik++ )   // k over fine children of j, increasing
					{
						const int wk = wks[ik];
						Ix< Dim > Kf; for( int d=0 ; d<Dim ; d++ ) Kf[d] = 2*J[d] + W3< Dim >( wk )[d];
						double pkj ; P( Kf , J , pkj );
						size_t k = FIndex< Dim >( Kf , F.res );
						double t = 0; bool tany = false;
						const int nm = nms[ik] , *wms = wmss[ik];
						for( int im=0 ; im<nm ; im++ )
						{
							const int wm = wms[im];
							Ix< Dim > Mi = Kf + W3< Dim >( wm );
							double pmi;
							if( !P( Mi , I , pmi ) ) continue;
							double mv = Mf[ k*NW+wm ];
							if( mv==0 ) continue;
							if( tany ) t += pmi*mv; else t = pmi*mv , tany = true;
						}
						if( !tany ) continue;
						if( any ) acc += t*pkj; else acc = t*pkj , any = true;
					}
					Mc[ j*NW+wi ] = acc;
				}
			}
		};
		Do( F.R , C.R );
		Do( F.B , C.B );
	}

	// ---- quadratic fit of one node (SystemEnergy.h:542-643)
	void fitNode( const Level &L , size_t idx , const double *x , const double *y , double Q[5] ) const
	{
		double q20=0 , q02=0 , q11=0 , q10=0 , q01=0;
		Ix< Dim > F1 = FIndex< Dim >( idx , L.res );
		long long nb[NW];
		for( int w=0 ; w<NW ; w++ )
		{
			Ix< Dim > G = F1 + W3< Dim >( w );
			bool in = true;
			for( int d=0 ; d<Dim ; d++ ) if( G[d]<0 || G[d]>L.res ) in = false;
			nb[w] = in ? (long long)FIndex< Dim >( G , L.res ) : -1;
		}
		for( int w=0 ; w<NW ; w++ ) if( nb[w]>=0 )
		{
			double v = L.R[ idx*NW+w ];
			if( v==0 && !( w==NW/2 ) ) continue;
			size_t r = (size_t)nb[w];
			if( r==idx ){ q20 += v ; q02 += v; }
			q10 += 2. * v * x[r];
			q01 += 2. * v * y[r];
		}
		if( prm.dWeight )
		{
			const auto &row = L.K->rows[ StencilIndex< Dim >( F1 , L.res ) ];
			for( auto &e : row )
			{
				long long _idx = nb[ e.first ];
				double value = e.second * prm.dWeight;
				if( _idx!=-1 )
				{
					if( (size_t)_idx==idx ){ q20 += value ; q02 += value; }
					q10 += 2. * value * x[_idx];
					q01 += 2. * value * y[_idx];
				}
			}
		}
		for( int w=0 ; w<NW ; w++ ) if( nb[w]>=0 && w!=NW/2 )
		{
			double v = L.B[ idx*NW+w ];
			if( v==0 ) continue;
			size_t r = (size_t)nb[w];
			q10 -= 2. * v * y[r];
			q01 += 2. * v * x[r];
		}
		const auto &rows = L.W->rows[ StencilIndex< Dim >( F1 , L.res ) ];
		long long i1 = (long long)idx;
		for( auto &row : rows )
		{
			long long i2 = nb[ row.f2 ];
			if( i2==-1 ) continue;
			for( auto &e : row.entries )
			{
				long long j1 = nb[ e.g1 ] , j2 = nb[ e.g2 ];
				if( j1==-1 || j2==-1 ) continue;
				double xy = ( x[j1] * y[j2] - x[j2] * y[j1] );
				double s = e.value;
				q10 += 2 * s * y[i2] * xy;
				q01 -= 2 * s * x[i2] * xy;
				if( j1==i1 )
				{
					q20 +=     s * y[i2] * y[j2];
					q02 +=     s * x[i2] * x[j2];
					q11 -= 2 * s * x[i2] * y[j2];
				}
				if( j2==i1 )
				{
					q20 -=     s * y[i2] * y[j1];
					q02 -=     s * x[i2] * x[j1];
					q11 += 2 * s * x[i2] * y[j1];
				}
			}
		}
		Q[0] = q20 , Q[1] = q02 , Q[2] = q11 , Q[3] = q10 , Q[4] = q01;
	}

	void qfit( int l , const double *x , const double *y , double *q ) const
	{
		const Level &L = levels[l];
		#pragma omp parallel for schedule(dynamic,64)
		for( long long i=0 ; i<(long long)L.N ; i++ ) fitNode( L , (size_t)i , x , y , q+5*i );
	}

	// the ω=½ block-Newton node update (ExteriorPoissonSolvers.h:173-181, Misha/Geometry.inl:319-331)
	static void NodeStep( const double Q[5] , double &dx , double &dy , bool &ok )
	{
		double H00 = 2*Q[0] , H11 = 2*Q[1] , H01 = Q[2] , H10 = Q[2];
		double det = H00*H11 - H01*H10;
		ok = fabs( det )>1e-20;
		dx = dy = 0;
		if( !ok ) return;
		double d = 1. / det;
		double i00 = H11 * d , i11 = H00 * d , i01 = -H01 * d , i10 = -H10 * d;
		double m00 = -i00 , m11 = -i11 , m01 = -i01 , m10 = -i10;
		double o0 = 0 , o1 = 0;
		o0 += Q[3]*m00 ; o1 += Q[3]*m01;
		o0 += Q[4]*m10 ; o1 += Q[4]*m11;
		dx = o0 * ( 1./2 ) , dy = o1 * ( 1./2 );
	}

	// SeparateGaussSeidelSolve's node update (ExteriorPoissonSolvers.h:211-236):
	// p = 0, then p[0] with p[1] fixed, then p[1] with the new p[0]; each guarded
	// by fabs( H(k,k)>1e-20 ), which is true exactly when H(k,k)>1e-20
	static void NodeStepSeparate( const double Q[5] , double &dx , double &dy )
	{
		const double H00 = 2*Q[0] , H11 = 2*Q[1] , H01 = Q[2] , H10 = Q[2] , g0 = Q[3] , g1 = Q[4];
		double p[2] = { 0 , 0 };
		if( fabs( (double)( H00>1e-20 ) ) ) p[0] = - ( ( H10 + H01 )*p[1] + g0 ) / ( 2 * H00 );
		if( fabs( (double)( H11>1e-20 ) ) ) p[1] = - ( ( H10 + H01 )*p[0] + g1 ) / ( 2 * H11 );
		dx = p[0] , dy = p[1];
	}

	void gs( int l , double *x , double *y , int iters ) const
	{
		const Level &L = levels[l];
		for( int it=0 ; it<iters ; it++ ) for( int c=0 ; c<NC ; c++ )
		{
			const std::vector< unsigned int > &col = L.colours[c];
			#pragma omp parallel for schedule(dynamic,128)
			for( long long j=0 ; j<(long long)col.size() ; j++ )
			{
				unsigned int idx = col[j];
				double Q[5] , dx , dy ; bool ok;
				fitNode( L , idx , x , y , Q );
				if( prm.separate ){ NodeStepSeparate( Q , dx , dy ) ; ok = true; }
				else NodeStep( Q , dx , dy , ok );
				if( ok ) x[idx] += dx , y[idx] += dy;
			}
		}
	}

	// wedge energy (Hat.products.inl:189-231), serial node order
	double wedge( int l , const double *x1 , const double *x2 ) const { return wedge4( l , x1 , x2 , x1 , x2 ); }
	double wedge4( int l , const double *x1 , const double *x2 , const double *y1 , const double *y2 ) const
	{
		const Level &L = levels[l];
		double integral = 0;
		for( size_t f1=0 ; f1<L.N ; f1++ )
		{
			Ix< Dim > F1 = FIndex< Dim >( f1 , L.res );
			long long nb[NW];
			for( int w=0 ; w<NW ; w++ )
			{
				Ix< Dim > G = F1 + W3< Dim >( w );
				bool in = true;
				for( int d=0 ; d<Dim ; d++ ) if( G[d]<0 || G[d]>L.res ) in = false;
				nb[w] = in ? (long long)FIndex< Dim >( G , L.res ) : -1;
			}
			for( auto &row : L.W->rows[ StencilIndex< Dim >( F1 , L.res ) ] )
			{
				long long f2 = nb[ row.f2 ];
				if( f2==-1 || (long long)f1>f2 ) continue;
				double xValue = x1[f1]*x2[f2] - x1[f2]*x2[f1];
				double value = 0;
				for( auto &e : row.entries )
				{
					long long g1 = nb[ e.g1 ] , g2 = nb[ e.g2 ];
					if( g1!=-1 && g2!=-1 ) value += e.value * ( y1[g1]*y2[g2] - y1[g2]*y2[g1] );
				}
				integral += value * xValue;
			}
		}
		return integral;
	}

	// scalar stiffness form x^T K y (Hat.scalars.inl:635-664)
	double stiffnessForm( int l , const double *x , const double *y ) const
	{
		const Level &L = levels[l];
		double integral = 0;
		for( size_t f1=0 ; f1<L.N ; f1++ )
		{
			Ix< Dim > F1 = FIndex< Dim >( f1 , L.res );
			for( auto &e : L.K->rows[ StencilIndex< Dim >( F1 , L.res ) ] )
			{
				Ix< Dim > G = F1 + W3< Dim >( e.first );
				integral += e.second * x[f1] * y[ FIndex< Dim >( G , L.res ) ];
			}
		}
		return integral;
	}

	void matvec( int l , char which , const double *x , double *y ) const
	{
		if( which=='P' ){ prolong( l , x , y ) ; return; }
		const Level &L = levels[l];
		const std::vector< double > &M = which=='R' ? L.R : L.B;
		for( size_t f=0 ; f<L.N ; f++ ) y[f] = 0;
		for( size_t f=0 ; f<L.N ; f++ )
		{
			Ix< Dim > F = FIndex< Dim >( f , L.res );
			for( int w=0 ; w<NW ; w++ )
			{
				double v = M[ f*NW+w ];
				if( v==0 ) continue;
				Ix< Dim > G = F + W3< Dim >( w );
				y[ FIndex< Dim >( G , L.res ) ] += v * x[f];
			}
		}
	}

	void energies( int l , const double *x , const double *y , double out[2] ) const
	{
		const Level &L = levels[l];
		std::vector< double > t( L.N );
		matvec( l , 'B' , x , &t[0] );
		double yBx = 0 ; for( size_t i=0 ; i<L.N ; i++ ) yBx += y[i]*t[i];
		out[0] = wedge( l , x , y ) - 2. * yBx + c_;
		matvec( l , 'R' , x , &t[0] );
		double xRx = 0 ; for( size_t i=0 ; i<L.N ; i++ ) xRx += t[i]*x[i];
		matvec( l , 'R' , y , &t[0] );
		double yRy = 0 ; for( size_t i=0 ; i<L.N ; i++ ) yRy += t[i]*y[i];
		out[1] = xRx + yRy;
		if( prm.dWeight ) out[1] += ( stiffnessForm( l , x , x ) + stiffnessForm( l , y , y ) ) * prm.dWeight;
	}

	// Random<double>() (Misha/Geometry.inl:429-438)
	static double Random( void )
	{
		static const unsigned long long ULL_RAND_MAX = (unsigned long long)( RAND_MAX ) + 1;
		static const double RAND_MAX_SQUARED = double( ULL_RAND_MAX * ULL_RAND_MAX - 1 );
		unsigned long long r1 = rand() , r2 = rand();
		long long foo = r1 * ULL_RAND_MAX + r2;
		return double( foo ) / RAND_MAX_SQUARED;
	}

	int minD( void ) const { return std::min( prm.minSolveDepth , D ); }
	int maxD( void ) const { return prm.maxSolveDepth<0 ? D : std::min( prm.maxSolveDepth , D ); }

	void init( double *x , double *y ) const
	{
		int l = minD();
		srand( prm.seed );
		if( adapted() ){ for( size_t f : levels[l].order ) x[f] = Random() , y[f] = Random(); }   // the hierarchy's numbering
		else for( size_t i=0 ; i<levels[l].N ; i++ ) x[i] = Random() , y[i] = Random();
		double s = wedge( l , x , y );
		s = sqrt( sqrt( c_ / s ) );
		for( size_t i=0 ; i<levels[l].N ; i++ ) x[i] *= s , y[i] *= s;
	}

	// cascadic solve (ExteriorPoisson.h:316-409)
	void solve( double *hx , double *hy , double *fx , double *fy , double E[2] ) const
	{
		int lo = minD() , hi = maxD();
		std::vector< std::vector< double > > X( D+1 ) , Y( D+1 );
		for( int l=0 ; l<=D ; l++ ) X[l].assign( levels[l].N , 0 ) , Y[l].assign( levels[l].N , 0 );
		init( &X[lo][0] , &Y[lo][0] );
		auto Iters = [&]( int d ){ return (int)ceil( pow( prm.iMult , D-d ) * prm.gsIters ); };
		std::vector< double > t;
		if( prm.singleLevel )
		{
			for( int d=lo ; d<hi ; d++ ){ prolong( d , &X[d][0] , &X[d+1][0] ) ; prolong( d , &Y[d][0] , &Y[d+1][0] ); }
			for( int d=lo ; d<hi ; d++ ) std::fill( X[d].begin() , X[d].end() , 0. ) , std::fill( Y[d].begin() , Y[d].end() , 0. );
			gs( hi , &X[hi][0] , &Y[hi][0] , prm.gsIters );
		}
		else
		{
			for( int d=lo ; d<=hi ; d++ )
			{
				gs( d , &X[d][0] , &Y[d][0] , Iters( d ) );
				if( d<hi )
				{
					t.resize( levels[d+1].N );
					prolong( d , &X[d][0] , &t[0] ) ; for( size_t i=0 ; i<t.size() ; i++ ) X[d+1][i] += t[i];
					prolong( d , &Y[d][0] , &t[0] ) ; for( size_t i=0 ; i<t.size() ; i++ ) Y[d+1][i] += t[i];
				}
			}
		}
		energies( hi , &X[hi][0] , &Y[hi][0] , E );
		for( int d=hi ; d<D ; d++ ){ prolong( d , &X[d][0] , &X[d+1][0] ) ; prolong( d , &Y[d][0] , &Y[d+1][0] ); }
		if( fx ) std::copy( X[D].begin() , X[D].end() , fx );
		if( fy ) std::copy( Y[D].begin() , Y[D].end() , fy );
		for( int d=D ; d>0 ; d-- )
		{
			t.resize( levels[d].N );
			prolong( d-1 , &X[d-1][0] , &t[0] ) ; for( size_t i=0 ; i<t.size() ; i++ ) X[d][i] -= t[i];
			prolong( d-1 , &Y[d-1][0] , &t[0] ) ; for( size_t i=0 ; i<t.size() ; i++ ) Y[d][i] -= t[i];
		}
		size_t o = 0;
		for( int d=0 ; d<=D ; d++ )
		{
			std::copy( X[d].begin() , X[d].end() , hx+o );
			std::copy( Y[d].begin() , Y[d].end() , hy+o );
			o += levels[d].N;
		}
	}

	size_t get( const std::string &name , double *out , size_t cap ) const
	{
		std::vector< double > v;
		if( name=="unit_pos" ) v = pos;
		else if( name=="skew_packed" ){ for( auto &s : skew ) for( int c=0 ; c<CD ; c++ ) v.push_back( s.p[c] ); }
		else if( name=="unit_to_world" ){ for( int c=0 ; c<=Dim ; c++ ) for( int r=0 ; r<=Dim ; r++ ) v.push_back( unitCubeToWorld[c][r] ); }
		else if( name=="cells" ){ for( auto &c : cells ) for( int d=0 ; d<Dim ; d++ ) v.push_back( c.first[d] ); }
		else if( name=="cell_counts" ){ for( auto &c : cells ) v.push_back( (double)c.second.size() ); }
		else if( name=="measure" ) v = measureV;
		else if( name=="noise" ) v = noiseV;
		else if( name=="sample_depth" ) v = depthV;
		else if( name=="S_elem" ) v = S;
		else if( name=="c" ) v.push_back( c_ );
		else return 0;
		if( out ) std::copy( v.begin() , v.begin()+std::min( cap , v.size() ) , out );
		return v.size();
	}

	// .tree layout (ExteriorPoissonRecon.cpp:259-270, Misha/Streams.h, Misha/EigenStreams.h:38-54)
	int writeTree( const char *path , const double *hx , const double *hy , unsigned long long hash ) const
	{
		FILE *fp = fopen( path , "wb" );
		if( !fp ) return -1;
		unsigned int dim = Dim;
		fwrite( &dim , sizeof(dim) , 1 , fp );
		fwrite( &hash , sizeof(hash) , 1 , fp );
		double v2w[Dim+1][Dim+1];
		{
			double g2c[Dim+1][Dim+1];
			for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ) g2c[i][j] = ( i==j ? 1. : 0. );
			for( int d=0 ; d<Dim ; d++ ) g2c[d][d] = 1./(1<<D);
			MatMul( unitCubeToWorld , g2c , v2w );
		}
		const unsigned long long vptr = 0;   // XForm written raw: the vtable pointer of its Algebra base first (Misha/Algebra.h:168)
		fwrite( &vptr , sizeof(vptr) , 1 , fp );
		fwrite( v2w , sizeof(double) , (Dim+1)*(Dim+1) , fp );
		for( int pass=0 ; pass<2 ; pass++ )
		{
			const double *h = pass==0 ? hx : hy;
			unsigned long long levelsN = D+1;
			fwrite( &levelsN , sizeof(levelsN) , 1 , fp );
			size_t o = 0;
			for( int d=0 ; d<=D ; d++ )
			{
				long long sz = (long long)levels[d].N;
				fwrite( &sz , sizeof(sz) , 1 , fp );
				fwrite( h+o , sizeof(double) , levels[d].N , fp );
				o += levels[d].N;
			}
		}
		unsigned int maxDepth = D;
		fwrite( &maxDepth , sizeof(maxDepth) , 1 , fp );
		fclose( fp );
		return 0;
	}
};
} // namespace

extern "C" {

void orc_default_params( orc_params *p )
{
	p->depth = 5 , p->seed = 0 , p->tRadius = 1 , p->dRadius = 1 , p->minSolveDepth = 0 , p->maxSolveDepth = -1 , p->gsIters = 10 , p->singleLevel = 0;
	p->separate = 0 , p->jitter = 0;
	p->nSigma = 0.25 , p->scale = 1.1 , p->sWeight = 100. , p->dWeight = 0.00625 , p->iMult = 2.;
}

void *orc_create( int dim , const orc_params *p , size_t n , const double *pos , const double *skew )
{
	return orc_create_adapted( dim , p , n , pos , skew , nullptr , nullptr , nullptr , nullptr );
}

void *orc_create_adapted( int dim , const orc_params *p , size_t n , const double *pos , const double *skew ,
                          const unsigned char *fmask , const unsigned char *emask , const long long *forder , const size_t *nforder )
{
	try
	{
		if( fmask && ( !emask || !forder || !nforder ) ){ g_error = "incomplete hierarchy" ; return nullptr; }
		if( dim==3 ){ Recon< 3 >::AdaptedInput a{ fmask , emask , forder , nforder } ; return static_cast< Base* >( new Recon< 3 >( *p , n , pos , skew , fmask ? &a : nullptr ) ); }
		if( dim==4 ){ Recon< 4 >::AdaptedInput a{ fmask , emask , forder , nforder } ; return static_cast< Base* >( new Recon< 4 >( *p , n , pos , skew , fmask ? &a : nullptr ) ); }
		g_error = "dimension must be 3 or 4";
	}
	catch( const std::string &s ){ g_error = s; }
	return nullptr;
}
void orc_destroy( void *h ){ delete static_cast< Base* >( h ); }
const char *orc_error( void ){ return g_error.c_str(); }
int orc_dim( void *h ){ return static_cast< Base* >( h )->dim(); }
size_t orc_num_functions( void *h , int l ){ return static_cast< Base* >( h )->numFunctions( l ); }
size_t orc_num_elements( void *h , int l ){ return static_cast< Base* >( h )->numElements( l ); }
size_t orc_get( void *h , const char *name , double *out , size_t cap ){ return static_cast< Base* >( h )->get( name , out , cap ); }
void orc_matvec( void *h , int l , char which , const double *x , double *y ){ static_cast< Base* >( h )->matvec( l , which , x , y ); }
void orc_energies( void *h , int l , const double *x , const double *y , double out[2] ){ static_cast< Base* >( h )->energies( l , x , y , out ); }
double orc_wedge( void *h , int l , const double *x , const double *y ){ return static_cast< Base* >( h )->wedge( l , x , y ); }
void orc_qfit( void *h , int l , const double *x , const double *y , double *q ){ static_cast< Base* >( h )->qfit( l , x , y , q ); }
void orc_gs( void *h , int l , double *x , double *y , int iters ){ static_cast< Base* >( h )->gs( l , x , y , iters ); }
void orc_init( void *h , double *x , double *y ){ static_cast< Base* >( h )->init( x , y ); }
void orc_solve( void *h , double *hx , double *hy , double *fx , double *fy , double E[2] ){ static_cast< Base* >( h )->solve( hx , hy , fx , fy , E ); }
int orc_write_tree( void *h , const char *path , const double *hx , const double *hy , unsigned long long hash ){ return static_cast< Base* >( h )->writeTree( path , hx , hy , hash ); }

}
