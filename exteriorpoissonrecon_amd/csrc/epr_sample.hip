// The reference's synthetic input generator (Sample/Sample.cpp with
// Include/Samples.h), host code: curve samples in 3D and surface samples in
// 4D with their frames, the unit-cube map, positional, angular and spot noise.
//
// Every value is computed with the operations, operands and order of the
// reference's build (clang -O3 for x86-64: no contraction into FMA, function
// arguments evaluated left to right), so the samples -- and the PLY files
// written from them -- are the reference Sample's bit for bit:
//  * the samplers' derivatives restate what Misha/AutoDiff's expression
//    templates evaluate (a scalar factor folded into the scale of a scaled
//    term, f - s as f + (-s), sums as right folds, d(f g) = f' g + f g',
//    d(c(u)) = c'(u) u' with d cos = sin * -1, constants differentiating to
//    an added 0);
//  * the matrix and point operations restate Misha/Geometry.inl (products
//    accumulated from 0 in index order, the Gauss-Jordan inverse, p / s as
//    p * (1/s), the 2x2/3x3 determinants of the generalised cross product);
//  * rand() is glibc's sequence (epr_rand.h: the program never seeds it),
//    the other draws use std::default_random_engine with libstdc++'s
//    normal and uniform distributions, as the reference does.
#pragma clang fp contract( off )
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/epr.h"
#include "epr_dev.h"
#include "epr_misha.h"
#include "epr_rand.h"

namespace epr
{
namespace sample
{
// the determinants of Geometry.inl:139-151 for the (N-1)x(N-1) minors
inline double Det2( const double c[2][2] ){ return c[0][0]*c[1][1] - c[0][1]*c[1][0]; }
inline double Det3( const double c[3][3] )
{
	return c[0][0]*( c[1][1]*c[2][2] - c[2][1]*c[1][2] ) + c[1][0]*( c[2][1]*c[0][2] - c[0][1]*c[2][2] ) + c[2][0]*( c[0][1]*c[1][2] - c[0][2]*c[1][1] );
}
template< int N >
double SubDeterminant( const Mat< N > &m , int col , int row )
{
	double t[N-1][N-1];
	for( int i=0 , ii=0 ; i<N ; i++ )
	{
		if( i==col ) continue;
		for( int j=0 , jj=0 ; j<N ; j++ ){ if( j==row ) continue ; t[ii][jj] = m.c[i][j] ; jj++; }
		ii++;
	}
	if constexpr( N==3 ) return Det2( t );
	else return Det3( t );
}
template< int N > double SquareNorm( const double *p ){ double s = 0 ; for( int i=0 ; i<N ; i++ ) s += p[i]*p[i] ; return s; }

// Geometry.inl:441-446 (the first rand() is the high part)
inline double Random2( GlibcRand &g )
{
	const long long r1 = g.rand();
	const long long t = r1*2147483647LL + g.rand();
	return ( double( t )/2147483648. )/2147483648.;
}
// Geometry.inl:447-505: a random orthonormal frame (rows of R) from rand()
template< int N >
Mat< N > RandomRotationMatrix( GlibcRand &g )
{
	double f[N][N];
	auto spherePoint = [&]( double *p )
	{
		while( true )
		{
			for( int d=0 ; d<N ; d++ ) p[d] = 1.0 - 2.0*Random2( g );
			if( SquareNorm< N >( p )<=1 ) break;
		}
		const double l = sqrt( SquareNorm< N >( p ) ) , s = 1./l;
		for( int d=0 ; d<N ; d++ ) p[d] *= s;
	};
	for( int d=0 ; d<N-1 ; d++ )
	{
		spherePoint( f[d] );
		while( true )
		{
			for( int dd=0 ; dd<d ; dd++ )
			{
				double dot = 0;
				for( int i=0 ; i<N ; i++ ) dot += f[d][i]*f[dd][i];
				for( int i=0 ; i<N ; i++ ) f[d][i] += -( f[dd][i]*dot );
			}
			if( SquareNorm< N >( f[d] )>1e-10 )
			{
				const double s = 1./sqrt( SquareNorm< N >( f[d] ) );
				for( int i=0 ; i<N ; i++ ) f[d][i] *= s;
				break;
			}
		}
	}
	Mat< N > x;
	for( int d=0 ; d<N-1 ; d++ ) for( int c=0 ; c<N ; c++ ) x( d , c ) = f[d][c];
	for( int d=0 ; d<N ; d++ ){ const double s = SubDeterminant( x , N-1 , d ) ; f[N-1][d] = ( d&1 ) ? -s : s; }
	Mat< N > R;
	for( int i=0 ; i<N ; i++ ) for( int j=0 ; j<N ; j++ ) R( i , j ) = f[i][j];
	return R;
}

// Hat::SkewSymmetricMatrix (Hat.inl:33-54): packed entries p[idx], idx over i>j
template< int N > struct Skew { double p[ N*(N-1)/2 ]; };
template< int N > Skew< N > SkewOf( const Mat< N > &m ){ Skew< N > s ; for( int i=0 , k=0 ; i<N ; i++ ) for( int j=0 ; j<i ; j++ , k++ ) s.p[k] = ( m( i , j ) - m( j , i ) )/2. ; return s; }
template< int N > Skew< N > SkewOf( const double *v1 , const double *v2 ){ Skew< N > s ; for( int i=0 , k=0 ; i<N ; i++ ) for( int j=0 ; j<i ; j++ , k++ ) s.p[k] = ( v1[j]*v2[i] - v1[i]*v2[j] )/2. ; return s; }
template< int N > Mat< N > Full( const Skew< N > &s ){ Mat< N > m ; for( int i=0 , k=0 ; i<N ; i++ ) for( int j=0 ; j<i ; j++ , k++ ) m( i , j ) = s.p[k] , m( j , i ) = -s.p[k] ; return m; }

// Samples.h:74-87: tangent <-> normal in 3D
inline Skew< 3 > Dual3( const double *t )
{
	Mat< 3 > n;
	for( int d=0 ; d<3 ; d++ ) n( (d+1)%3 , (d+2)%3 ) = t[d] , n( (d+2)%3 , (d+1)%3 ) = -t[d];
	return SkewOf( n );
}
inline void Dual3( const Skew< 3 > &s , double *t )
{
	const Mat< 3 > n = Full( s );
	for( int d=0 ; d<3 ; d++ ) t[d] = ( n( (d+1)%3 , (d+2)%3 ) - n( (d+2)%3 , (d+1)%3 ) )/2;
}
// Samples.h:798-808: the Hodge dual of a 2-blade in 4D
inline Skew< 4 > Dual4( const Skew< 4 > &t )
{
	const Mat< 4 > T = Full( t );
	Mat< 4 > n;
	n( 0 , 0 ) = 0         , n( 1 , 0 ) = T( 2 , 3 ) , n( 2 , 0 ) = T( 3 , 1 ) , n( 3 , 0 ) = T( 1 , 2 );
	n( 0 , 1 ) = T( 3 , 2 ) , n( 1 , 1 ) = 0         , n( 2 , 1 ) = T( 0 , 3 ) , n( 3 , 1 ) = T( 2 , 0 );
	n( 0 , 2 ) = T( 1 , 3 ) , n( 1 , 2 ) = T( 3 , 0 ) , n( 2 , 2 ) = 0         , n( 3 , 2 ) = T( 0 , 1 );
	n( 0 , 3 ) = T( 2 , 1 ) , n( 1 , 3 ) = T( 0 , 2 ) , n( 2 , 3 ) = T( 1 , 0 ) , n( 3 , 3 ) = 0;
	return SkewOf( n );
}

// ---------------- curves (Samples.h:91-645) ----------------
struct CurveSampler
{
	virtual ~CurveSampler(){}
	// the point and the tangent at parameter s
	virtual void eval( double s , double *p , double *t ) const = 0;
};

// spiral:<r> (Samples.h:321-331): (cos(X s), -1 + 2 s, sin(X s)), X = 2 pi r
struct Spiral : public CurveSampler
{
	double X;
	explicit Spiral( int r ) : X( 2.*M_PI*r ) {}
	void eval( double s , double *p , double *t ) const
	{
		const double u = s*X;
		p[0] = cos( u ) , p[1] = -1. + s*2. , p[2] = sin( u );
		t[0] = ( sin( u )*-1. )*( 1.*X ) , t[1] = 0. + 1.*2. , t[2] = cos( u )*( 1.*X );
	}
};

// torus_knot:<p>:<q> (Samples.h:441-556)
struct TorusKnot : public CurveSampler
{
	int p , q;
	unsigned int gcd;
	double ir , orad;
	TorusKnot( int _p , int _q , double inner=2./3 , double outer=1./3 ) : p( _p ) , q( _q ) , ir( inner ) , orad( outer )
	{
		// the reference's int-by-unsigned arithmetic (p and q converted to unsigned)
		const unsigned int mx = std::min< unsigned int >( std::abs( p ) , std::abs( q ) );
		if( mx==0 ) gcd = 1;
		else
		{
			for( gcd=mx ; gcd>=1 ; gcd-- ) if( ( (unsigned int)p%gcd )==0 && ( (unsigned int)q%gcd )==0 ) break;
			p = (int)( (unsigned int)p/gcd ) , q = (int)( (unsigned int)q/gcd );
		}
	}
	void eval( double s , double *P , double *T ) const
	{
		const int is = (int)floor( s*gcd );
		const double off = q ? ( 2.*M_PI*is )/(double)( gcd*q ) : 0.;
		// Theta = (S gcd + (-is)) 2 pi; Theta p and Theta q fold the factor into the scale
		const double th = s*(double)gcd + ( -(double)is );
		const double kp = 2.*M_PI*(double)p , kq = 2.*M_PI*(double)q;
		const double up = th*kp + off , uq = th*kq;
		const double sp = sin( up ) , cp = cos( up ) , sq = sin( uq ) , cq = cos( uq );
		const double R = cq*orad + ir;
		P[0] = R*cp , P[1] = R*sp , P[2] = sq*( -1.*orad );
		const double dth = 1.*(double)gcd + 0.;        // d(S gcd + const)
		const double dup = dth*kp + 0. , duq = dth*kq;
		const double dR = ( ( sq*-1. )*duq )*orad + 0.;
		const double dcp = ( sp*-1. )*dup , dsp = cp*dup;
		T[0] = dR*cp + R*dcp;
		T[1] = dR*sp + R*dsp;
		T[2] = ( cq*duq )*( -1.*orad );
	}
};

// borromean_rings (Samples.h:209-318): three rings of four circular arcs
struct Borromean : public CurveSampler
{
	double r , h , alpha , theta , length , centers[4][3] , borders[4];
	const int orient[3][2] = { { 0 , 1 } , { 1 , 2 } , { 2 , 0 } };
	explicit Borromean( double _r=0.25 ) : r( _r )
	{
		h = 0.5*( sqrt( 7 )-1 )*r;
		alpha = sin( h/( 2.*r ) )/( 2.*M_PI );
		theta = 1. - 2.*acos( h/( 2.*r ) )/( 2.*M_PI );
		length = 4*alpha + 2*theta;
		const double c[4][3] = { { r+h , 0. , 0. } , { 0. , h , 0. } , { -( r+h ) , 0. , 0. } , { 0. , -h , 0. } };
		memcpy( centers , c , sizeof( c ) );
		const double b[4] = { 2.*alpha , 2.*alpha+theta , 4.*alpha+theta , 4.*alpha + 2.*theta };
		memcpy( borders , b , sizeof( b ) );
	}
	void eval( double s , double *P , double *T ) const
	{
		int k = 0;
		if( 1./3<=s && s<2./3 ) k = 1 , s -= 1./3.;
		else if( 2./3<=s ) k = 2 , s -= 2./3.;
		const double L3 = 3.*length;          // L = S 3 length
		const double l = length*s*3;
		double a , da;                        // Angle and its derivative; the arc's centre
		int ci;
		if( l<borders[0] )
		{
			// (0.5 + alpha) - L: Constant + (S * (3 length * -1))
			a = ( 0.5 + alpha ) + s*( L3*-1. ) , da = 0. + 1.*( L3*-1. ) , ci = 0;
		}
		else if( borders[0]<=l && l<borders[1] )
		{
			a = ( 0.75 + ( 1.-theta )/2. - borders[0] ) + s*L3 , da = 0. + 1.*L3 , ci = 1;
		}
		else if( borders[1]<=l && l<borders[2] )
		{
			// alpha - ( L - b1 ): Constant + ( S 3 length + (-b1) ) * -1
			a = alpha + ( s*L3 + ( -borders[1] ) )*-1. , da = 0. + ( 1.*L3 + 0. )*-1. , ci = 2;
		}
		else
		{
			a = ( 0.75 - theta/2. - borders[2] ) + s*L3 , da = 0. + 1.*L3 , ci = 3;
		}
		// F1 = c0 + cos( Angle 2 pi ) r, F2 = c1 + sin( Angle 2 pi ) r
		const double u = a*( 2.*M_PI ) , du = da*( 2.*M_PI );
		double p[3] = { 0 , 0 , 0 } , t[3] = { 0 , 0 , 0 };
		p[ orient[k][0] ] = centers[ci][0] + cos( u )*r;
		p[ orient[k][1] ] = centers[ci][1] + sin( u )*r;
		t[ orient[k][0] ] = 0. + ( ( sin( u )*-1. )*du )*r;
		t[ orient[k][1] ] = 0. + ( cos( u )*du )*r;
		const double sc = 1./( r + h );
		for( int d=0 ; d<3 ; d++ ) P[d] = p[d]*sc , T[d] = t[d]*sc;
	}
};

// spherical:1:<n>:<l> (Samples.h:577-630, the type the Hopf torus uses):
// (cos(l cos(a s)) cos(b s), cos(l cos(a s)) sin(b s), sin(l cos(a s))), a = 2 n pi, b = 2 pi
struct SphericalCurve1 : public CurveSampler
{
	int n;
	double l;
	SphericalCurve1( int _n , double _l ) : n( _n ) , l( _l ) {}
	void eval( double s , double *P , double *T ) const
	{
		const double a = 2.*(double)n*M_PI , b = 2.*M_PI;
		const double u = s*a , cu = cos( u ) , su = sin( u );
		const double v = cu*l;                                  // l Cos(S a)
		const double dv = ( ( su*-1. )*( 1.*a ) )*l;
		const double ct1 = cos( v ) , st1 = sin( v );
		const double w = s*b , ct2 = cos( w ) , st2 = sin( w );
		const double dct1 = ( st1*-1. )*dv , dst1 = ct1*dv;
		const double dct2 = ( st2*-1. )*( 1.*b ) , dst2 = ct2*( 1.*b );
		P[0] = ct1*ct2 , P[1] = ct1*st2 , P[2] = st1;
		T[0] = dct1*ct2 + ct1*dct2;
		T[1] = dct1*st2 + ct1*dst2;
		T[2] = dst1;
	}
};

CurveSampler *GetCurveSampler( const std::string &str )
{
	std::istringstream ss( str );
	std::string head;
	if( !std::getline( ss , head , ':' ) ) return nullptr;
	auto arg = [&]( const char *what ) -> std::string
	{
		std::string a;
		if( !std::getline( ss , a , ':' ) ) throw std::string( "Could not read " ) + what + ": " + str;
		return a;
	};
	if( head=="spiral" ) return new Spiral( std::stoi( arg( "spiral r" ) ) );
	if( head=="torus_knot" ){ const int p = std::stoi( arg( "torus knot p" ) ) , q = std::stoi( arg( "torus knot q" ) ) ; return new TorusKnot( p , q ); }
	if( head=="borromean_rings" ) return new Borromean();
	if( head=="spherical" )
	{
		const int type = std::stoi( arg( "spherical curve type" ) ) , n = std::stoi( arg( "spherical curve n" ) );
		const double l = std::stod( arg( "spherical curve l" ) );
		if( type!=1 ) throw std::string( "spherical curves of type 0 are not supported: " ) + str;
		return new SphericalCurve1( n , l );
	}
	return nullptr;
}

// Samples.h:733-784 (no multiplier): random parameters from the default engine, or (i+0.5)/count
bool CurveSamples( unsigned int count , const std::string &str , bool random , std::vector< double > &pos , std::vector< double > &skew )
{
	std::unique_ptr< CurveSampler > smp( GetCurveSampler( str ) );
	if( !smp ) return false;
	pos.resize( (size_t)count*3 ) , skew.resize( (size_t)count*3 );
	auto one = [&]( unsigned int i , double s )
	{
		double t[3];
		smp->eval( s , &pos[ (size_t)i*3 ] , t );
		const Skew< 3 > n = Dual3( t );
		for( int k=0 ; k<3 ; k++ ) skew[ (size_t)i*3+k ] = n.p[k];
	};
	if( random )
	{
		std::default_random_engine generator;
		std::uniform_real_distribution< double > distribution( 0. , 1. );
		for( unsigned int i=0 ; i<count ; i++ ) one( i , distribution( generator ) );
	}
	else for( unsigned int i=0 ; i<count ; i++ ) one( i , (i+0.5)/count );
	return true;
}

// ---------------- surfaces (Samples.h:811-1361) ----------------
struct SurfaceSampler
{
	virtual ~SurfaceSampler(){}
	virtual void eval( double s , double t , double *p , Skew< 4 > &n ) const = 0;
};

// clifford_torus (Samples.h:818-857): (cos X s, sin X s, cos X t, sin X t)/sqrt(2), frame = d/ds ^ d/dt
struct Clifford : public SurfaceSampler
{
	void eval( double s , double t , double *P , Skew< 4 > &N ) const
	{
		const double X = 2.*M_PI , k = 1./sqrt( 2. );
		const double us = s*X , ut = t*X;
		P[0] = cos( us )*k , P[1] = sin( us )*k , P[2] = cos( ut )*k , P[3] = sin( ut )*k;
		const double d1 = ( ( sin( us )*-1. )*( 1.*X ) )*k , d2 = ( cos( us )*( 1.*X ) )*k;
		const double d3 = ( ( sin( ut )*-1. )*( 1.*X ) )*k , d4 = ( cos( ut )*( 1.*X ) )*k;
		const double v1[4] = { d1 , d2 , 0 , 0 } , v2[4] = { 0 , 0 , d3 , d4 };
		N = SkewOf< 4 >( v1 , v2 );
	}
};

// hopf_torus:<n>:<l> (Samples.h:1069-1107): the Hopf preimage of the spherical curve
struct Hopf : public SurfaceSampler
{
	SphericalCurve1 curve;
	Hopf( int n , double l ) : curve( n , l ) {}
	void eval( double s , double t , double *P , Skew< 4 > &N ) const
	{
		double cp[3] , tg[3] , ct[3];
		curve.eval( s , cp , tg );
		Dual3( Dual3( tg ) , ct );   // the curve's normal, back to the tangent (Curve::Dual twice)
		const double _t = t*2*M_PI;
		const double cost = cos( _t ) , sint = sin( _t );
		const double dcosdt = -sin( _t ) , dsindt = cos( _t );
		const double a = cp[0] , b = cp[1] , c = cp[2];
		const double dads = ct[0] , dbds = ct[1] , dcds = ct[2];
		const double factor = 1./sqrt( 2*( 1+c ) );
		const double dfactords = -dcds/( pow( 2*( 1+c ) , 1.5 ) );
		P[0] = ( 1+c )*cost*factor;
		P[1] = ( a*sint - b*cost )*factor;
		P[2] = ( a*cost + b*sint )*factor;
		P[3] = ( 1+c )*sint*factor;
		const double dt[4] = { ( 1+c )*dcosdt*factor , ( a*dsindt - b*dcosdt )*factor , ( a*dcosdt + b*dsindt )*factor , ( 1+c )*dsindt*factor };
		const double ds[4] = { dcds*cost*factor + ( 1+c )*cost*dfactords , ( dads*sint - dbds*cost )*factor + ( a*sint - b*cost )*dfactords ,
		                       ( dads*cost + dbds*sint )*factor + ( a*cost + b*sint )*dfactords , dcds*sint*factor + ( 1+c )*sint*dfactords };
		N = Dual4( SkewOf< 4 >( dt , ds ) );
	}
};

SurfaceSampler *GetSurfaceSampler( const std::string &str )
{
	std::istringstream ss( str );
	std::string head;
	if( !std::getline( ss , head , ':' ) ) return nullptr;
	if( head=="clifford_torus" ) return new Clifford();
	if( head=="hopf_torus" )
	{
		std::string n , l;
		if( !std::getline( ss , n , ':' ) ) throw std::string( "Could not read Hopf torus n: " ) + str;
		if( !std::getline( ss , l , ':' ) ) throw std::string( "Could not read Hopf torus l: " ) + str;
		return new Hopf( std::stoi( n ) , std::stod( l ) );
	}
	return nullptr;
}

// Samples.h:1282-1361: count^2 samples, (s, t) drawn in that order, or the (i+0.5)/count grid
bool SurfaceSamples( unsigned int count , const std::string &str , bool random , std::vector< double > &pos , std::vector< double > &skew )
{
	std::unique_ptr< SurfaceSampler > smp( GetSurfaceSampler( str ) );
	if( !smp ) return false;
	const size_t n = (size_t)count*count;
	pos.resize( n*4 ) , skew.resize( n*6 );
	auto one = [&]( size_t i , double s , double t )
	{
		Skew< 4 > N;
		smp->eval( s , t , &pos[ i*4 ] , N );
		for( int k=0 ; k<6 ; k++ ) skew[ i*6+k ] = N.p[k];
	};
	if( random )
	{
		std::default_random_engine generator;
		std::uniform_real_distribution< double > distribution( 0. , 1. );
		for( size_t i=0 ; i<n ; i++ ){ const double s = distribution( generator ) ; const double t = distribution( generator ) ; one( i , s , t ); }
	}
	else for( unsigned int i=0 ; i<count ; i++ ) for( unsigned int j=0 ; j<count ; j++ ) one( (size_t)i*count+j , (i+0.5)/count , (j+0.5)/count );
	return true;
}

// ---------------- Sample.cpp: unit cube and noise ----------------
// Sample.cpp:87-118: the bounding-box centre to 0.5, the shape diameter (twice
// the largest distance to the centroid) to 1
template< int N >
Mat< N+1 > ToUnitCube( const std::vector< double > &pos , size_t n )
{
	if( !n ) throw std::string( "Expected some samples: 0" );
	double b0[N] , b1[N];
	for( int d=0 ; d<N ; d++ ) b0[d] = b1[d] = pos[d];
	for( size_t i=1 ; i<n ; i++ ) for( int d=0 ; d<N ; d++ ){ b0[d] = std::min< double >( b0[d] , pos[i*N+d] ) ; b1[d] = std::max< double >( b1[d] , pos[i*N+d] ); }
	double center[N] = {};
	for( size_t i=0 ; i<n ; i++ ) for( int d=0 ; d<N ; d++ ) center[d] += pos[i*N+d];
	{
		const double s = 1./(double)n;
		for( int d=0 ; d<N ; d++ ) center[d] *= s;
	}
	double sc = 0;
	for( size_t i=0 ; i<n ; i++ )
	{
		double v[N];
		for( int d=0 ; d<N ; d++ ) v[d] = -pos[i*N+d] + center[d];
		sc = std::max< double >( sc , SquareNorm< N >( v ) );
	}
	sc = sqrt( sc )*2;
	Mat< N+1 > t1 = Mat< N+1 >::Identity() , s = Mat< N+1 >::Identity() , t2 = Mat< N+1 >::Identity();
	for( int d=0 ; d<N ; d++ ) t2( N , d ) = -( b0[d] + b1[d] )/2.;
	for( int d=0 ; d<N ; d++ ) s( d , d ) = 1./sc;
	for( int d=0 ; d<N ; d++ ) t1( N , d ) = 0.5;
	return ( t1*s )*t2;
}

// Sample.cpp:120-201, in the unit cube: positional noise (a per-sample delta
// redrawn until the moved sample stays in the open cube), angular noise (the
// frame conjugated by a random rotation: planar rotations by normal angles,
// in a random orthonormal frame), spot noise (ceil(n sNoise) uniform samples
// with uniform skew entries, appended)
template< int N >
void AddNoise( std::vector< double > &pos , std::vector< double > &skew , double pNoise , double aNoise , double sNoise , GlibcRand &rnd )
{
	constexpr int K = N*(N-1)/2;
	const size_t n = pos.size()/N;
	if( pNoise>0 )
	{
		std::default_random_engine generator;
		std::normal_distribution< double > distribution( 0. , pNoise );
		for( size_t i=0 ; i<n ; i++ )
		{
			double delta[N] = {};
			while( true )
			{
				bool in = true;
				for( int d=0 ; d<N && in ; d++ )
				{
					delta[d] = distribution( generator );
					in &= pos[i*N+d] + delta[d]>0 && pos[i*N+d] + delta[d]<1;
				}
				if( in ) break;
			}
			for( int d=0 ; d<N ; d++ ) pos[i*N+d] += delta[d];
		}
	}
	if( aNoise>0 )
	{
		std::default_random_engine generator;
		std::normal_distribution< double > distribution( 0. , aNoise*M_PI/180. );
		for( size_t i=0 ; i<n ; i++ )
		{
			// Samples.h:53-69
			Mat< N > R = Mat< N >::Identity();
			for( int d=0 ; d<N/2 ; d++ )
			{
				const double th = distribution( generator );
				const double ct = cos( th ) , st = sin( th );
				R( 2*d , 2*d ) = ct , R( 2*d , 2*d+1 ) = st , R( 2*d+1 , 2*d ) = -st , R( 2*d+1 , 2*d+1 ) = ct;
			}
			const Mat< N > Rr = RandomRotationMatrix< N >( rnd );
			R = ( Rr.transpose()*R )*Rr;
			Skew< N > sk;
			for( int k=0 ; k<K ; k++ ) sk.p[k] = skew[i*K+k];
			const Skew< N > out = SkewOf< N >( ( R.transpose()*Full( sk ) )*R );
			for( int k=0 ; k<K ; k++ ) skew[i*K+k] = out.p[k];
		}
	}
	if( sNoise )
	{
		std::default_random_engine generator;
		std::uniform_real_distribution< double > pd( 0. , 1. ) , sd( -1. , 1. );
		const unsigned int m = (unsigned int)ceil( n*sNoise );
		for( unsigned int i=0 ; i<m ; i++ )
		{
			for( int d=0 ; d<N ; d++ ) pos.push_back( pd( generator ) );
			Mat< N > sk;
			for( int j=0 ; j<N ; j++ ) for( int k=0 ; k<j ; k++ ){ sk( j , k ) = sd( generator ) ; sk( k , j ) = -sk( j , k ); }
			const Skew< N > s = SkewOf( sk );
			for( int k=0 ; k<K ; k++ ) skew.push_back( s.p[k] );
		}
	}
}

// Sample.cpp:203-230: to the unit cube, the noise, back to world coordinates;
// frames as full matrices (m(c,r) at [i][c][r])
template< int N >
void Finish( std::vector< double > &pos , std::vector< double > &skew , double pNoise , double aNoise , double sNoise , std::vector< double > &full )
{
	constexpr int K = N*(N-1)/2;
	const size_t n0 = pos.size()/N;
	const Mat< N+1 > w2u = ToUnitCube< N >( pos , n0 ) , u2w = w2u.inverse();
	for( size_t i=0 ; i<n0 ; i++ ){ double o[N] ; Apply< N >( w2u , &pos[i*N] , o ) ; for( int d=0 ; d<N ; d++ ) pos[i*N+d] = o[d]; }
	GlibcRand rnd( 1 );
	AddNoise< N >( pos , skew , pNoise , aNoise , sNoise , rnd );
	const size_t n = pos.size()/N;
	for( size_t i=0 ; i<n ; i++ ){ double o[N] ; Apply< N >( u2w , &pos[i*N] , o ) ; for( int d=0 ; d<N ; d++ ) pos[i*N+d] = o[d]; }
	full.resize( n*N*N );
	for( size_t i=0 ; i<n ; i++ )
	{
		Skew< N > s;
		for( int k=0 ; k<K ; k++ ) s.p[k] = skew[i*K+k];
		const Mat< N > m = Full( s );
		for( int c=0 ; c<N ; c++ ) for( int r=0 ; r<N ; r++ ) full[ i*N*N + c*N + r ] = m( c , r );
	}
}
} // namespace sample
} // namespace epr

extern "C" int epr_sample_generate( const char *type , unsigned int res , int regular , double aNoise , double pNoise , double sNoise ,
                                    int *dim , size_t *count , double *pos , double *skew , size_t cap )
{
	return epr::ContextFree( [&]()
	{
		using namespace epr::sample;
		if( !type || !dim || !count ) throw std::string( "bad arguments" );
		std::vector< double > p , s , full;
		int D = 3;
		if( !CurveSamples( res , type , !regular , p , s ) )
		{
			D = 4;
			if( !SurfaceSamples( res , type , !regular , p , s ) ) throw std::string( "Could not parse sample type: " ) + type;
		}
		if( D==3 ) Finish< 3 >( p , s , pNoise , aNoise , sNoise , full );
		else       Finish< 4 >( p , s , pNoise , aNoise , sNoise , full );
		const size_t n = p.size()/D;
		*dim = D , *count = n;
		if( pos && skew && cap>=n )
		{
			memcpy( pos , p.data() , p.size()*sizeof(double) );
			memcpy( skew , full.data() , full.size()*sizeof(double) );
		}
	} );
}
