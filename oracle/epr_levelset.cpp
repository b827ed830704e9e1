// CPU restatement (test infrastructure, the oracle) of the reference's
// simplicial extraction: the triangulation of a list of grid cells and the
// codimension-1/2 level sets of vertex values over a simplicial mesh.
//
//   HyperCubeSimplices<Dim>            Misha/HyperCubeSimplices.h:47-170
//   RegularSubGridTriangulation        Misha/MarchingSimplices.h:622-706
//   LevelSet / _AddLevelSet            Misha/MarchingSimplices.h:410-576
//   OrientedBoundaryFace               Misha/MarchingSimplices.h:82-90
//   MultiIndex (sorted index tuples)   Misha/MarchingSimplices.h:775-795
//
// The reference solves the 2x2 barycentric system with Eigen 3.3.7's
// FullPivHouseholderQR (vendored at Eigen/src/QR/FullPivHouseholderQR.h:456-560,
// Householder reflections of Eigen/src/Householder/Householder.h:62-127);
// SolveQR2 restates that algorithm's arithmetic for the 2x2 case (full
// pivoting by the first strictly largest |entry| in column-major order, one
// Householder reflection, upper-triangular back substitution), so that the
// interior/snapping decisions and the interpolated vertex data agree bit for
// bit.  Vertex numbering follows the reference's first-creation order.
//
// Only tests/ and the CPU checks use this file; the product never links it.
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "epr_oracle.h"

namespace
{
// ---------------------------------------------------------------- cube simplices
// Vertex v of the unit Dim-cube has coordinate d = bit d of v; simplices of the
// Dim-cube are the cones from the origin over the triangulated far faces
// {x_dir = 1}, dir = 0..Dim-1 (HyperCubeSimplices.h:126-134).
struct CubeSimplices
{
	int dim;
	std::vector< std::vector< int > > simplices;          // vertex indices (bits = coordinates)
	explicit CubeSimplices( int d ) : dim( d )
	{
		if( d==1 ){ simplices = { { 0 , 1 } } ; return; }
		CubeSimplices lower( d-1 );
		for( int dir=0 ; dir<d ; dir++ ) for( const std::vector< int > &ls : lower.simplices )
		{
			std::vector< int > s;
			for( int lv : ls )
			{
				// insert coordinate `dir` = 1 into the (d-1)-dimensional vertex lv
				int v = 0;
				for( int c=0 , lc=0 ; c<d ; c++ ) v |= ( c==dir ? 1 : ( lv>>lc++ ) & 1 ) << c;
				s.push_back( v );
			}
			s.push_back( 0 );
			simplices.push_back( s );
		}
	}
};

// exact determinant of a small integer matrix (entries in {-1,0,1}), Laplace expansion
static long long IntDet( std::vector< std::vector< long long > > m )
{
	const int n = (int)m.size();
	if( n==1 ) return m[0][0];
	long long det = 0;
	for( int c=0 ; c<n ; c++ )
	{
		std::vector< std::vector< long long > > sub;
		for( int r=1 ; r<n ; r++ ){ std::vector< long long > row; for( int cc=0 ; cc<n ; cc++ ) if( cc!=c ) row.push_back( m[r][cc] ) ; sub.push_back( row ); }
		det += ( c&1 ? -1 : 1 ) * m[0][c] * IntDet( sub );
	}
	return det;
}

// ---------------------------------------------------------------- 2x2 QR solve
static void SolveQR2( const double M[2][2] , const double b[2] , double x[2] )
{
	double m[2][2] = { { M[0][0] , M[0][1] } , { M[1][0] , M[1][1] } };
	int rowT[2] = { 0 , 1 } , colT[2] = { 0 , 1 };
	double hc[2] = { 0 , 0 };
	int nonzero = 2;
	double maxpivot = 0 , biggest = 0;
	const double prec = DBL_EPSILON * 2;
	for( int k=0 ; k<2 ; k++ )
	{
		int bi = k , bj = k;
		double best = std::fabs( m[k][k] );
		for( int j=k ; j<2 ; j++ ) for( int i=k ; i<2 ; i++ )
		{
			if( i==k && j==k ) continue;
			const double v = std::fabs( m[i][j] );
			if( v>best ) best = v , bi = i , bj = j;
		}
		if( k==0 ) biggest = best;
		if( best<=biggest*prec )
		{
			nonzero = k;
			for( int i=k ; i<2 ; i++ ) rowT[i] = colT[i] = i , hc[i] = 0;
			break;
		}
		rowT[k] = bi , colT[k] = bj;
		if( k!=bi ) for( int c=k ; c<2 ; c++ ) std::swap( m[k][c] , m[bi][c] );
		if( k!=bj ) for( int r=0 ; r<2 ; r++ ) std::swap( m[r][k] , m[r][bj] );
		double tau , beta;
		if( k==1 ) tau = 0 , beta = m[1][1];
		else
		{
			const double tailSq = m[1][0]*m[1][0] , c0 = m[0][0];
			if( tailSq<=DBL_MIN ) tau = 0 , beta = c0 , m[1][0] = 0;
			else
			{
				beta = std::sqrt( c0*c0 + tailSq );
				if( c0>=0 ) beta = -beta;
				m[1][0] = m[1][0] / ( c0 - beta );
				tau = ( beta - c0 ) / beta;
			}
		}
		m[k][k] = beta , hc[k] = tau;
		if( std::fabs( beta )>maxpivot ) maxpivot = std::fabs( beta );
		if( k==0 && tau!=0 )
		{
			const double ess = m[1][0];
			double tmp = ess*m[1][1];
			tmp += m[0][1];
			m[0][1] -= tau*tmp;
			m[1][1] -= ( tau*ess )*tmp;
		}
	}
	int perm[2] = { 0 , 1 };
	for( int k=0 ; k<2 ; k++ ) std::swap( perm[k] , perm[ colT[k] ] );
	int rank = 0;
	for( int i=0 ; i<nonzero ; i++ ) rank += std::fabs( m[i][i] ) > maxpivot*prec;
	if( rank==0 ){ x[0] = x[1] = 0 ; return; }
	double c[2] = { b[0] , b[1] };
	for( int k=0 ; k<rank ; k++ )
	{
		std::swap( c[k] , c[ rowT[k] ] );
		if( k==1 ) c[1] *= 1. - hc[1];
		else if( hc[0]!=0 )
		{
			const double ess = m[1][0];
			double tmp = ess*c[1];
			tmp += c[0];
			c[0] -= hc[0]*tmp;
			c[1] -= ( hc[0]*ess )*tmp;
		}
	}
	for( int i=rank-1 ; i>=0 ; i-- ) if( c[i]!=0 )
	{
		c[i] /= m[i][i];
		for( int r=0 ; r<i ; r++ ) c[r] -= c[i]*m[r][i];
	}
	for( int i=0 ; i<rank ; i++ ) x[ perm[i] ] = c[i];
	for( int i=rank ; i<2 ; i++ ) x[ perm[i] ] = 0;
}

// ---------------------------------------------------------------- level set
typedef std::vector< unsigned > Simp;                   // vertex indices of one simplex

struct LevelSetState
{
	int codim , vdim;                                   // values per vertex, data per vertex
	const double *data , *values , *iso;
	std::map< std::vector< unsigned > , unsigned > vMap;   // sorted index tuple -> output vertex
	std::vector< double > outData;
	double value( unsigned v , int c ) const { return values[ (size_t)v*codim + c ]; }
};

static Simp OrientedBoundaryFace( const Simp &si , unsigned v )
{
	const unsigned K = (unsigned)si.size()-1;      // the simplex dimension
	Simp f( K );
	for( unsigned d=0 ; d<K ; d++ ) f[d] = si[ ( v+d+1 )%( K+1 ) ];
	if( ( K&1 ) && ( v&1 ) ) std::swap( f[0] , f[1] );
	return f;
}

static std::vector< unsigned > Sorted( const Simp &s ){ std::vector< unsigned > t = s ; std::sort( t.begin() , t.end() ) ; return t; }

// out gets simplices of dimension (si.size()-1) - codim
static void AddLevelSet( std::vector< Simp > &out , const Simp &si , LevelSetState &S )
{
	const int Dim = (int)si.size()-1 , CoDim = S.codim;
	// bounding-sphere rejection in value space (MarchingSimplices.h:424-431)
	{
		double center[2] = { 0 , 0 };
		for( int d=0 ; d<=Dim ; d++ ) for( int c=0 ; c<CoDim ; c++ ) center[c] += S.value( si[d] , c );
		const double inv = 1. / (double)( Dim+1 );
		for( int c=0 ; c<CoDim ; c++ ) center[c] *= inv;
		double r2 = 0;
		for( int d=0 ; d<=Dim ; d++ )
		{
			double n2 = 0;
			for( int c=0 ; c<CoDim ; c++ ){ const double t = -S.value( si[d] , c ) + center[c] ; n2 += t*t; }
			r2 = std::max< double >( r2 , n2 );
		}
		double n2 = 0;
		for( int c=0 ; c<CoDim ; c++ ){ const double t = -S.iso[c] + center[c] ; n2 += t*t; }
		if( n2>r2 ) return;
	}
	if( Dim==CoDim && CoDim==2 )
	{
		auto InList = [&]( unsigned v ){ for( const Simp &s : out ) if( s[0]==v ) return true ; return false; };
		std::vector< unsigned > idx = Sorted( si );
		auto iter = S.vMap.find( idx );
		if( iter!=S.vMap.end() && !InList( iter->second ) ){ out.push_back( Simp{ iter->second } ) ; return; }
		double M[2][2] , rhs[2] , a_[2];
		for( int c=0 ; c<2 ; c++ ) for( int r=0 ; r<2 ; r++ ) M[r][c] = S.value( si[c+1] , r ) - S.value( si[0] , r );
		for( int i=0 ; i<2 ; i++ ) rhs[i] = S.iso[i] - S.value( si[0] , i );
		SolveQR2( M , rhs , a_ );
		double a[3];
		a[0] = 1.;
		for( int i=0 ; i<2 ; i++ ) a[0] -= a_[i] , a[i+1] = a_[i];
		const double EPS = 1e-15;
		bool interior = true , recheck = false;
		Simp _si = si;
		for( int i=0 ; i<=2 ; i++ )
		{
			interior &= a[i]>=0 && a[i]<=1.;
			if( std::fabs( a[i] )<EPS ) a[i] = 0 , _si[i] = (unsigned)-1 , recheck = true;
		}
		if( !interior ) return;
		if( recheck )
		{
			idx = Sorted( _si );
			auto it2 = S.vMap.find( idx );
			if( it2!=S.vMap.end() ){ if( !InList( it2->second ) ) out.push_back( Simp{ it2->second } ) ; return; }
		}
		const unsigned v = (unsigned)( S.outData.size() / S.vdim );
		S.vMap[idx] = v;
		for( int k=0 ; k<S.vdim ; k++ )
		{
			double d = S.data[ (size_t)si[0]*S.vdim+k ] * a[0];
			for( int i=1 ; i<=2 ; i++ ) d += S.data[ (size_t)si[i]*S.vdim+k ] * a[i];
			S.outData.push_back( d );
		}
		out.push_back( Simp{ v } );
	}
	else if( CoDim==1 && Dim==2 )
	{
		double cv[3];
		for( int i=0 ; i<3 ; i++ ){ cv[i] = S.value( si[i] , 0 ) - S.iso[0] ; if( cv[i]==0 ) cv[i] = 1e-10; }
		int ii = -1;
		for( int d=0 ; d<3 ; d++ ) if( cv[d]*cv[(d+1)%3]<0 && cv[d]*cv[(d+2)%3]<0 ){ ii = d ; break; }
		if( ii<0 ) return;
		Simp isoEdge( 2 );
		for( int i=0 ; i<2 ; i++ )
		{
			const unsigned e0 = si[ii] , e1 = si[ (ii+1+i)%3 ];
			std::vector< unsigned > idx = Sorted( Simp{ e0 , e1 } );
			auto iter = S.vMap.find( idx );
			if( iter!=S.vMap.end() ) isoEdge[i] = iter->second;
			else
			{
				const double s = - cv[ii] / ( cv[ (ii+1+i)%3 ] - cv[ii] );
				const unsigned v = (unsigned)( S.outData.size() / S.vdim );
				isoEdge[i] = S.vMap[idx] = v;
				for( int k=0 ; k<S.vdim ; k++ ) S.outData.push_back( S.data[ (size_t)e0*S.vdim+k ] * ( 1.-s ) + S.data[ (size_t)e1*S.vdim+k ] * s );
			}
		}
		if( cv[ii]<0 ) std::swap( isoEdge[0] , isoEdge[1] );
		out.push_back( isoEdge );
	}
	else
	{
		std::vector< Simp > sub;
		for( int i=0 ; i<=Dim ; i++ ) AddLevelSet( sub , OrientedBoundaryFace( si , i ) , S );
		if( sub.empty() ) return;
		const unsigned v = sub[0][0];
		const int K = Dim - CoDim;                     // vertices per sub-simplex
		for( size_t i=1 ; i<sub.size() ; i++ )
		{
			bool containsV = false;
			for( int j=0 ; j<K ; j++ ) containsV |= sub[i][j]==v;
			if( !containsV ){ Simp s( K+1 ) ; s[0] = v ; for( int j=0 ; j<K ; j++ ) s[j+1] = sub[i][j] ; out.push_back( s ); }
		}
	}
}

template< typename T > static T *Copy( const std::vector< T > &v )
{
	T *p = (T *)malloc( std::max< size_t >( 1 , v.size() )*sizeof(T) );
	if( p && !v.empty() ) memcpy( p , v.data() , v.size()*sizeof(T) );
	return p;
}
} // namespace

extern "C" {

int orc_cube_simplices( int dim , int *out )
{
	if( dim<1 || dim>4 ) return -1;
	CubeSimplices cs( dim );
	int k = 0;
	for( const std::vector< int > &s : cs.simplices ) for( int v : s ) out[k++] = v;
	return (int)cs.simplices.size();
}

int orc_subgrid_triangulation( int dim , size_t nCells , const unsigned *cells , int orient ,
                               size_t *nV , double **V , size_t *nS , unsigned **S )
{
	if( dim<2 || dim>4 ) return -1;
	CubeSimplices cs( dim );
	std::vector< std::vector< int > > simp = cs.simplices;
	if( orient ) for( std::vector< int > &s : simp )
	{
		std::vector< std::vector< long long > > m( dim , std::vector< long long >( dim ) );
		for( int j=0 ; j<dim ; j++ ) for( int k=0 ; k<dim ; k++ ) m[j][k] = (long long)( ( s[j+1]>>k ) & 1 ) - (long long)( ( s[0]>>k ) & 1 );
		if( IntDet( m )<0 ) std::swap( s[0] , s[1] );
	}
	std::map< std::vector< unsigned > , unsigned > vMap;
	std::vector< double > verts;
	std::vector< unsigned > sims;
	for( size_t e=0 ; e<nCells ; e++ ) for( const std::vector< int > &s : simp ) for( int j=0 ; j<=dim ; j++ )
	{
		std::vector< unsigned > p( dim );
		for( int d=0 ; d<dim ; d++ ) p[d] = cells[ e*dim+d ] + (unsigned)( ( s[j]>>d ) & 1 );
		auto iter = vMap.find( p );
		if( iter==vMap.end() )
		{
			const unsigned id = (unsigned)vMap.size();
			vMap[p] = id;
			for( int d=0 ; d<dim ; d++ ) verts.push_back( (double)p[d] );
			sims.push_back( id );
		}
		else sims.push_back( iter->second );
	}
	*nV = verts.size() / dim , *V = Copy( verts );
	*nS = sims.size() / ( dim+1 ) , *S = Copy( sims );
	return 0;
}

int orc_level_set( int dim , int codim , int vdim , size_t nV , const double *vdata , size_t nS , const unsigned *S ,
                   const double *values , const double *iso , size_t *nLV , double **LV , size_t *nLS , unsigned **LS )
{
	if( dim<2 || dim>4 || codim<1 || codim>2 || codim>=dim ) return -1;
	(void)nV;
	LevelSetState st;
	st.codim = codim , st.vdim = vdim , st.data = vdata , st.values = values , st.iso = iso;
	std::vector< Simp > out;
	for( size_t i=0 ; i<nS ; i++ )
	{
		Simp si( S + i*( dim+1 ) , S + ( i+1 )*( dim+1 ) );
		AddLevelSet( out , si , st );
	}
	std::vector< unsigned > flat;
	for( const Simp &s : out ) flat.insert( flat.end() , s.begin() , s.end() );
	*nLV = st.outData.size() / vdim , *LV = Copy( st.outData );
	*nLS = out.size() , *LS = Copy( flat );
	return 0;
}

void orc_free( void *p ){ free( p ); }

}
