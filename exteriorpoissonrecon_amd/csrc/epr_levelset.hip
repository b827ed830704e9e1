// Simplicial extraction on the device: the triangulation of a list of grid
// cells (MarchingSimplices::RegularSubGridTriangulation,
// Misha/MarchingSimplices.h:622-706) and the codimension-2 zero set of two
// values per vertex over a simplicial mesh (MarchingSimplices::LevelSet,
// Misha/MarchingSimplices.h:410-576), with the reference's vertex numbering
// and simplex order.
//
// The reference is sequential: output vertices are numbered in the order in
// which a scan over (cell, simplex, vertex) -- or over (simplex, face) --
// first meets them, through std::map / std::unordered_map lookups.  Here every
// occurrence is evaluated in parallel, the first occurrence of each key is
// found with atomicMin (dense grid keys) or a stable radix sort (index-tuple
// keys), and an exclusive scan over "creates a vertex" flags in occurrence
// order yields the reference's numbers.
//
// Level-set semantics per triangle occurrence o (in the reference's scan
// order), with P(o) its sorted vertex triple and R(o) the triple with the
// snapped barycentric corners replaced by 0xffffffff:
//   creatorP(P) = first occurrence with key P that is interior and not snapped
//   o uses creatorP's vertex when creatorP(P(o)) < o;
//   otherwise, if o is interior and snapped, it uses the first such
//   occurrence with the same R (creating the vertex if it is that one);
//   otherwise, if o is interior, o creates the vertex for P(o).
// Within one parent the reference skips a vertex already in the parent's list
// (InList, MarchingSimplices.h:434-439); the list is then fanned from its first
// vertex (:557-572).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/epr.h"
#include "epr_dev.h"
#include "epr_levelset_dev.h"

namespace epr
{
namespace ls
{
constexpr unsigned NONE = 0xffffffffu;

// ---------------------------------------------------------------- cube simplices
// Vertex bits: coordinate d = bit d.  The Dim-cube is coned from the origin
// over its triangulated far faces {x_dir = 1}, dir = 0..Dim-1, the origin last
// (Misha/HyperCubeSimplices.h:126-134).
static std::vector< std::vector< int > > CubeSimplices( int dim )
{
	if( dim==1 ) return { { 0 , 1 } };
	std::vector< std::vector< int > > lower = CubeSimplices( dim-1 ) , out;
	for( int dir=0 ; dir<dim ; dir++ ) for( const std::vector< int > &ls : lower )
	{
		std::vector< int > s;
		for( int lv : ls )
		{
			int v = 0;
			for( int c=0 , lc=0 ; c<dim ; c++ ) v |= ( c==dir ? 1 : ( lv>>lc++ ) & 1 ) << c;
			s.push_back( v );
		}
		s.push_back( 0 );
		out.push_back( s );
	}
	return out;
}

static long long IntDet( const std::vector< std::vector< long long > > &m )
{
	const int n = (int)m.size();
	if( n==1 ) return m[0][0];
	long long det = 0;
	for( int c=0 ; c<n ; c++ )
	{
		std::vector< std::vector< long long > > sub;
		for( int r=1 ; r<n ; r++ ){ std::vector< long long > row ; for( int cc=0 ; cc<n ; cc++ ) if( cc!=c ) row.push_back( m[r][cc] ) ; sub.push_back( row ); }
		det += ( c&1 ? -1 : 1 ) * m[0][c] * IntDet( sub );
	}
	return det;
}

struct CubeTab { unsigned char v[24][5]; int n; };

static CubeTab MakeCubeTab( int dim , bool orient )
{
	CubeTab t;
	std::vector< std::vector< int > > simp = CubeSimplices( dim );
	t.n = (int)simp.size();
	for( int i=0 ; i<t.n ; i++ )
	{
		std::vector< int > &s = simp[i];
		if( orient )
		{
			// MarchingSimplices.h:658-664: swap the first two vertices of negatively oriented simplices
			std::vector< std::vector< long long > > m( dim , std::vector< long long >( dim ) );
			for( int j=0 ; j<dim ; j++ ) for( int k=0 ; k<dim ; k++ ) m[j][k] = (long long)( ( s[j+1]>>k ) & 1 ) - (long long)( ( s[0]>>k ) & 1 );
			if( IntDet( m )<0 ) std::swap( s[0] , s[1] );
		}
		for( int j=0 ; j<=dim ; j++ ) t.v[i][j] = (unsigned char)s[j];
	}
	return t;
}

struct KeyGrid { long long stride[4]; };

__device__ __forceinline__ long long PointKey( const unsigned *cell , unsigned bits , int dim , const KeyGrid &kg , unsigned *p )
{
	long long key = 0;
	for( int d=0 ; d<dim ; d++ ){ p[d] = cell[d] + ( ( bits>>d ) & 1 ) ; key += (long long)p[d] * kg.stride[d]; }
	return key;
}

__global__ void __launch_bounds__( 256 ) KTriFirst( const unsigned *cells , long long nOcc , int dim , CubeTab tab , KeyGrid kg , unsigned *first )
{
	const long long o = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( o>=nOcc ) return;
	const int per = tab.n*( dim+1 );
	const long long e = o / per;
	const int r = (int)( o % per );
	unsigned p[4];
	const long long key = PointKey( cells + e*dim , tab.v[ r/( dim+1 ) ][ r%( dim+1 ) ] , dim , kg , p );
	atomicMin( first + key , (unsigned)o );
}

__global__ void __launch_bounds__( 256 ) KTriFlag( const unsigned *cells , long long nOcc , int dim , CubeTab tab , KeyGrid kg , const unsigned *first , unsigned *flag )
{
	const long long o = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( o>=nOcc ) return;
	const int per = tab.n*( dim+1 );
	const long long e = o / per;
	const int r = (int)( o % per );
	unsigned p[4];
	const long long key = PointKey( cells + e*dim , tab.v[ r/( dim+1 ) ][ r%( dim+1 ) ] , dim , kg , p );
	flag[o] = first[key]==(unsigned)o ? 1u : 0u;
}

// first occurrences write their vertex and publish its number under the key
__global__ void __launch_bounds__( 256 ) KTriVertex( const unsigned *cells , long long nOcc , int dim , CubeTab tab , KeyGrid kg , const unsigned *flag , const unsigned *vid , unsigned *first , double *V )
{
	const long long o = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( o>=nOcc || !flag[o] ) return;
	const int per = tab.n*( dim+1 );
	const long long e = o / per;
	const int r = (int)( o % per );
	unsigned p[4];
	const long long key = PointKey( cells + e*dim , tab.v[ r/( dim+1 ) ][ r%( dim+1 ) ] , dim , kg , p );
	const unsigned v = vid[o];
	for( int d=0 ; d<dim ; d++ ) V[ (size_t)v*dim+d ] = (double)p[d];
	first[key] = v;
}

__global__ void __launch_bounds__( 256 ) KTriSimplex( const unsigned *cells , long long nOcc , int dim , CubeTab tab , KeyGrid kg , const unsigned *vertexOf , unsigned *S )
{
	const long long o = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( o>=nOcc ) return;
	const int per = tab.n*( dim+1 );
	const long long e = o / per;
	const int r = (int)( o % per );
	unsigned p[4];
	const long long key = PointKey( cells + e*dim , tab.v[ r/( dim+1 ) ][ r%( dim+1 ) ] , dim , kg , p );
	S[o] = vertexOf[key];
}

__global__ void __launch_bounds__( 256 ) KFillU32( unsigned *p , long long n , unsigned v )
{
	for( long long i=(long long)blockIdx.x*blockDim.x + threadIdx.x ; i<n ; i+=(long long)gridDim.x*blockDim.x ) p[i] = v;
}

// ---------------------------------------------------------------- level set (codim 2)
#pragma clang fp contract(off)

// Eigen 3.3.7 FullPivHouseholderQR<Matrix2d> compute + solve
// (Eigen/src/QR/FullPivHouseholderQR.h:456-560, Householder.h:62-127) for a
// 2x2 system, operation for operation.
__device__ static void SolveQR2( const double M[2][2] , const double b[2] , double x[2] )
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
		double best = fabs( m[k][k] );
		for( int j=k ; j<2 ; j++ ) for( int i=k ; i<2 ; i++ )
		{
			if( i==k && j==k ) continue;
			const double v = fabs( m[i][j] );
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
		if( k!=bi ) for( int c=k ; c<2 ; c++ ){ const double t = m[k][c] ; m[k][c] = m[bi][c] ; m[bi][c] = t; }
		if( k!=bj ) for( int r=0 ; r<2 ; r++ ){ const double t = m[r][k] ; m[r][k] = m[r][bj] ; m[r][bj] = t; }
		double tau , beta;
		if( k==1 ) tau = 0 , beta = m[1][1];
		else
		{
			const double tailSq = m[1][0]*m[1][0] , c0 = m[0][0];
			if( tailSq<=DBL_MIN ) tau = 0 , beta = c0 , m[1][0] = 0;
			else
			{
				beta = sqrt( c0*c0 + tailSq );
				if( c0>=0 ) beta = -beta;
				m[1][0] = m[1][0] / ( c0 - beta );
				tau = ( beta - c0 ) / beta;
			}
		}
		m[k][k] = beta , hc[k] = tau;
		if( fabs( beta )>maxpivot ) maxpivot = fabs( beta );
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
	for( int k=0 ; k<2 ; k++ ){ const int t = perm[k] ; perm[k] = perm[ colT[k] ] ; perm[ colT[k] ] = t; }
	int rank = 0;
	for( int i=0 ; i<nonzero ; i++ ) rank += fabs( m[i][i] ) > maxpivot*prec;
	if( rank==0 ){ x[0] = x[1] = 0 ; return; }
	double c[2] = { b[0] , b[1] };
	for( int k=0 ; k<rank ; k++ )
	{
		{ const double t = c[k] ; c[k] = c[ rowT[k] ] ; c[ rowT[k] ] = t; }
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

// bounding-sphere rejection of a K-simplex in value space (MarchingSimplices.h:424-431)
template< int K >
__device__ __forceinline__ bool Reaches( const double (*val)[2] , const int *idx , double iso0 , double iso1 )
{
	double c0 = 0 , c1 = 0;
	for( int d=0 ; d<=K ; d++ ) c0 += val[ idx[d] ][0] , c1 += val[ idx[d] ][1];
	const double inv = 1. / (double)( K+1 );
	c0 *= inv , c1 *= inv;
	double r2 = 0;
	for( int d=0 ; d<=K ; d++ )
	{
		const double t0 = -val[ idx[d] ][0] + c0 , t1 = -val[ idx[d] ][1] + c1;
		double n2 = 0;
		n2 += t0*t0;
		n2 += t1*t1;
		r2 = fmax( r2 , n2 );
	}
	const double t0 = -iso0 + c0 , t1 = -iso1 + c1;
	double n2 = 0;
	n2 += t0*t0;
	n2 += t1*t1;
	return !( n2>r2 );
}

// face v of a K-simplex given by local positions, with the reference's orientation (MarchingSimplices.h:82-90)
template< int K >
__device__ __forceinline__ void Face( const int *si , int v , int *f )
{
	for( int d=0 ; d<K ; d++ ) f[d] = si[ ( v+d+1 )%( K+1 ) ];
	if( ( K&1 ) && ( v&1 ) ){ const int t = f[0] ; f[0] = f[1] ; f[1] = t; }
}

struct LsMesh
{
	const unsigned *S;           // nS x (Dim+1)
	const double *values;        // nV x 2
	const double *data;          // nV x vdim
	int vdim;
	long long nS;
	double iso0 , iso1;
};

struct LsRecords
{
	unsigned *pk0 , *pk1 , *pk2;   // sorted vertex triple
	unsigned *rk0 , *rk1 , *rk2;   // sorted triple with snapped corners = NONE
	unsigned *flags;               // bit 0 interior, bit 1 snapped
	unsigned *parent;              // simplex
	unsigned char *local;          // g*4 + f (4D), f (3D)
	double *data;                  // vdim per record
};

__device__ __forceinline__ void Sort3( unsigned &a , unsigned &b , unsigned &c )
{
	unsigned t;
	if( a>b ) t = a , a = b , b = t;
	if( b>c ) t = b , b = c , c = t;
	if( a>b ) t = a , a = b , b = t;
}

// The reached triangle occurrences of simplex s in the reference's order.
// WRITE=false counts them, WRITE=true writes their records from `at` on.
template< int Dim , bool WRITE >
__device__ int SimplexOccurrences( const LsMesh &M , long long s , LsRecords R , unsigned at )
{
	unsigned gi[Dim+1];
	double val[Dim+1][2];
	for( int d=0 ; d<=Dim ; d++ )
	{
		gi[d] = M.S[ s*( Dim+1 ) + d ];
		val[d][0] = M.values[ 2*(size_t)gi[d] ] , val[d][1] = M.values[ 2*(size_t)gi[d]+1 ];
	}
	int top[Dim+1];
	for( int d=0 ; d<=Dim ; d++ ) top[d] = d;
	if( !Reaches< Dim >( val , top , M.iso0 , M.iso1 ) ) return 0;
	int count = 0;
	constexpr int NT = Dim==4 ? 5 : 1;
	for( int g=0 ; g<NT ; g++ )
	{
		int tet[4];
		if constexpr( Dim==4 ){ Face< 4 >( top , g , tet ) ; if( !Reaches< 3 >( val , tet , M.iso0 , M.iso1 ) ) continue; }
		else for( int d=0 ; d<4 ; d++ ) tet[d] = top[d];
		for( int f=0 ; f<4 ; f++ )
		{
			int tri[3];
			Face< 3 >( tet , f , tri );
			if( !Reaches< 2 >( val , tri , M.iso0 , M.iso1 ) ) continue;
			if constexpr( WRITE )
			{
				const unsigned r = at + count;
				unsigned p0 = gi[ tri[0] ] , p1 = gi[ tri[1] ] , p2 = gi[ tri[2] ];
				// barycentric coordinates of the iso-value (MarchingSimplices.h:447-475)
				double Mx[2][2] , rhs[2] , a_[2];
				for( int c=0 ; c<2 ; c++ ) for( int rr=0 ; rr<2 ; rr++ ) Mx[rr][c] = val[ tri[c+1] ][rr] - val[ tri[0] ][rr];
				rhs[0] = M.iso0 - val[ tri[0] ][0] , rhs[1] = M.iso1 - val[ tri[0] ][1];
				SolveQR2( Mx , rhs , a_ );
				double a[3];
				a[0] = 1.;
				a[0] -= a_[0] , a[1] = a_[0];
				a[0] -= a_[1] , a[2] = a_[1];
				bool interior = true , snapped = false;
				unsigned q[3] = { p0 , p1 , p2 };
				for( int i=0 ; i<3 ; i++ )
				{
					interior &= a[i]>=0 && a[i]<=1.;
					if( fabs( a[i] )<1e-15 ) a[i] = 0 , q[i] = NONE , snapped = true;
				}
				Sort3( p0 , p1 , p2 );
				Sort3( q[0] , q[1] , q[2] );
				R.pk0[r] = p0 , R.pk1[r] = p1 , R.pk2[r] = p2;
				R.rk0[r] = q[0] , R.rk1[r] = q[1] , R.rk2[r] = q[2];
				R.flags[r] = ( interior ? 1u : 0u ) | ( snapped ? 2u : 0u );
				R.parent[r] = (unsigned)s;
				R.local[r] = (unsigned char)( g*4 + f );
				for( int k=0 ; k<M.vdim ; k++ )
				{
					double d = M.data[ (size_t)gi[ tri[0] ]*M.vdim+k ] * a[0];
					d += M.data[ (size_t)gi[ tri[1] ]*M.vdim+k ] * a[1];
					d += M.data[ (size_t)gi[ tri[2] ]*M.vdim+k ] * a[2];
					R.data[ (size_t)r*M.vdim+k ] = d;
				}
			}
			count++;
		}
	}
	return count;
}
#pragma clang fp contract(on)

template< int Dim >
__global__ void __launch_bounds__( 256 ) KLsCount( LsMesh M , unsigned *count )
{
	const long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=M.nS ) return;
	LsRecords none{};
	count[s] = (unsigned)SimplexOccurrences< Dim , false >( M , s , none , 0 );
}

template< int Dim >
__global__ void __launch_bounds__( 256 ) KLsRecords( LsMesh M , const unsigned *offset , LsRecords R )
{
	const long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=M.nS ) return;
	SimplexOccurrences< Dim , true >( M , s , R , offset[s] );
}

__global__ void __launch_bounds__( 256 ) KIota( unsigned *p , unsigned n ){ const unsigned i = blockIdx.x*blockDim.x + threadIdx.x ; if( i<n ) p[i] = i; }

// 64-bit sort keys (k0, k1) of the records in `order`
__global__ void __launch_bounds__( 256 ) KHiKeys( const unsigned *k0 , const unsigned *k1 , const unsigned *order , unsigned n , unsigned long long *out )
{
	const unsigned i = blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=n ) return;
	const unsigned r = order[i];
	out[i] = ( (unsigned long long)k0[r]<<32 ) | k1[r];
}

// keys of the snapped lookup: ineligible records sort last under (NONE, NONE, NONE)
__global__ void __launch_bounds__( 256 ) KEligibleKeys( LsRecords R , const unsigned *creatorP , unsigned n , unsigned *e0 , unsigned *e1 , unsigned *e2 )
{
	const unsigned r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=n ) return;
	const bool eligible = R.flags[r]==3u && !( creatorP[r]!=NONE && creatorP[r]<r );
	e0[r] = eligible ? R.rk0[r] : NONE , e1[r] = eligible ? R.rk1[r] : NONE , e2[r] = eligible ? R.rk2[r] : NONE;
}

// For each run of equal keys in sorted order, the first record (in record
// order, kept by the stable sort) with (flags & mask)==want; written to every
// record of the run.
__global__ void __launch_bounds__( 256 ) KFirstInRun( const unsigned *k0 , const unsigned *k1 , const unsigned *k2 , const unsigned *order , unsigned n ,
                                                      const unsigned *flags , unsigned mask , unsigned want , const unsigned *skip , unsigned *first )
{
	const unsigned j = blockIdx.x*blockDim.x + threadIdx.x;
	if( j>=n ) return;
	const unsigned r = order[j];
	auto same = [&]( unsigned a , unsigned b ){ return k0[a]==k0[b] && k1[a]==k1[b] && k2[a]==k2[b]; };
	if( j>0 && same( order[j-1] , r ) ) return;
	if( k0[r]==NONE ) return;              // the run of ineligible records keeps NONE
	unsigned found = NONE;
	for( unsigned t=j ; t<n && same( order[t] , r ) ; t++ )
	{
		const unsigned q = order[t];
		const bool ok = ( flags[q] & mask )==want && !( skip && ( skip[q]!=NONE && skip[q]<q ) );
		if( ok && found==NONE ) found = q;
	}
	for( unsigned t=j ; t<n && same( order[t] , r ) ; t++ ) first[ order[t] ] = found;
}

__global__ void __launch_bounds__( 256 ) KCreatorFlags( const unsigned *flags , const unsigned *creatorP , const unsigned *creatorR , unsigned n , unsigned *isCreator )
{
	const unsigned r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=n ) return;
	const bool plain = flags[r]==1u && creatorP[r]==r;
	const bool eligible = flags[r]==3u && !( creatorP[r]!=NONE && creatorP[r]<r );
	isCreator[r] = ( plain || ( eligible && creatorR[r]==r ) ) ? 1u : 0u;
}

__global__ void __launch_bounds__( 256 ) KResolve( const unsigned *flags , const unsigned *creatorP , const unsigned *creatorR , const unsigned *isCreator , const unsigned *vid ,
                                                   const double *recData , int vdim , unsigned n , unsigned *vertexOf , double *outV )
{
	const unsigned r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=n ) return;
	if( isCreator[r] ) for( int k=0 ; k<vdim ; k++ ) outV[ (size_t)vid[r]*vdim+k ] = recData[ (size_t)r*vdim+k ];
	unsigned v = NONE;
	if( creatorP[r]!=NONE && creatorP[r]<r ) v = vid[ creatorP[r] ];
	else if( flags[r]==1u ) v = vid[r];
	else if( flags[r]==3u && creatorR[r]!=NONE ) v = vid[ creatorR[r] ];
	vertexOf[r] = v;
}

// Output simplices of simplex s from its records [b, e): per tetrahedron the
// vertex list without repeats, fanned into segments from its first vertex;
// in 4D the segments of the five tetrahedra fanned into triangles.
template< int Dim , bool WRITE >
__device__ int AssembleSimplex( const unsigned *vertexOf , const unsigned char *local , unsigned b , unsigned e , unsigned *out )
{
	unsigned seg[15][2];
	int nSeg = 0;
	unsigned r = b;
	constexpr int NT = Dim==4 ? 5 : 1;
	for( int g=0 ; g<NT ; g++ )
	{
		unsigned L[4];
		int nL = 0;
		for( ; r<e && ( Dim==3 || local[r]/4==g ) ; r++ )
		{
			const unsigned v = vertexOf[r];
			if( v==NONE ) continue;
			bool in = false;
			for( int i=0 ; i<nL ; i++ ) in |= L[i]==v;
			if( !in ) L[ nL++ ] = v;
		}
		for( int i=1 ; i<nL ; i++ ) seg[nSeg][0] = L[0] , seg[nSeg][1] = L[i] , nSeg++;
	}
	if constexpr( Dim==3 )
	{
		if( WRITE ) for( int i=0 ; i<nSeg ; i++ ) out[2*i] = seg[i][0] , out[2*i+1] = seg[i][1];
		return nSeg;
	}
	else
	{
		if( !nSeg ) return 0;
		const unsigned v = seg[0][0];
		int n = 0;
		for( int i=1 ; i<nSeg ; i++ ) if( seg[i][0]!=v && seg[i][1]!=v )
		{
			if( WRITE ) out[3*n] = v , out[3*n+1] = seg[i][0] , out[3*n+2] = seg[i][1];
			n++;
		}
		return n;
	}
}

template< int Dim , bool WRITE >
__global__ void __launch_bounds__( 256 ) KAssemble( const unsigned *recOffset , long long nS , const unsigned *vertexOf , const unsigned char *local , unsigned *count , const unsigned *outOffset , unsigned *out )
{
	const long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=nS ) return;
	const unsigned b = recOffset[s] , e = recOffset[s+1];
	if( WRITE ) AssembleSimplex< Dim , true >( vertexOf , local , b , e , out + (size_t)outOffset[s]*( Dim-1 ) );
	else count[s] = b==e ? 0u : (unsigned)AssembleSimplex< Dim , false >( vertexOf , local , b , e , nullptr );
}

// ---------------------------------------------------------------- level set (codim 1)
// One value per vertex.  A triangle (the reference's CoDim==1, Dim==2 branch,
// MarchingSimplices.h:497-541) with a vertex whose sign differs from both
// others yields an oriented segment between the iso points of its two edges
// at that vertex; the iso point of an edge is created by the first edge
// occurrence that asks for it (key = sorted vertex pair), with the parameter
// s = -c_ii / (c_o - c_ii) of that occurrence.  Higher simplices fan their
// faces' simplices from the first vertex (:555-573).
#pragma clang fp contract(off)
template< int K >
__device__ __forceinline__ bool Reaches1( const double *val , const int *idx , double iso )
{
	double c = 0;
	for( int d=0 ; d<=K ; d++ ) c += val[ idx[d] ];
	c *= 1. / (double)( K+1 );
	double r2 = 0;
	for( int d=0 ; d<=K ; d++ ){ const double t = -val[ idx[d] ] + c ; double n2 = 0 ; n2 += t*t ; r2 = fmax( r2 , n2 ); }
	const double t = -iso + c;
	double n2 = 0;
	n2 += t*t;
	return !( n2>r2 );
}

struct Ls1Records
{
	unsigned long long *key;     // sorted vertex pair of the edge
	unsigned char *local;        // ( triangle index )*2 + edge (0, 1)
	unsigned char *swap;         // the triangle's segment is reversed
	double *data;                // vdim per record
};

template< int Dim , bool WRITE >
__device__ int SimplexEdges( const LsMesh &M , long long s , Ls1Records R , unsigned at )
{
	unsigned gi[Dim+1];
	double val[Dim+1];
	for( int d=0 ; d<=Dim ; d++ ){ gi[d] = M.S[ s*( Dim+1 ) + d ] ; val[d] = M.values[ gi[d] ]; }
	int top[Dim+1];
	for( int d=0 ; d<=Dim ; d++ ) top[d] = d;
	if( !Reaches1< Dim >( val , top , M.iso0 ) ) return 0;
	int count = 0;
	constexpr int NT = Dim==4 ? 5 : 1;
	for( int g=0 ; g<NT ; g++ )
	{
		int tet[4];
		if constexpr( Dim==4 ){ Face< 4 >( top , g , tet ) ; if( !Reaches1< 3 >( val , tet , M.iso0 ) ) continue; }
		else for( int d=0 ; d<4 ; d++ ) tet[d] = top[d];
		for( int f=0 ; f<4 ; f++ )
		{
			int tri[3];
			Face< 3 >( tet , f , tri );
			if( !Reaches1< 2 >( val , tri , M.iso0 ) ) continue;
			double cv[3];
			for( int i=0 ; i<3 ; i++ ){ cv[i] = val[ tri[i] ] - M.iso0 ; if( cv[i]==0 ) cv[i] = 1e-10; }
			int ii = -1;
			for( int d=0 ; d<3 ; d++ ) if( ii<0 && cv[d]*cv[(d+1)%3]<0 && cv[d]*cv[(d+2)%3]<0 ) ii = d;
			if( ii<0 ) continue;
			if constexpr( WRITE ) for( int i=0 ; i<2 ; i++ )
			{
				const unsigned r = at + count + i;
				const int o = ( ii+1+i )%3;
				const unsigned a = gi[ tri[ii] ] , b = gi[ tri[o] ];
				R.key[r] = a<b ? ( (unsigned long long)a<<32 ) | b : ( (unsigned long long)b<<32 ) | a;
				R.local[r] = (unsigned char)( ( g*4 + f )*2 + i );
				R.swap[r] = cv[ii]<0 ? 1 : 0;
				const double sp = - cv[ii] / ( cv[o] - cv[ii] );
				for( int k=0 ; k<M.vdim ; k++ ) R.data[ (size_t)r*M.vdim+k ] = M.data[ (size_t)a*M.vdim+k ]*( 1.-sp ) + M.data[ (size_t)b*M.vdim+k ]*sp;
			}
			count += 2;
		}
	}
	return count;
}
#pragma clang fp contract(on)

template< int Dim >
__global__ void __launch_bounds__( 256 ) KLs1Count( LsMesh M , unsigned *count )
{
	const long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=M.nS ) return;
	Ls1Records none{};
	count[s] = (unsigned)SimplexEdges< Dim , false >( M , s , none , 0 );
}

template< int Dim >
__global__ void __launch_bounds__( 256 ) KLs1Records( LsMesh M , const unsigned *offset , Ls1Records R )
{
	const long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=M.nS ) return;
	SimplexEdges< Dim , true >( M , s , R , offset[s] );
}

// the first record of every run of equal keys creates the vertex
__global__ void __launch_bounds__( 256 ) KFirstOfKey( const unsigned long long *skey , const unsigned *order , unsigned n , unsigned *first )
{
	const unsigned j = blockIdx.x*blockDim.x + threadIdx.x;
	if( j>=n ) return;
	if( j>0 && skey[j]==skey[j-1] ) return;
	const unsigned f = order[j];
	for( unsigned t=j ; t<n && skey[t]==skey[j] ; t++ ) first[ order[t] ] = f;
}

__global__ void __launch_bounds__( 256 ) KLs1Vertices( const unsigned *first , const unsigned *vid , const double *recData , int vdim , unsigned n , unsigned *vertexOf , double *outV )
{
	const unsigned r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=n ) return;
	if( first[r]==r ) for( int k=0 ; k<vdim ; k++ ) outV[ (size_t)vid[r]*vdim+k ] = recData[ (size_t)r*vdim+k ];
	vertexOf[r] = vid[ first[r] ];
}

__global__ void __launch_bounds__( 256 ) KIsFirst( const unsigned *first , unsigned n , unsigned *flag )
{
	const unsigned r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r<n ) flag[r] = first[r]==r ? 1u : 0u;
}

// Output simplices of simplex s: per triangle its oriented segment, per
// tetrahedron the fan of its segments into triangles; in 4D the fan of the
// five tetrahedra's triangles into tetrahedra.
template< int Dim , bool WRITE >
__device__ int AssembleSimplex1( const unsigned *vertexOf , const unsigned char *local , const unsigned char *swp , unsigned b , unsigned e , unsigned *out )
{
	unsigned tris[15][3];
	int nTri = 0;
	unsigned r = b;
	constexpr int NT = Dim==4 ? 5 : 1;
	for( int g=0 ; g<NT ; g++ )
	{
		unsigned seg[4][2];
		int nSeg = 0;
		for( ; r<e && ( Dim==3 || ( local[r]>>1 )/4==g ) ; r+=2 )
		{
			unsigned v0 = vertexOf[r] , v1 = vertexOf[r+1];
			if( swp[r] ){ const unsigned t = v0 ; v0 = v1 ; v1 = t; }
			seg[nSeg][0] = v0 , seg[nSeg][1] = v1 , nSeg++;
		}
		if( !nSeg ) continue;
		const unsigned v = seg[0][0];
		for( int i=1 ; i<nSeg ; i++ ) if( seg[i][0]!=v && seg[i][1]!=v ){ tris[nTri][0] = v , tris[nTri][1] = seg[i][0] , tris[nTri][2] = seg[i][1] ; nTri++; }
	}
	if constexpr( Dim==3 )
	{
		if( WRITE ) for( int i=0 ; i<nTri ; i++ ) for( int k=0 ; k<3 ; k++ ) out[3*i+k] = tris[i][k];
		return nTri;
	}
	else
	{
		if( !nTri ) return 0;
		const unsigned v = tris[0][0];
		int n = 0;
		for( int i=1 ; i<nTri ; i++ ) if( tris[i][0]!=v && tris[i][1]!=v && tris[i][2]!=v )
		{
			if( WRITE ){ out[4*n] = v ; for( int k=0 ; k<3 ; k++ ) out[4*n+1+k] = tris[i][k]; }
			n++;
		}
		return n;
	}
}

template< int Dim , bool WRITE >
__global__ void __launch_bounds__( 256 ) KAssemble1( const unsigned *recOffset , long long nS , const unsigned *vertexOf , const unsigned char *local , const unsigned char *swp ,
                                                     unsigned *count , const unsigned *outOffset , unsigned *out )
{
	const long long s = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=nS ) return;
	const unsigned b = recOffset[s] , e = recOffset[s+1];
	if( WRITE ) AssembleSimplex1< Dim , true >( vertexOf , local , swp , b , e , out + (size_t)outOffset[s]*Dim );
	else count[s] = b==e ? 0u : (unsigned)AssembleSimplex1< Dim , false >( vertexOf , local , swp , b , e , nullptr );
}

static inline unsigned int Grid( long long n ){ return (unsigned int)std::max< long long >( 1 , ( n + 255 )/256 ); }

// exclusive prefix sum of n values into out[0..n] (out[n] = total); returns the total
static unsigned ExclusiveScan( const unsigned *in , unsigned *out , long long n , DevBuf< unsigned char > &tmp , hipStream_t st )
{
	size_t bytes = 0;
	HIPCHECK( hipcub::DeviceScan::ExclusiveSum( nullptr , bytes , in , out , (int)n + 1 , st ) );
	tmp.ensure( bytes );
	HIPCHECK( hipcub::DeviceScan::ExclusiveSum( tmp.p , bytes , in , out , (int)n + 1 , st ) );
	unsigned total = 0;
	HIPCHECK( hipMemcpyAsync( &total , out + n , sizeof(unsigned) , hipMemcpyDeviceToHost , st ) );
	HIPCHECK( hipStreamSynchronize( st ) );
	return total;
}

// record order sorted stably by the triple (k0, k1, k2): least significant key first
static void SortTriples( const unsigned *k0 , const unsigned *k1 , const unsigned *k2 , unsigned n , DevBuf< unsigned > &order ,
                         DevBuf< unsigned > &t0 , DevBuf< unsigned > &t1 , DevBuf< unsigned long long > &h0 , DevBuf< unsigned long long > &h1 ,
                         DevBuf< unsigned char > &tmp , hipStream_t st )
{
	order.ensure( n ) , t0.ensure( n ) , t1.ensure( n ) , h0.ensure( n ) , h1.ensure( n );
	KIota<<< Grid( n ) , 256 , 0 , st >>>( t0.p , n );
	size_t bytes = 0;
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( nullptr , bytes , k2 , t1.p , t0.p , order.p , (int)n , 0 , 32 , st ) );
	tmp.ensure( bytes );
	// pass 1: by k2 (keys sorted into t1, record ids into order)
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( tmp.p , bytes , k2 , t1.p , t0.p , order.p , (int)n , 0 , 32 , st ) );
	// pass 2: stably by (k0, k1)
	KHiKeys<<< Grid( n ) , 256 , 0 , st >>>( k0 , k1 , order.p , n , h0.p );
	size_t bytes2 = 0;
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( nullptr , bytes2 , h0.p , h1.p , order.p , t0.p , (int)n , 0 , 64 , st ) );
	tmp.ensure( bytes2 );
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( tmp.p , bytes2 , h0.p , h1.p , order.p , t0.p , (int)n , 0 , 64 , st ) );
	std::swap( order , t0 );
}

void TriangulateDev( int dim , const unsigned *dCells , size_t nCells , const unsigned ext[4] , bool orient , hipStream_t st , DevMesh &out )
{
	CubeTab tab = MakeCubeTab( dim , orient );
	const long long nOcc = (long long)nCells * tab.n * ( dim+1 );
	if( nOcc>=(long long)NONE ) throw std::string( "too many cells for 32-bit occurrence indices" );
	KeyGrid kg{};
	unsigned long long keys = 1;
	for( int d=dim-1 ; d>=0 ; d-- ){ kg.stride[d] = (long long)keys ; keys *= (unsigned long long)ext[d] + 2; }
	if( keys>( 1ull<<32 ) ) throw std::string( "cell coordinates span too large a box" );
	out.vdim = dim , out.k = dim+1 , out.nV = out.nS = 0;
	if( !nCells ) return;
	DevBuf< unsigned > first , flag , vid;
	DevBuf< unsigned char > tmp;
	first.alloc( keys ) , flag.alloc( nOcc+1 ) , vid.alloc( nOcc+1 ) , out.S.alloc( nOcc );
	KFillU32<<< 4096 , 256 , 0 , st >>>( first.p , (long long)keys , NONE );
	KTriFirst<<< Grid( nOcc ) , 256 , 0 , st >>>( dCells , nOcc , dim , tab , kg , first.p );
	KTriFlag<<< Grid( nOcc ) , 256 , 0 , st >>>( dCells , nOcc , dim , tab , kg , first.p , flag.p );
	KFillU32<<< 1 , 256 , 0 , st >>>( flag.p + nOcc , 1 , 0 );
	const unsigned nV = ExclusiveScan( flag.p , vid.p , nOcc , tmp , st );
	out.V.alloc( (size_t)nV*dim );
	KTriVertex<<< Grid( nOcc ) , 256 , 0 , st >>>( dCells , nOcc , dim , tab , kg , flag.p , vid.p , first.p , out.V.p );
	KTriSimplex<<< Grid( nOcc ) , 256 , 0 , st >>>( dCells , nOcc , dim , tab , kg , first.p , out.S.p );
	HIPCHECK( hipGetLastError() );
	HIPCHECK( hipStreamSynchronize( st ) );
	out.nV = nV , out.nS = (size_t)nOcc / ( dim+1 );
}

static epr_mesh *Download( const DevMesh &d , hipStream_t st )
{
	epr_mesh *m = new epr_mesh;
	m->vdim = d.vdim , m->k = d.k;
	m->v.resize( d.nV*d.vdim ) , m->s.resize( d.nS*d.k );
	d.V.download( m->v.data() , m->v.size() , st );
	d.S.download( m->s.data() , m->s.size() , st );
	HIPCHECK( hipStreamSynchronize( st ) );
	return m;
}

static epr_mesh *Triangulate( int dim , size_t nCells , const unsigned *cells , bool orient )
{
	unsigned ext[4] = { 0 , 0 , 0 , 0 };
	for( size_t e=0 ; e<nCells ; e++ ) for( int d=0 ; d<dim ; d++ ) ext[d] = std::max( ext[d] , cells[ e*dim+d ] );
	hipStream_t st = nullptr;
	DevBuf< unsigned > dCells;
	dCells.upload( cells , nCells*dim , st );
	DevMesh d;
	TriangulateDev( dim , dCells.p , nCells , ext , orient , st , d );
	return Download( d , st );
}

template< int Dim >
void LevelSet2Dev( int vdim , size_t nV , const double *dVdata , size_t nS , const unsigned *dS , const double *dValues , const double iso[2] , hipStream_t st , DevMesh &out )
{
	out.vdim = vdim , out.k = Dim-1 , out.nV = out.nS = 0;
	if( !nS ) return;
	if( (unsigned long long)nS*( Dim==4 ? 20 : 4 )>=NONE ) throw std::string( "too many simplices for 32-bit occurrence indices" );
	DevBuf< unsigned > cnt , off;
	DevBuf< unsigned char > tmp;
	LsMesh M{ dS , dValues , dVdata , vdim , (long long)nS , iso[0] , iso[1] };
	cnt.alloc( nS+1 ) , off.alloc( nS+1 );
	KLsCount< Dim ><<< Grid( nS ) , 256 , 0 , st >>>( M , cnt.p );
	KFillU32<<< 1 , 256 , 0 , st >>>( cnt.p + nS , 1 , 0 );
	const unsigned n = ExclusiveScan( cnt.p , off.p , (long long)nS , tmp , st );
	if( !n ) return;
	DevBuf< unsigned > k[6] , flags , parent , cP , cR , isC , vidB , vOf , order , t0 , t1 , e0 , e1 , e2;
	DevBuf< unsigned char > local;
	DevBuf< double > recData;
	DevBuf< unsigned long long > h0 , h1;
	for( int i=0 ; i<6 ; i++ ) k[i].alloc( n );
	flags.alloc( n ) , parent.alloc( n ) , local.alloc( n ) , recData.alloc( (size_t)n*vdim );
	LsRecords R{ k[0].p , k[1].p , k[2].p , k[3].p , k[4].p , k[5].p , flags.p , parent.p , local.p , recData.p };
	KLsRecords< Dim ><<< Grid( nS ) , 256 , 0 , st >>>( M , off.p , R );
	// first plain (interior, not snapped) occurrence of every vertex triple
	cP.alloc( n ) , cR.alloc( n );
	SortTriples( k[0].p , k[1].p , k[2].p , n , order , t0 , t1 , h0 , h1 , tmp , st );
	KFirstInRun<<< Grid( n ) , 256 , 0 , st >>>( k[0].p , k[1].p , k[2].p , order.p , n , flags.p , 3u , 1u , nullptr , cP.p );
	// first eligible snapped occurrence of every snapped triple
	e0.alloc( n ) , e1.alloc( n ) , e2.alloc( n );
	KEligibleKeys<<< Grid( n ) , 256 , 0 , st >>>( R , cP.p , n , e0.p , e1.p , e2.p );
	SortTriples( e0.p , e1.p , e2.p , n , order , t0 , t1 , h0 , h1 , tmp , st );
	KFillU32<<< Grid( n ) , 256 , 0 , st >>>( cR.p , n , NONE );
	KFirstInRun<<< Grid( n ) , 256 , 0 , st >>>( e0.p , e1.p , e2.p , order.p , n , flags.p , 3u , 3u , cP.p , cR.p );
	// creators in occurrence order -> vertex numbers
	isC.alloc( n+1 ) , vidB.alloc( n+1 ) , vOf.alloc( n );
	KCreatorFlags<<< Grid( n ) , 256 , 0 , st >>>( flags.p , cP.p , cR.p , n , isC.p );
	KFillU32<<< 1 , 256 , 0 , st >>>( isC.p + n , 1 , 0 );
	const unsigned nLV = ExclusiveScan( isC.p , vidB.p , n , tmp , st );
	out.V.alloc( (size_t)nLV*vdim );
	KResolve<<< Grid( n ) , 256 , 0 , st >>>( flags.p , cP.p , cR.p , isC.p , vidB.p , recData.p , vdim , n , vOf.p , out.V.p );
	// output simplices in simplex order
	DevBuf< unsigned > oc , oo;
	oc.alloc( nS+1 ) , oo.alloc( nS+1 );
	KAssemble< Dim , false ><<< Grid( nS ) , 256 , 0 , st >>>( off.p , (long long)nS , vOf.p , local.p , oc.p , nullptr , nullptr );
	KFillU32<<< 1 , 256 , 0 , st >>>( oc.p + nS , 1 , 0 );
	const unsigned nLS = ExclusiveScan( oc.p , oo.p , (long long)nS , tmp , st );
	out.S.alloc( (size_t)nLS*( Dim-1 ) );
	if( nLS ) KAssemble< Dim , true ><<< Grid( nS ) , 256 , 0 , st >>>( off.p , (long long)nS , vOf.p , local.p , nullptr , oo.p , out.S.p );
	HIPCHECK( hipGetLastError() );
	// the temporaries above are freed when this returns: finish their work first
	HIPCHECK( hipStreamSynchronize( st ) );
	out.nV = nLV , out.nS = nLS;
}

template< int Dim >
void LevelSet1Dev( int vdim , size_t nV , const double *dVdata , size_t nS , const unsigned *dS , const double *dValues , double iso , hipStream_t st , DevMesh &out )
{
	out.vdim = vdim , out.k = Dim , out.nV = out.nS = 0;
	if( !nS ) return;
	if( (unsigned long long)nS*( Dim==4 ? 40 : 8 )>=NONE ) throw std::string( "too many simplices for 32-bit occurrence indices" );
	DevBuf< unsigned > cnt , off;
	DevBuf< unsigned char > tmp;
	LsMesh M{ dS , dValues , dVdata , vdim , (long long)nS , iso , 0. };
	cnt.alloc( nS+1 ) , off.alloc( nS+1 );
	KLs1Count< Dim ><<< Grid( nS ) , 256 , 0 , st >>>( M , cnt.p );
	KFillU32<<< 1 , 256 , 0 , st >>>( cnt.p + nS , 1 , 0 );
	const unsigned n = ExclusiveScan( cnt.p , off.p , (long long)nS , tmp , st );
	if( !n ) return;
	DevBuf< unsigned long long > key , skey;
	DevBuf< unsigned char > local , swp;
	DevBuf< double > recData;
	DevBuf< unsigned > iota , order , first , flag , vid , vOf;
	key.alloc( n ) , skey.alloc( n ) , local.alloc( n ) , swp.alloc( n ) , recData.alloc( (size_t)n*vdim );
	Ls1Records R{ key.p , local.p , swp.p , recData.p };
	KLs1Records< Dim ><<< Grid( nS ) , 256 , 0 , st >>>( M , off.p , R );
	// the first occurrence of every edge (stable sort keeps occurrence order within a key)
	iota.alloc( n ) , order.alloc( n ) , first.alloc( n ) , flag.alloc( n+1 ) , vid.alloc( n+1 ) , vOf.alloc( n );
	KIota<<< Grid( n ) , 256 , 0 , st >>>( iota.p , n );
	size_t bytes = 0;
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( nullptr , bytes , key.p , skey.p , iota.p , order.p , (int)n , 0 , 64 , st ) );
	tmp.ensure( bytes );
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( tmp.p , bytes , key.p , skey.p , iota.p , order.p , (int)n , 0 , 64 , st ) );
	KFirstOfKey<<< Grid( n ) , 256 , 0 , st >>>( skey.p , order.p , n , first.p );
	KIsFirst<<< Grid( n ) , 256 , 0 , st >>>( first.p , n , flag.p );
	KFillU32<<< 1 , 256 , 0 , st >>>( flag.p + n , 1 , 0 );
	const unsigned nLV = ExclusiveScan( flag.p , vid.p , n , tmp , st );
	out.V.alloc( (size_t)nLV*vdim );
	KLs1Vertices<<< Grid( n ) , 256 , 0 , st >>>( first.p , vid.p , recData.p , vdim , n , vOf.p , out.V.p );
	DevBuf< unsigned > oc , oo;
	oc.alloc( nS+1 ) , oo.alloc( nS+1 );
	KAssemble1< Dim , false ><<< Grid( nS ) , 256 , 0 , st >>>( off.p , (long long)nS , vOf.p , local.p , swp.p , oc.p , nullptr , nullptr );
	KFillU32<<< 1 , 256 , 0 , st >>>( oc.p + nS , 1 , 0 );
	const unsigned nLS = ExclusiveScan( oc.p , oo.p , (long long)nS , tmp , st );
	out.S.alloc( (size_t)nLS*Dim );
	if( nLS ) KAssemble1< Dim , true ><<< Grid( nS ) , 256 , 0 , st >>>( off.p , (long long)nS , vOf.p , local.p , swp.p , nullptr , oo.p , out.S.p );
	HIPCHECK( hipGetLastError() );
	HIPCHECK( hipStreamSynchronize( st ) );
	out.nV = nLV , out.nS = nLS;
}

template void LevelSet2Dev< 3 >( int , size_t , const double * , size_t , const unsigned * , const double * , const double [2] , hipStream_t , DevMesh & );
template void LevelSet2Dev< 4 >( int , size_t , const double * , size_t , const unsigned * , const double * , const double [2] , hipStream_t , DevMesh & );
template void LevelSet1Dev< 3 >( int , size_t , const double * , size_t , const unsigned * , const double * , double , hipStream_t , DevMesh & );
template void LevelSet1Dev< 4 >( int , size_t , const double * , size_t , const unsigned * , const double * , double , hipStream_t , DevMesh & );

template< int Dim >
static epr_mesh *LevelSet( int codim , int vdim , size_t nV , const double *vdata , size_t nS , const unsigned *S , const double *values , const double *iso )
{
	for( size_t i=0 ; i<nS*( Dim+1 ) ; i++ ) if( S[i]>=nV ) throw std::string( "simplex vertex index out of range" );
	hipStream_t st = nullptr;
	DevBuf< unsigned > dS;
	DevBuf< double > dVal , dData;
	DevMesh d;
	if( nS )
	{
		dS.upload( S , nS*( Dim+1 ) , st );
		dVal.upload( values , nV*codim , st );
		dData.upload( vdata , nV*vdim , st );
	}
	if( codim==2 ) LevelSet2Dev< Dim >( vdim , nV , dData.p , nS , dS.p , dVal.p , iso , st , d );
	else           LevelSet1Dev< Dim >( vdim , nV , dData.p , nS , dS.p , dVal.p , iso[0] , st , d );
	return Download( d , st );
}
} // namespace ls
} // namespace epr

namespace epr
{
std::string &ThreadError( void ){ static thread_local std::string e ; return e; }
}
const char *EprMeshLastError( void ){ return epr::ThreadError().c_str(); }

extern "C" {

int epr_subgrid_triangulation( int dim , size_t nCells , const unsigned *cells , int orient , epr_mesh **out )
{
	if( !out || ( dim!=3 && dim!=4 ) || ( nCells && !cells ) ){ epr::ThreadError() = "bad argument" ; return EPR_ERR_ARG; }
	*out = nullptr;
	return epr::ContextFree( [&](){ *out = epr::ls::Triangulate( dim , nCells , cells , orient!=0 ); } );
}

int epr_level_set( int dim , int codim , int vdim , size_t nV , const double *vdata , size_t nS , const unsigned *simplices , const double *values , const double *iso , epr_mesh **out )
{
	if( !out || ( dim!=3 && dim!=4 ) || ( codim!=1 && codim!=2 ) || vdim<1 || vdim>16 || ( nS && ( !simplices || !values || !vdata ) ) ){ epr::ThreadError() = "bad argument" ; return EPR_ERR_ARG; }
	*out = nullptr;
	const double zero[2] = { 0 , 0 };
	if( !iso ) iso = zero;
	return epr::ContextFree( [&]()
	{
		if( dim==3 ) *out = epr::ls::LevelSet< 3 >( codim , vdim , nV , vdata , nS , simplices , values , iso );
		else         *out = epr::ls::LevelSet< 4 >( codim , vdim , nV , vdata , nS , simplices , values , iso );
	} );
}

int epr_mesh_info( const epr_mesh *m , size_t *nVertices , int *vertexDim , size_t *nSimplices , int *simplexVertices )
{
	if( !m ) return EPR_ERR_ARG;
	if( nVertices ) *nVertices = m->vdim ? m->v.size() / m->vdim : 0;
	if( vertexDim ) *vertexDim = m->vdim;
	if( nSimplices ) *nSimplices = m->k ? m->s.size() / m->k : 0;
	if( simplexVertices ) *simplexVertices = m->k;
	return EPR_OK;
}

int epr_mesh_copy( const epr_mesh *m , double *vertices , unsigned *simplices )
{
	if( !m ) return EPR_ERR_ARG;
	if( vertices && !m->v.empty() ) memcpy( vertices , m->v.data() , m->v.size()*sizeof(double) );
	if( simplices && !m->s.empty() ) memcpy( simplices , m->s.data() , m->s.size()*sizeof(unsigned) );
	return EPR_OK;
}

void epr_mesh_destroy( epr_mesh *m ){ delete m; }

}
