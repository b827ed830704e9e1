// Extraction of the reconstructed surface from a solved hierarchy on the
// device: the reference's Extract (Extract/Extract.cpp:96-330) over its
// IsoTree (Include/IsoTree.inl).
//
// Pipeline (one device-resident chain; only the orientation pass runs on the
// host, since it is a sequential traversal):
//   1. read the .tree (dense or adapted hierarchy) and form the finest regular
//      values V = x_D + P( x_{D-1} + P( ... ) ), parents summed in the
//      IsoTree's window order (IsoTree.inl:314-338)                 KRegular
//   2. zero-crossing flags of every finest element (IsoTree.inl:135-163)
//                                                                   KZeroCross
//   3. adapted trees: keep the face-connected zero-crossing components that
//      contain an indexed finest element (refineZeroCrossing, :165-222)
//                                             KUfLink / KUfRoot / KKeep
//   4. the kept elements in tree (Morton) order (levelSetNodes, :225-234),
//      their triangulation and vertex values (Extract.cpp:161-199)
//   5. the joint zero set (Extract.cpp:205), oriented (:208)  (host Orient)
//   6. per level-set vertex: the IsoTree gradient (IsoTree.inl:101-133),
//      scaled as Extract.cpp:240-257, and the voxel-to-world map (:259)
//                                                                   KFrames
// Floating-point steps follow the reference's operation order without fused
// multiply-adds, so results are bit-identical to a build of the reference
// without FMA contraction.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <bits/hash_bytes.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/epr.h"
#include "epr_dev.h"
#include "epr_levelset_dev.h"

#pragma clang fp contract(off)

namespace epr
{
namespace iso
{
constexpr unsigned long long NONE64 = 0xffffffffffffffffull;
using Clock = std::chrono::steady_clock;
static double Since( Clock::time_point t ){ return std::chrono::duration< double >( Clock::now() - t ).count(); }

// ------------------------------------------------------------------ .tree reading
struct TreeData
{
	int dim = 0 , D = 0;
	bool adapted = false;
	unsigned long long hash = 0;
	std::vector< double > xform;                          // (dim+1)^2, coords[c][r] c-major
	std::vector< std::vector< double > > x , y;           // hierarchical coefficients per depth (file order)
	// adapted: per depth, the multi-index of every function in the hierarchy's numbering
	std::vector< std::vector< unsigned > > fOff;
	std::vector< unsigned > eFinest;                      // adapted: indexed elements of depth D (dim each)
};

struct Reader
{
	FILE *fp;
	std::string path;
	void read( void *p , size_t n ){ if( n && fread( p , 1 , n , fp )!=n ) throw std::string( "truncated .tree file: " ) + path; }
	template< typename T > T get( void ){ T v ; read( &v , sizeof(T) ) ; return v; }
	// the indexer is the last thing in the file: the read position must be the end
	void end( void )
	{
		const long here = ftell( fp );
		if( fseek( fp , 0 , SEEK_END ) || ftell( fp )!=here ) throw std::string( here>ftell( fp ) ? "truncated .tree file: " : "trailing bytes in .tree file: " ) + path;
	}
	std::vector< std::vector< double > > vectors( void )
	{
		const unsigned long long n = get< unsigned long long >();
		if( n>64 ) throw std::string( "bad depth count in " ) + path;
		std::vector< std::vector< double > > out( n );
		for( auto &v : out )
		{
			const unsigned long long s = get< unsigned long long >();
			if( s>( 1ull<<36 ) ) throw std::string( "bad vector size in " ) + path;
			v.resize( s );
			read( v.data() , s*sizeof(double) );
		}
		return out;
	}
};

// typeid( MishaK::Hat::Hierarchical{Regular,Adapted}Indexer< dim > ).hash_code() of a
// libstdc++ build: std::_Hash_bytes of the mangled name, seed 0xc70f6907 (the
// value ExteriorPoissonRecon.cpp:263 writes and Extract.cpp:352-361 dispatches on).
static unsigned long long IndexerHash( int dim , bool adapted )
{
	const std::string name = std::string( adapted ? "N6MishaK3Hat26HierarchicalAdaptedIndexerILj" : "N6MishaK3Hat26HierarchicalRegularIndexerILj" ) + std::to_string( dim ) + "EEE";
	return (unsigned long long)std::_Hash_bytes( name.c_str() , name.size() , static_cast< size_t >( 0xc70f6907UL ) );
}

static TreeData ReadTree( const char *path , const unsigned long long hashes[2][2] /* [dim-3][adapted] */ )
{
	TreeData t;
	Reader r{ fopen( path , "rb" ) , path };
	if( !r.fp ) throw std::string( "cannot open " ) + path;
	struct Closer { FILE *f ; ~Closer(){ fclose( f ); } } closer{ r.fp };
	t.dim = (int)r.get< unsigned >();
	if( t.dim!=3 && t.dim!=4 ) throw std::string( "Only dimensions 3 and 4 supported: " ) + std::to_string( t.dim );
	t.hash = r.get< unsigned long long >();
	if     ( t.hash==hashes[ t.dim-3 ][0] ) t.adapted = false;
	else if( t.hash==hashes[ t.dim-3 ][1] ) t.adapted = true;
	else throw std::string( "Unrecognized hierarchical indexer hash: " ) + std::to_string( t.hash );
	const int dim = t.dim;
	(void)r.get< unsigned long long >();                  // XForm's vtable slot
	t.xform.resize( ( dim+1 )*( dim+1 ) );
	r.read( t.xform.data() , t.xform.size()*sizeof(double) );
	t.x = r.vectors() , t.y = r.vectors();
	if( t.x.empty() || t.x.size()!=t.y.size() ) throw std::string( "mismatched coefficient depths in " ) + path;
	t.D = (int)t.x.size() - 1;
	if( t.D*dim>30 ) throw std::string( "depth too large for the extraction's 32-bit element indices" );
	for( int d=0 ; d<=t.D ; d++ ) if( t.x[d].size()!=t.y[d].size() ) throw std::string( "mismatched coefficient sizes in " ) + path;
	if( !t.adapted )
	{
		const unsigned maxDepth = r.get< unsigned >();
		r.end();
		if( (int)maxDepth!=t.D ) throw std::string( "indexer depth does not match the coefficients" );
		for( int d=0 ; d<=t.D ; d++ )
		{
			unsigned long long n = 1 ; for( int k=0 ; k<dim ; k++ ) n *= ( 1ull<<d ) + 1;
			if( t.x[d].size()!=n ) throw std::string( "dense coefficient count mismatch at depth " ) + std::to_string( d );
		}
		return t;
	}
	// HierarchicalAdaptedIndexer::write (Hat.adaptive.inl:408-414): the tree
	// depth first (RegularTree.inl:616-662), then active sizes and the tables.
	const size_t head = ( ( 2*( dim+1 ) + 7 )/8 )*8 , rec = head + 32;
	std::vector< unsigned char > buf( rec );
	t.fOff.assign( t.D+1 , {} );
	for( int d=0 ; d<=t.D ; d++ ) t.fOff[d].assign( t.x[d].size()*dim , 0xffffffffu );
	std::vector< unsigned > fin;
	auto Record = [&]( void )
	{
		r.read( buf.data() , rec );
		unsigned short h[5];
		memcpy( h , buf.data() , sizeof(unsigned short)*( dim+1 ) );
		unsigned long long w[4];
		memcpy( w , buf.data() + head , 32 );
		const int depth = (int)h[0] - 1;
		if( w[2]!=NONE64 )
		{
			if( depth<0 || depth>t.D || w[2]>=t.x[depth].size() ) throw std::string( "bad function index in the adapted tree" );
			for( int k=0 ; k<dim ; k++ ) t.fOff[depth][ w[2]*dim+k ] = h[1+k];
		}
		if( w[3]!=NONE64 && depth==t.D ) for( int k=0 ; k<dim ; k++ ) fin.push_back( h[1+k] );
		return w[1]!=0;
	};
	// a node with children is followed by its 2^dim children, then each child's subtree
	struct Walk
	{
		int dim;
		decltype( Record ) &rec;
		void children( bool has , int depth )
		{
			if( !has ) return;
			if( depth>32 ) throw std::string( "adapted tree too deep" );
			bool c[16];
			for( int i=0 ; i<( 1<<dim ) ; i++ ) c[i] = rec();
			for( int i=0 ; i<( 1<<dim ) ; i++ ) children( c[i] , depth+1 );
		}
	} walk{ dim , Record };
	walk.children( Record() , 0 );
	t.eFinest = fin;
	for( int d=0 ; d<=t.D ; d++ ) for( size_t i=0 ; i<t.fOff[d].size() ; i++ ) if( t.fOff[d][i]==0xffffffffu ) throw std::string( "adapted tree misses a function node" );
	// active sizes and the node tables (their 8-byte pointers carry nothing)
	const unsigned long long na = r.get< unsigned long long >();
	for( unsigned long long i=0 ; i<na ; i++ ) (void)r.get< unsigned long long >();
	for( int tab=0 ; tab<2 ; tab++ )
	{
		const unsigned long long nd = r.get< unsigned long long >();
		if( (int)nd!=t.D+1 ) throw std::string( "adapted table depth mismatch" );
		for( unsigned long long d=0 ; d<nd ; d++ )
		{
			const unsigned long long s = r.get< unsigned long long >();
			if( tab==0 && s!=t.x[d].size() ) throw std::string( "adapted function count mismatch" );
			if( s>( 1ull<<36 ) || fseek( r.fp , (long)( s*8 ) , SEEK_CUR ) ) throw std::string( "truncated .tree file: " ) + path;
		}
	}
	r.end();
	return t;
}

// ------------------------------------------------------------------ device kernels
template< int Dim >
__device__ __forceinline__ long long FlatF( const int F[Dim] , int res1 )
{
	long long f = 0;
	for( int d=0 ; d<Dim ; d++ ) f = f*res1 + F[d];
	return f;
}

// bit b of coordinate d sits at bit b*Dim+d: tree order, child index bit d = coordinate d
template< int Dim >
__device__ __forceinline__ void MortonDecode( unsigned m , int D , int E[Dim] )
{
	for( int d=0 ; d<Dim ; d++ ) E[d] = 0;
	for( int b=0 ; b<D ; b++ ) for( int d=0 ; d<Dim ; d++ ) E[d] |= (int)( ( m>>( b*Dim+d ) ) & 1u ) << b;
}
template< int Dim >
__device__ __forceinline__ unsigned MortonEncode( const int E[Dim] , int D )
{
	unsigned m = 0;
	for( int b=0 ; b<D ; b++ ) for( int d=0 ; d<Dim ; d++ ) m |= (unsigned)( ( E[d]>>b ) & 1 ) << ( b*Dim+d );
	return m;
}

// V_f = x_f + sum over the parents of f, in window order (coordinate 0 most
// significant, lower parent first), of V_parent * 0.5^(odd coordinates)
// (IsoTree.inl:314-338); pairs interleaved (component 0, component 1).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KRegular( const double *Vc , const double *hx , const double *hy , double *Vf , int cres )
{
	const int fres = 2*cres;
	long long nf = 1 ; for( int d=0 ; d<Dim ; d++ ) nf *= fres+1;
	const long long f = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( f>=nf ) return;
	int F[Dim];
	{
		long long r = f;
		for( int d=Dim-1 ; d>=0 ; d-- ){ F[d] = (int)( r % ( fres+1 ) ) ; r /= fres+1; }
	}
	int nodd = 0;
	for( int d=0 ; d<Dim ; d++ ) nodd += F[d]&1;
	double w = 1;
	for( int k=0 ; k<nodd ; k++ ) w *= 0.5;
	double s0 = 0 , s1 = 0;
	for( int m=0 ; m<( 1<<nodd ) ; m++ )
	{
		long long c = 0;
		int oi = 0;
		for( int d=0 ; d<Dim ; d++ )
		{
			int cd = F[d]>>1;
			if( F[d]&1 ){ cd += ( m>>( nodd-1-oi ) ) & 1 ; oi++; }
			c = c*( cres+1 ) + cd;
		}
		s0 = s0 + Vc[2*c]*w;
		s1 = s1 + Vc[2*c+1]*w;
	}
	Vf[2*f] = hx[f] + s0;
	Vf[2*f+1] = hy[f] + s1;
}

__global__ void __launch_bounds__( 256 ) KInterleave( const double *a , const double *b , double *out , long long n )
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( i<n ) out[2*i] = a[i] , out[2*i+1] = b[i];
}

// IsoTree::isZeroCrossingNode (IsoTree.inl:135-163) of every finest element, by Morton index
template< int Dim >
__global__ void __launch_bounds__( 256 ) KZeroCross( const double *V , int D , unsigned nE , int all , unsigned char *zc )
{
	const unsigned m = blockIdx.x*blockDim.x + threadIdx.x;
	if( m>=nE ) return;
	int E[Dim];
	MortonDecode< Dim >( m , D , E );
	const int res1 = ( 1<<D ) + 1;
	unsigned f0 = 0 , f1 = 0;
	for( int i=0 ; i<( 1<<Dim ) ; i++ )
	{
		int F[Dim];
		for( int d=0 ; d<Dim ; d++ ) F[d] = E[d] + ( ( i>>d ) & 1 );
		const long long f = FlatF< Dim >( F , res1 );
		const double v0 = V[2*f] , v1 = V[2*f+1];
		if( v0<0 ) f0 |= 1 ; else if( v0>0 ) f0 |= 2 ; else f0 |= 3;
		if( v1<0 ) f1 |= 1 ; else if( v1>0 ) f1 |= 2 ; else f1 |= 3;
	}
	zc[m] = all ? ( f0==3 && f1==3 ) : ( f0==3 || f1==3 );
}

// union-find over the zero-crossing elements, face adjacency (IsoTree.inl:205-212:
// offsets differing in exactly one coordinate, by one, inside the grid).
// Roots only change through the compare-and-swap; path halving rewrites
// non-roots with one of their ancestors.
__device__ __forceinline__ unsigned UfLoad( const unsigned *p , unsigned i ){ return __hip_atomic_load( p + i , __ATOMIC_RELAXED , __HIP_MEMORY_SCOPE_AGENT ); }
__device__ __forceinline__ unsigned UfFind( unsigned *parent , unsigned x )
{
	unsigned p = UfLoad( parent , x );
	while( p!=x )
	{
		const unsigned g = UfLoad( parent , p );
		if( g!=p ) __hip_atomic_store( parent + x , g , __ATOMIC_RELAXED , __HIP_MEMORY_SCOPE_AGENT );
		x = p , p = g;
	}
	return x;
}

__global__ void __launch_bounds__( 256 ) KUfInit( unsigned *parent , unsigned nE )
{
	const unsigned m = blockIdx.x*blockDim.x + threadIdx.x;
	if( m<nE ) parent[m] = m;
}

template< int Dim >
__global__ void __launch_bounds__( 256 ) KUfLink( const unsigned char *zc , int D , unsigned nE , unsigned *parent )
{
	const unsigned m = blockIdx.x*blockDim.x + threadIdx.x;
	if( m>=nE || !zc[m] ) return;
	int E[Dim];
	MortonDecode< Dim >( m , D , E );
	for( int d=0 ; d<Dim ; d++ )
	{
		if( E[d]+1>=( 1<<D ) ) continue;
		E[d]++;
		const unsigned n = MortonEncode< Dim >( E , D );
		E[d]--;
		if( !zc[n] ) continue;
		unsigned a = m , b = n;
		while( true )
		{
			a = UfFind( parent , a ) , b = UfFind( parent , b );
			if( a==b ) break;
			if( a>b ){ const unsigned t = a ; a = b ; b = t; }
			// link the larger root under the smaller one
			unsigned expected = b;
			if( __hip_atomic_compare_exchange_strong( parent + b , &expected , a , __ATOMIC_RELAXED , __ATOMIC_RELAXED , __HIP_MEMORY_SCOPE_AGENT ) ) break;
			b = expected;
		}
	}
}

// roots of components that hold an indexed (seed) element
__global__ void __launch_bounds__( 256 ) KUfSeed( const unsigned char *zc , const unsigned char *indexed , unsigned nE , unsigned *parent , unsigned char *seedRoot )
{
	const unsigned m = blockIdx.x*blockDim.x + threadIdx.x;
	if( m>=nE || !zc[m] || !indexed[m] ) return;
	seedRoot[ UfFind( parent , m ) ] = 1;
}

__global__ void __launch_bounds__( 256 ) KUfKeep( const unsigned char *zc , unsigned nE , unsigned *parent , const unsigned char *seedRoot , unsigned char *keep )
{
	const unsigned m = blockIdx.x*blockDim.x + threadIdx.x;
	if( m>=nE ) return;
	keep[m] = zc[m] && seedRoot[ UfFind( parent , m ) ];
}

template< int Dim >
__global__ void __launch_bounds__( 256 ) KCells( const unsigned *morton , unsigned n , int D , unsigned *cells )
{
	const unsigned i = blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=n ) return;
	int E[Dim];
	MortonDecode< Dim >( morton[i] , D , E );
	for( int d=0 ; d<Dim ; d++ ) cells[ (size_t)i*Dim+d ] = (unsigned)E[d];
}

// IsoTree::functionValue at the rounded vertex positions (Extract.cpp:191-197)
template< int Dim >
__global__ void __launch_bounds__( 256 ) KVertexValues( const double *X , size_t nV , const double *V , int D , double *vals )
{
	const size_t v = (size_t)blockIdx.x*blockDim.x + threadIdx.x;
	if( v>=nV ) return;
	int F[Dim];
	for( int d=0 ; d<Dim ; d++ ) F[d] = (int)floor( X[ v*Dim+d ] + 0.5 );
	const long long f = FlatF< Dim >( F , ( 1<<D ) + 1 );
	vals[2*v] = V[2*f] , vals[2*v+1] = V[2*f+1];
}

__global__ void __launch_bounds__( 256 ) KColumn( const double *pairs , size_t n , int c , double *out )
{
	const size_t i = (size_t)blockIdx.x*blockDim.x + threadIdx.x;
	if( i<n ) out[i] = pairs[2*i+c];
}

struct XF { double m[5][5]; };

// voxelToWorld( p ) (Misha/Geometry.inl:413-427): out_i = sum_j p_j m[j][i] + m[Dim][i], over 1 + sum_i m[i][Dim]
template< int Dim >
__device__ __forceinline__ void Transform( const XF &xf , const double p[Dim] , double out[Dim] )
{
	double scale = 1;
	for( int i=0 ; i<Dim ; i++ )
	{
		out[i] = 0;
		for( int j=0 ; j<Dim ; j++ ) out[i] = out[i] + p[j]*xf.m[j][i];
		out[i] = out[i] + xf.m[Dim][i];
		scale = scale + xf.m[i][Dim];
	}
	for( int i=0 ; i<Dim ; i++ ) out[i] = out[i] / scale;
}

// Per level-set vertex (grid units): the IsoTree gradient of both functions
// (IsoTree.inl:101-133) on the leaf that getLeafNode reaches (a coordinate on
// a cell face descends to the lower cell: RegularTree.inl:248-262 with
// ChildIndex's strict '>'), with local coordinates from the clamped floor; the
// frame is grad * (1/res) * (1/scale) (Extract.cpp:253); the position is
// mapped to world space.  valid(leaf) is the IsoTree's valid-element test:
// an indexed finest element or a face neighbour of a kept one.
template< int Dim >
__global__ void __launch_bounds__( 256 ) KFrames( const double *X , size_t nV , const double *V , int D , double invScale , XF xf ,
                                                   const unsigned char *keep , const unsigned char *indexed , double *out , unsigned *invalid )
{
	const size_t v = (size_t)blockIdx.x*blockDim.x + threadIdx.x;
	if( v>=nV ) return;
	const int res = 1<<D;
	const double dres = (double)res , invRes = 1.0/dres;
	double g[Dim] , pl[Dim];
	int leaf[Dim] , E[Dim];
	for( int d=0 ; d<Dim ; d++ )
	{
		g[d] = X[ v*Dim+d ];
		const double p = g[d]*invRes;                    // vertex / res
		const double pr = p*dres;
		const double c = ceil( pr );
		leaf[d] = pr>0 ? (int)c - 1 : 0;
		leaf[d] = max( 0 , min( res-1 , leaf[d] ) );
		E[d] = max( 0 , min( (int)floor( pr ) , res-1 ) );
		pl[d] = pr - (double)E[d];
	}
	// the reference stops with "Leaf node is not a valid element node" otherwise
	{
		const unsigned m = MortonEncode< Dim >( leaf , D );
		bool ok = keep[m] || ( indexed ? indexed[m]!=0 : true );
		for( int d=0 ; d<Dim && !ok ; d++ ) for( int s=-1 ; s<=1 && !ok ; s+=2 )
		{
			const int e = leaf[d] + s;
			if( e<0 || e>=res ) continue;
			leaf[d] = e;
			ok = keep[ MortonEncode< Dim >( leaf , D ) ]!=0;
			leaf[d] -= s;
		}
		if( !ok ) atomicAdd( invalid , 1u );
	}
	double grad[2][Dim];
	for( int c=0 ; c<2 ; c++ ) for( int d=0 ; d<Dim ; d++ ) grad[c][d] = 0;
	const int res1 = res + 1;
	for( int i=0 ; i<( 1<<Dim ) ; i++ )                 // Range order: coordinate 0 slowest
	{
		int I[Dim] , F[Dim];
		for( int d=0 ; d<Dim ; d++ ) I[d] = ( i>>( Dim-1-d ) ) & 1 , F[d] = leaf[d] + I[d];
		const long long f = FlatF< Dim >( F , res1 );
		const double val[2] = { V[2*f] , V[2*f+1] };
		for( int d=0 ; d<Dim ; d++ )
		{
			double weight = dres;
			for( int dd=0 ; dd<Dim ; dd++ )
			{
				if( dd==d ) weight = weight * ( I[dd]==1 ? 1. : -1. );
				else        weight = weight * ( I[dd]==1 ? pl[dd] : 1.-pl[dd] );
			}
			for( int c=0 ; c<2 ; c++ ) grad[c][d] = grad[c][d] + val[c]*weight;
		}
	}
	double *o = out + v*3*Dim;
	double w[Dim];
	Transform< Dim >( xf , g , w );
	for( int d=0 ; d<Dim ; d++ ) o[d] = w[d];
	for( int c=0 ; c<2 ; c++ ) for( int d=0 ; d<Dim ; d++ ) o[ Dim*( 1+c )+d ] = ( grad[c][d]*invRes )*invScale;
}

template< int Dim >
__global__ void __launch_bounds__( 256 ) KToWorld( const double *X , size_t nV , XF xf , double *out )
{
	const size_t v = (size_t)blockIdx.x*blockDim.x + threadIdx.x;
	if( v>=nV ) return;
	double p[Dim] , w[Dim];
	for( int d=0 ; d<Dim ; d++ ) p[d] = X[ v*Dim+d ];
	Transform< Dim >( xf , p , w );
	for( int d=0 ; d<Dim ; d++ ) out[ v*Dim+d ] = w[d];
}

static inline unsigned Grid( long long n ){ return (unsigned)std::max< long long >( 1 , ( n + 255 )/256 ); }

// ------------------------------------------------------------------ orientation
// MarchingSimplices::Orient (Misha/MarchingSimplices.h:179-246) for (K+1)-vertex
// simplices.  Face keys are the sorted vertex sets; the two simplices incident
// on a face are the first two in (simplex, face) order, a third is an error.
// The face links are found on the device (keys, a radix sort, neighbours in
// the sorted order); the traversal runs on the host: seeds in index order, a
// depth-first stack visiting the faces of each simplex in the order of its
// current (possibly flipped) vertices, a flip swapping the first two vertices.
constexpr unsigned NONE32 = 0xffffffffu;

// the face opposite vertex v: for K=1 the other vertex (orientation: v==0),
// for K=2 the edge {si[v+1], si[v+2]} cyclically (orientation: increasing)
// (OrientedFaceIndex, MarchingSimplices.h:121-146; no swap for even K)
template< int K >
__global__ void __launch_bounds__( 256 ) KFaceRecords( const unsigned *S , unsigned nR , int vb , unsigned long long *key , unsigned *id , unsigned char *even )
{
	const unsigned r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=nR ) return;
	const unsigned *si = S + ( r/( K+1 ) )*( K+1 );
	const int v = r%( K+1 );
	if( K==1 ){ key[r] = si[ v==0 ? 1 : 0 ] , even[r] = v==0; }
	else
	{
		const unsigned a = si[ ( v+1 )%3 ] , b = si[ ( v+2 )%3 ];
		even[r] = a<b;
		key[r] = a<b ? ( (unsigned long long)a<<vb | b ) : ( (unsigned long long)b<<vb | a );
	}
	id[r] = r;
}

template< int K >
__global__ void __launch_bounds__( 256 ) KFaceLink( const unsigned long long *sk , const unsigned *sid , unsigned nR , const unsigned char *even ,
                                                     unsigned *other , unsigned char *same , unsigned *err )
{
	const unsigned i = blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=nR ) return;
	const unsigned long long k = sk[i];
	if( i+2<nR && sk[i+2]==k ) atomicOr( err , 1u );
	unsigned p = NONE32;
	if( i>0 && sk[i-1]==k ) p = sid[i-1];
	else if( i+1<nR && sk[i+1]==k ) p = sid[i+1];
	const unsigned r = sid[i];
	other[r] = p==NONE32 ? NONE32 : p/( K+1 );
	same[r] = p==NONE32 ? 0 : ( even[r]==even[p] );
}

// The face links on the device: other[r] / same[r] per (simplex, face) record r
struct Links
{
	DevBuf< unsigned > other;
	DevBuf< unsigned char > same;
};

template< int K >
static void LinkFaces( const unsigned *dS , size_t nS , size_t nV , hipStream_t st , Links &L )
{
	const size_t nR = nS*( K+1 );
	if( nR>=NONE32 ) throw std::string( "too many simplices for 32-bit face records" );
	int vb = 1;
	while( ( 1ull<<vb )<nV ) vb++;
	const int bits = K==1 ? vb : 2*vb;
	DevBuf< unsigned long long > key , skey;
	DevBuf< unsigned > id , sid , err;
	DevBuf< unsigned char > even , tmp;
	key.alloc( nR ) , skey.alloc( nR ) , id.alloc( nR ) , sid.alloc( nR ) , even.alloc( nR ) , L.other.alloc( nR ) , L.same.alloc( nR ) , err.alloc( 1 );
	err.zero( st );
	KFaceRecords< K ><<< Grid( nR ) , 256 , 0 , st >>>( dS , (unsigned)nR , vb , key.p , id.p , even.p );
	size_t bytes = 0;
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( nullptr , bytes , key.p , skey.p , id.p , sid.p , (int)nR , 0 , bits , st ) );
	tmp.ensure( bytes );
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( tmp.p , bytes , key.p , skey.p , id.p , sid.p , (int)nR , 0 , bits , st ) );
	KFaceLink< K ><<< Grid( nR ) , 256 , 0 , st >>>( skey.p , sid.p , (unsigned)nR , even.p , L.other.p , L.same.p , err.p );
	HIPCHECK( hipGetLastError() );
	unsigned e = 0;
	err.download( &e , 1 , st );
	HIPCHECK( hipStreamSynchronize( st ) );
	if( e ) throw std::string( "Both faces occupied" );
}

// ---- orientation of every orientable component on the device.  Crossing a
// shared face, the walk flips the neighbour exactly when both incidences carry
// the same orientation, so flip(o) = flip(s) xor same: a parity constraint per
// link.  A union-find with the parity to the parent (entry = parent << 1 |
// parity) solves it; the walk's seed is the component's smallest simplex and
// keeps its orientation, so flip(s) = parity(s) xor parity(seed).  A component
// whose links contradict each other is not orientable: there the walk's order
// decides, and the host walks it as the reference does.
__device__ __forceinline__ unsigned long long PLoad( const unsigned long long *p ){ return __hip_atomic_load( p , __ATOMIC_RELAXED , __HIP_MEMORY_SCOPE_AGENT ); }

__device__ __forceinline__ void PFind( unsigned long long *node , unsigned x , unsigned &root , unsigned &par )
{
	unsigned acc = 0;
	unsigned long long e = PLoad( node + x );
	while( (unsigned)( e>>1 )!=x )
	{
		const unsigned p = (unsigned)( e>>1 ) , a = (unsigned)( e&1 );
		const unsigned long long f = PLoad( node + p );
		const unsigned g = (unsigned)( f>>1 ) , b = (unsigned)( f&1 );
		if( g!=p )
		{
			// path halving: x's parent becomes its grandparent, with the combined parity
			__hip_atomic_store( node + x , ( (unsigned long long)g<<1 ) | ( a^b ) , __ATOMIC_RELAXED , __HIP_MEMORY_SCOPE_AGENT );
			acc ^= a^b , x = g , e = PLoad( node + g );
		}
		else acc ^= a , x = p , e = f;
	}
	root = x , par = acc;
}

__global__ void __launch_bounds__( 256 ) KParityInit( unsigned long long *node , unsigned nS , unsigned *seed )
{
	const unsigned s = blockIdx.x*blockDim.x + threadIdx.x;
	if( s<nS ) node[s] = (unsigned long long)s<<1 , seed[s] = NONE32;
}

template< int K >
__global__ void __launch_bounds__( 256 ) KParityLink( const unsigned *other , const unsigned char *same , unsigned nR , unsigned long long *node )
{
	const unsigned r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=nR ) return;
	const unsigned a = r/( K+1 ) , b = other[r];
	if( b==NONE32 || b<a ) return;                    // each link once
	const unsigned w = same[r];
	while( true )
	{
		unsigned ra , pa , rb , pb;
		PFind( node , a , ra , pa ) , PFind( node , b , rb , pb );
		if( ra==rb ) return;
		const unsigned lo = ra<rb ? ra : rb , hi = ra<rb ? rb : ra;
		unsigned long long expected = (unsigned long long)hi<<1;
		if( __hip_atomic_compare_exchange_strong( node + hi , &expected , ( (unsigned long long)lo<<1 ) | ( pa^pb^w ) ,
		                                          __ATOMIC_RELAXED , __ATOMIC_RELAXED , __HIP_MEMORY_SCOPE_AGENT ) ) return;
	}
}

__global__ void __launch_bounds__( 256 ) KParityRoot( unsigned long long *node , unsigned nS , unsigned *rootOf , unsigned char *par , unsigned *seed )
{
	const unsigned s = blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=nS ) return;
	unsigned r , p;
	PFind( node , s , r , p );
	rootOf[s] = r , par[s] = (unsigned char)p;
	atomicMin( seed + r , s );
}

template< int K >
__global__ void __launch_bounds__( 256 ) KParityCheck( const unsigned *other , const unsigned char *same , unsigned nR , const unsigned *rootOf ,
                                                        const unsigned char *par , unsigned char *bad )
{
	const unsigned r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=nR ) return;
	const unsigned a = r/( K+1 ) , b = other[r];
	if( b!=NONE32 && ( par[a]^par[b] )!=same[r] ) bad[ rootOf[a] ] = 1;
}

// orientable components: swap the first two vertices of the flipped simplices
template< int K >
__global__ void __launch_bounds__( 256 ) KParityApply( unsigned *S , unsigned nS , const unsigned *rootOf , const unsigned char *par , const unsigned *seed ,
                                                        const unsigned char *bad , unsigned char *flip )
{
	const unsigned s = blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=nS ) return;
	const unsigned r = rootOf[s];
	if( bad[r] ){ flip[s] = 2 ; return; }             // 2: left to the host walk
	const unsigned char f = par[s] ^ par[ seed[r] ];
	flip[s] = f;
	if( f ){ const unsigned t = S[ (size_t)s*( K+1 ) ] ; S[ (size_t)s*( K+1 ) ] = S[ (size_t)s*( K+1 )+1 ] ; S[ (size_t)s*( K+1 )+1 ] = t; }
}

// The reference's walk (seeds in index order, a stack, faces in the order of
// each simplex's current vertices, a flip swapping the first two vertices),
// over the simplices marked 2 (non-orientable components) only.
template< int K >
static void OrientWalk( std::vector< unsigned > &S , const std::vector< unsigned > &other , const std::vector< unsigned char > &same , std::vector< unsigned char > &flipped )
{
	const size_t nS = S.size()/( K+1 );
	std::vector< unsigned char > oriented( nS , 0 );
	std::vector< std::pair< unsigned , bool > > stack;
	for( size_t idx=0 ; idx<nS ; idx++ )
	{
		if( flipped[idx]!=2 || oriented[idx] ) continue;
		flipped[idx] = 0;
		oriented[idx] = 1;
		stack.push_back( { (unsigned)idx , true } );
		while( !stack.empty() )
		{
			const std::pair< unsigned , bool > s = stack.back();
			stack.pop_back();
			for( int v=0 ; v<=K ; v++ )
			{
				// after a flip the faces opposite vertices 0 and 1 trade places
				const int vo = flipped[ s.first ]==1 && v<2 ? 1-v : v;
				const size_t r = (size_t)s.first*( K+1 )+vo;
				const unsigned o = other[r];
				if( o==NONE32 || oriented[o] ) continue;
				oriented[o] = 1;
				if( s.second!=( same[r]!=0 ) ){ flipped[o] = 0 ; stack.push_back( { o , true } ); }
				else
				{
					std::swap( S[ (size_t)o*( K+1 ) ] , S[ (size_t)o*( K+1 )+1 ] );
					flipped[o] = 1;
					stack.push_back( { o , false } );
				}
			}
		}
	}
}

// MarchingSimplices::Orient of the device simplices dS (nS x (K+1)): S receives
// the oriented simplices, flipped the per-simplex flips; other (host) is filled
// when wantLinks or a non-orientable component needs the walk
template< int K >
static void Orient( unsigned *dS , size_t nS , size_t nV , hipStream_t st , bool wantLinks , std::vector< unsigned > &S ,
                    std::vector< unsigned char > &flipped , std::vector< unsigned > &other , double *walkSeconds=nullptr )
{
	S.resize( nS*( K+1 ) ) , flipped.assign( nS , 0 ) , other.clear();
	if( !nS ) return;
	Links L;
	LinkFaces< K >( dS , nS , nV , st , L );
	const size_t nR = nS*( K+1 );
	DevBuf< unsigned long long > node;
	DevBuf< unsigned > rootOf , seed;
	DevBuf< unsigned char > par , bad , flip;
	node.alloc( nS ) , rootOf.alloc( nS ) , seed.alloc( nS ) , par.alloc( nS ) , bad.alloc( nS ) , flip.alloc( nS );
	bad.zero( st );
	KParityInit<<< Grid( nS ) , 256 , 0 , st >>>( node.p , (unsigned)nS , seed.p );
	KParityLink< K ><<< Grid( nR ) , 256 , 0 , st >>>( L.other.p , L.same.p , (unsigned)nR , node.p );
	KParityRoot<<< Grid( nS ) , 256 , 0 , st >>>( node.p , (unsigned)nS , rootOf.p , par.p , seed.p );
	KParityCheck< K ><<< Grid( nR ) , 256 , 0 , st >>>( L.other.p , L.same.p , (unsigned)nR , rootOf.p , par.p , bad.p );
	KParityApply< K ><<< Grid( nS ) , 256 , 0 , st >>>( dS , (unsigned)nS , rootOf.p , par.p , seed.p , bad.p , flip.p );
	HIPCHECK( hipGetLastError() );
	flip.download( flipped.data() , nS , st );
	HIPCHECK( hipMemcpyAsync( S.data() , dS , nR*sizeof(unsigned) , hipMemcpyDeviceToHost , st ) );
	HIPCHECK( hipStreamSynchronize( st ) );
	bool anyBad = false;
	for( size_t s=0 ; s<nS && !anyBad ; s++ ) anyBad = flipped[s]==2;
	if( wantLinks || anyBad )
	{
		other.resize( nR );
		L.other.download( other.data() , nR , st );
		HIPCHECK( hipStreamSynchronize( st ) );
	}
	if( anyBad )
	{
		Clock::time_point t0 = Clock::now();
		std::vector< unsigned char > same( nR );
		L.same.download( same.data() , nR , st );
		HIPCHECK( hipStreamSynchronize( st ) );
		OrientWalk< K >( S , other , same , flipped );
		if( walkSeconds ) *walkSeconds = Since( t0 );
	}
}

// ------------------------------------------------------------------ density trimming
// The reference's density file (ExteriorPoissonRecon.cpp:272-277): the
// voxel-to-world XForm, the TreeEstimator's tree written depth first
// (RegularTreeNode::write, Misha/RegularTree.inl:614-640: the root, then every
// brood followed by its children's subtrees), each node a raw
// RegularTreeNode< Dim , DensityAndNoiseInfo< Dim > , unsigned short > whose
// data starts with weightSum, then _targetSamplesPerCell.
struct DensityTree
{
	int dim = 0 , maxDepth = 0;           // maxDepth: the TreeSplatter's _maxDepth (tree height - 1)
	double target = 0;
	std::vector< int > child;             // first child of every node (broods are contiguous), -1 for leaves
	std::vector< double > w;              // weightSum
};

// xform (may be null) receives the file's voxel-to-world matrix, (dim+1)^2 doubles
static DensityTree ReadDensity( const char *path , int dim , double *xform=nullptr )
{
	DensityTree T;
	T.dim = dim;
	Reader r{ fopen( path , "rb" ) , path };
	if( !r.fp ) throw std::string( "cannot open " ) + path;
	struct Closer { FILE *f ; ~Closer(){ fclose( f ); } } closer{ r.fp };
	std::vector< unsigned char > skip( 8 + ( dim+1 )*( dim+1 )*8 );
	r.read( skip.data() , skip.size() );
	if( xform ) memcpy( xform , skip.data()+8 , ( dim+1 )*( dim+1 )*sizeof(double) );
	const size_t head = ( ( 2*( dim+1 ) + 7 )/8 )*8 , rec = head + 16 + 8*( 1 + dim + 1 + dim*dim );
	std::vector< unsigned char > buf( rec );
	int height = 0;
	auto Record = [&]( int depth )
	{
		r.read( buf.data() , rec );
		unsigned long long ch;
		double w;
		memcpy( &ch , buf.data() + head + 8 , 8 );
		memcpy( &w , buf.data() + head + 16 , 8 );
		T.child.push_back( -1 ) , T.w.push_back( w );
		height = std::max( height , depth );
		return ch!=0;
	};
	struct Walk
	{
		int dim;
		DensityTree &T;
		decltype( Record ) &rec;
		void children( int node , bool has , int depth )
		{
			if( !has ) return;
			if( depth>30 ) throw std::string( "density tree too deep" );
			const int first = (int)T.child.size();
			T.child[node] = first;
			bool c[16];
			for( int i=0 ; i<( 1<<dim ) ; i++ ) c[i] = rec( depth+1 );
			for( int i=0 ; i<( 1<<dim ) ; i++ ) children( first+i , c[i] , depth+1 );
		}
	} walk{ dim , T , Record };
	walk.children( 0 , Record( 0 ) , 0 );
	T.target = r.get< double >();
	r.end();
	T.maxDepth = height - 1;
	if( T.child[0]<0 ) throw std::string( "density tree without a space root" );
	return T;
}

struct DensityQuery { int kind , d ; double a , b; };   // kind 0: extremal, 1: interior, 2: interior with no samples at d2

// TreeEstimator::_samplesPerCell at depth t (GridSamples.inl:783-815) over the
// 2^Dim window of the ancestor at t (left radius 0, right radius 1), with the
// linear B-spline weights of the clamped floor cell (the reference's weights
// and window need not align on a cell face; that is kept).
template< int Dim >
__device__ double SamplesPerCellAt( const int *win , int t , const double p[Dim] , const double *w )
{
	const int res = 1<<( t-1 );
	double val[Dim][2];
	for( int d=0 ; d<Dim ; d++ )
	{
		const double x = p[d]*res;
		const int e = (int)floor( x );
		const double f = x - e;
		val[d][0] = 1.*( 1.-f )*1. , val[d][1] = 1.*( f*1. );
	}
	double sum = 0;
	for( int i=0 ; i<( 1<<Dim ) ; i++ )                 // Range order: coordinate 0 slowest
	{
		const int n = win[i];
		if( n<0 ) continue;
		double scale = 1.;
		for( int d=0 ; d<Dim ; d++ ) scale = scale * val[d][ ( i>>( Dim-1-d ) ) & 1 ];
		sum = sum + w[n]*scale;
	}
	return sum / (double)( 1<<( ( t-1 )*( Dim-2 ) ) );
}

// per level-set vertex: getLeafNode from the space root, the neighbour windows
// of its ancestors (children of the parent's window), and the branch of
// TreeEstimator::_depth (GridSamples.inl:863-896) with the samples-per-cell
// values it needs; the logarithms are taken on the host.
template< int Dim >
__global__ void __launch_bounds__( 128 ) KDensityQuery( const double *X , size_t nV , int D , const int *child , const double *w , double target , DensityQuery *out )
{
	const size_t v = (size_t)blockIdx.x*blockDim.x + threadIdx.x;
	if( v>=nV ) return;
	constexpr int MAXT = 18 , NW = 1<<Dim;
	const double invRes = 1.0/(double)( 1<<D );
	double p[Dim];
	for( int d=0 ; d<Dim ; d++ ) p[d] = X[ v*Dim+d ]*invRes;
	int win[MAXT][NW];
	int off[Dim];
	for( int d=0 ; d<Dim ; d++ ) off[d] = 0;
	// depth 1: the space root (child 0 of the root) and its siblings
	for( int i=0 ; i<NW ; i++ )
	{
		int c = 0;
		for( int d=0 ; d<Dim ; d++ ) c |= ( ( i>>( Dim-1-d ) ) & 1 ) << d;
		win[1][i] = child[0] + c;
	}
	int node = child[0] , t = 1;
	double center[Dim] , width = 1.;
	for( int d=0 ; d<Dim ; d++ ) center[d] = 0.5;
	while( child[node]>=0 && t+1<MAXT )
	{
		int c = 0;
		for( int d=0 ; d<Dim ; d++ ) if( p[d]>center[d] ) c |= 1<<d;
		node = child[node] + c;
		width /= 2;
		for( int d=0 ; d<Dim ; d++ ) center[d] += ( c&( 1<<d ) ) ? width/2 : -width/2;
		int noff[Dim];
		for( int d=0 ; d<Dim ; d++ ) noff[d] = 2*off[d] + ( ( c>>d ) & 1 );
		for( int i=0 ; i<NW ; i++ )
		{
			int J = 0 , ci = 0;
			for( int d=0 ; d<Dim ; d++ )
			{
				const int o = noff[d] + ( ( i>>( Dim-1-d ) ) & 1 );
				J = 2*J + ( ( o>>1 ) - off[d] );
				ci |= ( o&1 ) << d;
			}
			const int par = win[t][J];
			win[t+1][i] = par>=0 && child[par]>=0 ? child[par] + ci : -1;
		}
		for( int d=0 ; d<Dim ; d++ ) off[d] = noff[d];
		t++;
	}
	const int maxD = t;
	int d1 = maxD , d2 = maxD+1;
	while( d2>1 && SamplesPerCellAt< Dim >( win[d1] , d1 , p , w )<target ) d1-- , d2--;
	DensityQuery q;
	if( d2==1 || d2==maxD+1 )
	{
		const int d = d2==1 ? 1 : maxD;
		q = { 0 , d , SamplesPerCellAt< Dim >( win[d] , d , p , w ) , 0. };
	}
	else
	{
		const double s1 = SamplesPerCellAt< Dim >( win[d1] , d1 , p , w ) , s2 = SamplesPerCellAt< Dim >( win[d2] , d2 , p , w );
		q = { s2!=0 ? 1 : 2 , d1 , s1 , s2 };
	}
	out[v] = q;
}

// TreeEstimator::samplesPerCell (GridSamples.inl:845-850) from a query
static double SamplesPerCell( const DensityTree &T , const DensityQuery &q , int dim )
{
	double depth;
	if( q.kind==0 ) depth = q.d + log( q.a/T.target ) / log( 2. );
	else if( q.kind==2 ) depth = q.d;
	else depth = q.d + log( T.target/q.a ) / log( q.b/q.a );
	depth -= 1.;
	// pow( 0.5 , y ) in the reference, which its clang -O3 build lowers to
	// exp2( -y ); written out so the result does not depend on that rewrite
	return T.target * exp2( -( (double)( dim-2 ) * ( T.maxDepth - depth ) ) );
}

// The voxel coordinates of every node of the (2^D+1)^Dim grid, in the
// function index order (coordinate 0 most significant).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KGridNodes( long long N , int n1 , double *X )
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=N ) return;
	long long t = i;
	for( int d=Dim-1 ; d>=0 ; d-- ){ X[ i*Dim+d ] = (double)( t % n1 ) ; t /= n1; }
}

// The density estimator's samples per cell (TreeEstimator::samplesPerCell,
// GridSamples.inl:845-850, the quantity Extract --trimDensity thresholds) at
// every node of the depth-`depth` grid, written as a .grid with one DOUBLE per
// node (Misha/RegularGrid.inl:440-471 layout, the file's voxel-to-world
// transform in the header, as Extract --grid writes the coefficients).
template< int Dim >
static void DensityGrid( const char *densityPath , int depth , const char *gridPath )
{
	if( depth<0 || depth>14 ) throw std::string( "depth must be in [0,14]" );
	double xform[ ( Dim+1 )*( Dim+1 ) ];
	const DensityTree T = ReadDensity( densityPath , Dim , xform );
	const int n1 = ( 1<<depth ) + 1;
	long long N = 1 ; for( int d=0 ; d<Dim ; d++ ) N *= n1;
	hipStream_t st;
	HIPCHECK( hipStreamCreate( &st ) );
	struct StreamCloser { hipStream_t s ; ~StreamCloser(){ (void)hipStreamDestroy( s ); } } closer{ st };
	DevBuf< double > X , dW;
	DevBuf< int > dChild;
	DevBuf< DensityQuery > q;
	X.alloc( (size_t)N*Dim ) , q.alloc( (size_t)N );
	dChild.upload( T.child.data() , T.child.size() , st );
	dW.upload( T.w.data() , T.w.size() , st );
	KGridNodes< Dim ><<< Blocks( N , 256 ) , 256 , 0 , st >>>( N , n1 , X.p );
	KDensityQuery< Dim ><<< Blocks( N , 128 ) , 128 , 0 , st >>>( X.p , (size_t)N , depth , dChild.p , dW.p , T.target , q.p );
	HIPCHECK( hipGetLastError() );
	std::vector< DensityQuery > hq( (size_t)N );
	q.download( hq.data() , hq.size() , st );
	HIPCHECK( hipStreamSynchronize( st ) );
	std::vector< double > v( (size_t)N );
	for( long long i=0 ; i<N ; i++ ) v[i] = SamplesPerCell( T , hq[i] , Dim );
	FILE *fp = fopen( gridPath , "wb" );
	if( !fp ) throw std::string( "Failed to open grid file for writing: " ) + gridPath;
	fprintf( fp , "G%d\n" , Dim );
	fprintf( fp , "1 DOUBLE\n" );
	for( int d=0 ; d<Dim ; d++ ) fprintf( fp , d==Dim-1 ? "%d\n" : "%d " , n1 );
	for( int j=0 ; j<=Dim ; j++ ) for( int i=0 ; i<=Dim ; i++ ) fprintf( fp , i==Dim ? "%f\n" : "%f " , (float)xform[ i*( Dim+1 )+j ] );
	const size_t wrote = fwrite( v.data() , sizeof(double) , (size_t)N , fp );
	fclose( fp );
	if( wrote!=(size_t)N ) throw std::string( "Failed to write " ) + gridPath;
}

// MarchingSimplices::FaceAdjacentConnectedComponents (MarchingSimplices.h:288-365)
// over the oriented simplices: seeds from the back of the unprocessed list,
// a stack, faces in the order of each simplex's current vertices; after a
// component, processed entries leave the list by swap-with-last in
// decreasing position (what the reference's backward scan does).
template< int K >
static std::vector< std::vector< unsigned > > Components( size_t nS , const std::vector< unsigned > &other , const std::vector< unsigned char > &flipped )
{
	std::vector< std::vector< unsigned > > comps;
	std::vector< unsigned char > processed( nS , 0 );
	std::vector< unsigned > un( nS ) , pos( nS );
	for( size_t i=0 ; i<nS ; i++ ) un[i] = (unsigned)i , pos[i] = (unsigned)i;
	std::vector< unsigned > Q;
	while( !un.empty() )
	{
		std::vector< unsigned > comp;
		Q.push_back( un.back() );
		processed[ un.back() ] = 1;
		pos[ un.back() ] = NONE32;
		un.pop_back();
		while( !Q.empty() )
		{
			const unsigned i = Q.back();
			Q.pop_back();
			comp.push_back( i );
			for( int v=0 ; v<=K ; v++ )
			{
				const int vo = flipped[i] && v<2 ? 1-v : v;
				const unsigned o = other[ (size_t)i*( K+1 )+vo ];
				if( o!=NONE32 && !processed[o] ) Q.push_back( o ) , processed[o] = 1;
			}
		}
		std::vector< unsigned > at;
		for( unsigned i : comp ) if( pos[i]!=NONE32 ) at.push_back( pos[i] );
		std::sort( at.begin() , at.end() , std::greater< unsigned >() );
		for( unsigned p : at )
		{
			const unsigned last = un.back();
			un[p] = last , pos[last] = p;
			un.pop_back();
		}
		for( unsigned i : comp ) pos[i] = NONE32;
		comps.push_back( std::move( comp ) );
	}
	return comps;
}
} // namespace iso
} // namespace epr

struct epr_iso
{
	int dim = 0 , depth = 0;
	bool adapted = false;
	unsigned long long hash = 0;
	std::vector< double > xform;
	hipStream_t st = nullptr;
	epr::DevBuf< double > V;                            // finest regular values, pairs
	epr::DevBuf< unsigned char > indexed;               // adapted: indexed finest elements by Morton index
	std::unique_ptr< epr::iso::DensityTree > density;   // a density file, when read
	epr::DevBuf< int > dChild;                          // its tree: first child per node
	epr::DevBuf< double > dW;                           // its weightSum per node
	std::vector< double > componentMax;                 // the last trimmed extraction's per-component maxima
	std::string err;
	double seconds[5] = { 0 , 0 , 0 , 0 , 0 };           // open, level-set nodes + triangulation, level set, orientation (+ trimming), frames
};

namespace epr
{
namespace iso
{
template< int Dim >
static void Upload( epr_iso &I , const TreeData &t )
{
	const int D = t.D;
	hipStream_t st = I.st;
	DevBuf< double > Vc , Vf , hx , hy;
	std::vector< double > dx , dy;
	for( int d=0 ; d<=D ; d++ )
	{
		const long long n1 = ( 1ll<<d ) + 1;
		long long N = 1 ; for( int k=0 ; k<Dim ; k++ ) N *= n1;
		const double *px = t.x[d].data() , *py = t.y[d].data();
		if( t.adapted )
		{
			dx.assign( N , 0. ) , dy.assign( N , 0. );
			const std::vector< unsigned > &F = t.fOff[d];
			for( size_t i=0 ; i<t.x[d].size() ; i++ )
			{
				long long f = 0;
				for( int k=0 ; k<Dim ; k++ )
				{
					if( F[ i*Dim+k ]>=(unsigned)n1 ) throw std::string( "adapted function offset outside the grid" );
					f = f*n1 + F[ i*Dim+k ];
				}
				dx[f] = t.x[d][i] , dy[f] = t.y[d][i];
			}
			px = dx.data() , py = dy.data();
		}
		hx.upload( px , N , st ) , hy.upload( py , N , st );
		Vf.alloc( 2*N );
		if( d==0 ) KInterleave<<< Grid( N ) , 256 , 0 , st >>>( hx.p , hy.p , Vf.p , N );
		else       KRegular< Dim ><<< Grid( N ) , 256 , 0 , st >>>( Vc.p , hx.p , hy.p , Vf.p , 1<<( d-1 ) );
		HIPCHECK( hipGetLastError() );
		HIPCHECK( hipStreamSynchronize( st ) );     // the host staging buffers are reused
		std::swap( Vc , Vf );
	}
	I.V = std::move( Vc );
	if( t.adapted )
	{
		const unsigned nE = 1u<<( D*Dim );
		std::vector< unsigned char > ind( nE , 0 );
		for( size_t i=0 ; i<t.eFinest.size()/Dim ; i++ )
		{
			unsigned m = 0;
			for( int b=0 ; b<D ; b++ ) for( int k=0 ; k<Dim ; k++ ) m |= ( ( t.eFinest[ i*Dim+k ]>>b ) & 1u ) << ( b*Dim+k );
			if( m>=nE ) throw std::string( "adapted element offset outside the grid" );
			ind[m] = 1;
		}
		I.indexed.upload( ind.data() , nE , st );
		HIPCHECK( hipStreamSynchronize( st ) );
	}
}

// the kept finest elements (levelSetNodes): keep flags by Morton index and the cell list in Morton order
template< int Dim >
static void LevelSetCells( epr_iso &I , bool all , DevBuf< unsigned char > &keep , DevBuf< unsigned > &cells , size_t &nCells )
{
	const int D = I.depth;
	const unsigned nE = 1u<<( D*Dim );
	hipStream_t st = I.st;
	DevBuf< unsigned char > zc , tmp;
	zc.alloc( nE );
	KZeroCross< Dim ><<< Grid( nE ) , 256 , 0 , st >>>( I.V.p , D , nE , all ? 1 : 0 , zc.p );
	if( I.adapted )
	{
		DevBuf< unsigned > parent;
		DevBuf< unsigned char > seedRoot;
		parent.alloc( nE ) , seedRoot.alloc( nE ) , keep.alloc( nE );
		seedRoot.zero( st );
		KUfInit<<< Grid( nE ) , 256 , 0 , st >>>( parent.p , nE );
		KUfLink< Dim ><<< Grid( nE ) , 256 , 0 , st >>>( zc.p , D , nE , parent.p );
		KUfSeed<<< Grid( nE ) , 256 , 0 , st >>>( zc.p , I.indexed.p , nE , parent.p , seedRoot.p );
		KUfKeep<<< Grid( nE ) , 256 , 0 , st >>>( zc.p , nE , parent.p , seedRoot.p , keep.p );
		HIPCHECK( hipGetLastError() );
		HIPCHECK( hipStreamSynchronize( st ) );
	}
	else keep = std::move( zc );
	DevBuf< unsigned > morton , count;
	morton.alloc( nE ) , count.alloc( 1 );
	size_t bytes = 0;
	hipcub::CountingInputIterator< unsigned > it( 0u );
	HIPCHECK( hipcub::DeviceSelect::Flagged( nullptr , bytes , it , keep.p , morton.p , count.p , (int)nE , st ) );
	tmp.ensure( bytes );
	HIPCHECK( hipcub::DeviceSelect::Flagged( tmp.p , bytes , it , keep.p , morton.p , count.p , (int)nE , st ) );
	unsigned n = 0;
	count.download( &n , 1 , st );
	HIPCHECK( hipStreamSynchronize( st ) );
	nCells = n;
	cells.alloc( (size_t)n*Dim );
	if( n ) KCells< Dim ><<< Grid( n ) , 256 , 0 , st >>>( morton.p , n , D , cells.p );
	HIPCHECK( hipGetLastError() );
	HIPCHECK( hipStreamSynchronize( st ) );
}

template< int Dim >
static void SubGrid( epr_iso &I , bool all , DevBuf< unsigned char > &keep , ls::DevMesh &tri , DevBuf< double > &vals )
{
	DevBuf< unsigned > cells;
	size_t nCells = 0;
	LevelSetCells< Dim >( I , all , keep , cells , nCells );
	const unsigned ext[4] = { ( 1u<<I.depth ) - 1 , ( 1u<<I.depth ) - 1 , ( 1u<<I.depth ) - 1 , ( 1u<<I.depth ) - 1 };
	ls::TriangulateDev( Dim , cells.p , nCells , ext , true , I.st , tri );
	vals.alloc( 2*tri.nV + 2 );
	if( tri.nV ) KVertexValues< Dim ><<< Grid( tri.nV ) , 256 , 0 , I.st >>>( tri.V.p , tri.nV , I.V.p , I.depth , vals.p );
	HIPCHECK( hipGetLastError() );
}

static XF MakeXF( const epr_iso &I )
{
	XF xf{};
	for( int c=0 ; c<=I.dim ; c++ ) for( int r=0 ; r<=I.dim ; r++ ) xf.m[c][r] = I.xform[ c*( I.dim+1 )+r ];
	return xf;
}

// Extract.cpp:149-151
static double FrameScale( const epr_iso &I )
{
	double scale = 1.;
	for( int d=0 ; d<I.dim ; d++ ) scale *= I.xform[ d*( I.dim+1 )+d ];
	return pow( scale , 1./I.dim );
}

// Extract.cpp:217-238: components of the oriented level set, the largest
// samples-per-cell estimate over each component's vertices, and only the
// components above the threshold kept, in component order
template< int Dim >
static void Trim( epr_iso &I , const ls::DevMesh &lvl , double trim , const std::vector< unsigned > &other , const std::vector< unsigned char > &flipped ,
                  std::vector< unsigned > &S )
{
	constexpr int K = Dim-2;
	const size_t nS = S.size()/( K+1 );
	std::vector< double > spc( lvl.nV );
	if( lvl.nV )
	{
		DevBuf< DensityQuery > q;
		q.alloc( lvl.nV );
		KDensityQuery< Dim ><<< (unsigned)( ( lvl.nV + 127 )/128 ) , 128 , 0 , I.st >>>( lvl.V.p , lvl.nV , I.depth , I.dChild.p , I.dW.p , I.density->target , q.p );
		HIPCHECK( hipGetLastError() );
		std::vector< DensityQuery > hq( lvl.nV );
		q.download( hq.data() , hq.size() , I.st );
		HIPCHECK( hipStreamSynchronize( I.st ) );
		for( size_t v=0 ; v<lvl.nV ; v++ ) spc[v] = SamplesPerCell( *I.density , hq[v] , Dim );
	}
	std::vector< std::vector< unsigned > > comps = Components< K >( nS , other , flipped );
	std::vector< unsigned > kept;
	for( const std::vector< unsigned > &c : comps )
	{
		double mx = 0;
		for( unsigned s : c ) for( int v=0 ; v<=K ; v++ ) mx = std::max< double >( mx , spc[ S[ (size_t)s*( K+1 )+v ] ] );
		I.componentMax.push_back( mx );
		if( mx>trim ) for( unsigned s : c ) for( int v=0 ; v<=K ; v++ ) kept.push_back( S[ (size_t)s*( K+1 )+v ] );
	}
	S.swap( kept );
}

template< int Dim >
static epr_mesh *Extract( epr_iso &I , bool all , double trim )
{
	hipStream_t st = I.st;
	Clock::time_point t0 = Clock::now();
	DevBuf< unsigned char > keep;
	ls::DevMesh tri , lvl;
	DevBuf< double > vals;
	SubGrid< Dim >( I , all , keep , tri , vals );
	HIPCHECK( hipStreamSynchronize( st ) );
	I.seconds[1] = Since( t0 ) , t0 = Clock::now();
	const double zero[2] = { 0 , 0 };
	ls::LevelSet2Dev< Dim >( Dim , tri.nV , tri.V.p , tri.nS , tri.S.p , vals.p , zero , st , lvl );
	HIPCHECK( hipStreamSynchronize( st ) );
	I.seconds[2] = Since( t0 ) , t0 = Clock::now();
	epr_mesh *m = new epr_mesh;
	m->vdim = 3*Dim , m->k = Dim-1;
	try
	{
		const bool trimming = I.density && trim>0;
		std::vector< unsigned > other;
		std::vector< unsigned char > flipped;
		Orient< Dim-2 >( lvl.S.p , lvl.nS , lvl.nV , st , trimming , m->s , flipped , other );
		I.componentMax.clear();
		if( trimming ) Trim< Dim >( I , lvl , trim , other , flipped , m->s );
	}
	catch( ... ){ delete m ; throw; }
	I.seconds[3] = Since( t0 ) , t0 = Clock::now();
	DevBuf< double > out;
	DevBuf< unsigned > invalid;
	out.alloc( lvl.nV*3*Dim + 1 ) , invalid.alloc( 1 );
	invalid.zero( st );
	const double scale = FrameScale( I );
	if( lvl.nV ) KFrames< Dim ><<< Grid( lvl.nV ) , 256 , 0 , st >>>( lvl.V.p , lvl.nV , I.V.p , I.depth , 1./scale , MakeXF( I ) , keep.p ,
	                                                                  I.adapted ? I.indexed.p : nullptr , out.p , invalid.p );
	HIPCHECK( hipGetLastError() );
	unsigned bad = 0;
	invalid.download( &bad , 1 , st );
	m->v.resize( lvl.nV*3*Dim );
	out.download( m->v.data() , m->v.size() , st );
	HIPCHECK( hipStreamSynchronize( st ) );
	I.seconds[4] = Since( t0 );
	if( bad ){ delete m ; throw std::string( "Leaf node is not a valid element node" ); }
	return m;
}

// Extract.cpp:296-326: the codimension-1 zero set of component c over the
// same triangulation (all = false), vertices in world space
template< int Dim >
static epr_mesh *LevelSet1( epr_iso &I , int c )
{
	hipStream_t st = I.st;
	DevBuf< unsigned char > keep;
	ls::DevMesh tri , lvl;
	DevBuf< double > vals , col , world;
	SubGrid< Dim >( I , false , keep , tri , vals );
	col.alloc( tri.nV + 1 );
	if( tri.nV ) KColumn<<< Grid( tri.nV ) , 256 , 0 , st >>>( vals.p , tri.nV , c , col.p );
	ls::LevelSet1Dev< Dim >( Dim , tri.nV , tri.V.p , tri.nS , tri.S.p , col.p , 0. , st , lvl );
	world.alloc( lvl.nV*Dim + 1 );
	if( lvl.nV ) KToWorld< Dim ><<< Grid( lvl.nV ) , 256 , 0 , st >>>( lvl.V.p , lvl.nV , MakeXF( I ) , world.p );
	HIPCHECK( hipGetLastError() );
	epr_mesh *m = new epr_mesh;
	m->vdim = Dim , m->k = Dim;
	m->v.resize( lvl.nV*Dim ) , m->s.resize( lvl.nS*Dim );
	world.download( m->v.data() , m->v.size() , st );
	lvl.S.download( m->s.data() , m->s.size() , st );
	HIPCHECK( hipStreamSynchronize( st ) );
	return m;
}

template< int Dim >
static void Cells( epr_iso &I , bool all , std::vector< unsigned > &out )
{
	DevBuf< unsigned char > keep;
	DevBuf< unsigned > cells;
	size_t n = 0;
	LevelSetCells< Dim >( I , all , keep , cells , n );
	out.resize( n*Dim );
	cells.download( out.data() , out.size() , I.st );
	HIPCHECK( hipStreamSynchronize( I.st ) );
}

// Extract.cpp:132-144: per component, RegularGrid<Dim,double>::write
// (Misha/RegularGrid.inl:441-471) of the finest regular coefficients
static void WriteGrid( epr_iso &I , const char *header )
{
	const int dim = I.dim;
	const long long n1 = ( 1ll<<I.depth ) + 1;
	long long N = 1 ; for( int k=0 ; k<dim ; k++ ) N *= n1;
	std::vector< double > V( 2*N );
	I.V.download( V.data() , V.size() , I.st );
	HIPCHECK( hipStreamSynchronize( I.st ) );
	std::vector< double > comp( N );
	for( int c=0 ; c<2 ; c++ )
	{
		for( long long i=0 ; i<N ; i++ ) comp[i] = V[ 2*i+c ];
		const std::string name = std::string( header ) + "." + std::to_string( c ) + ".grid";
		FILE *fp = fopen( name.c_str() , "wb" );
		if( !fp ) throw std::string( "Failed to open grid file for writing: " ) + name;
		fprintf( fp , "G%d\n" , dim );
		fprintf( fp , "1 DOUBLE\n" );
		for( int d=0 ; d<dim ; d++ ) fprintf( fp , d==dim-1 ? "%d\n" : "%d " , (int)n1 );
		for( int j=0 ; j<=dim ; j++ ) for( int i=0 ; i<=dim ; i++ )
			fprintf( fp , i==dim ? "%f\n" : "%f " , (float)I.xform[ i*( dim+1 )+j ] );
		fwrite( comp.data() , sizeof(double) , N , fp );
		fclose( fp );
	}
}
} // namespace iso
} // namespace epr

template< typename F >
static int IsoGuard( epr_iso *I , F f )
{
	if( !I ) return EPR_ERR_ARG;
	try { f() ; I->err.clear() ; return EPR_OK; }
	catch( const epr::HipError &e ){ I->err = e.msg ; return EPR_ERR_HIP; }
	catch( const std::string &s ){ I->err = s ; return EPR_ERR_STATE; }
	catch( const std::exception &e ){ I->err = e.what() ; return EPR_ERR_STATE; }
}

extern "C" {

int epr_iso_open( const char *path , epr_iso **out )
{
	if( !path || !out ){ epr::ThreadError() = "bad argument" ; return EPR_ERR_ARG; }
	*out = nullptr;
	return epr::ContextFree( [&]()
	{
		using namespace epr::iso;
		Clock::time_point t0 = Clock::now();
		const unsigned long long hashes[2][2] = { { IndexerHash( 3 , false ) , IndexerHash( 3 , true ) } , { IndexerHash( 4 , false ) , IndexerHash( 4 , true ) } };
		TreeData t = ReadTree( path , hashes );
		epr_iso *I = new epr_iso;
		try
		{
			I->dim = t.dim , I->depth = t.D , I->adapted = t.adapted , I->hash = t.hash , I->xform = t.xform;
			if( t.dim==3 ) Upload< 3 >( *I , t ) ; else Upload< 4 >( *I , t );
		}
		catch( ... ){ delete I ; throw; }
		I->seconds[0] = Since( t0 );
		*out = I;
	} );
}

const char *epr_iso_last_error( const epr_iso *iso ){ return iso ? iso->err.c_str() : epr::ThreadError().c_str(); }

int epr_iso_info( const epr_iso *iso , int *dim , int *depth , int *adapted , double xform[25] )
{
	if( !iso ) return EPR_ERR_ARG;
	if( dim ) *dim = iso->dim;
	if( depth ) *depth = iso->depth;
	if( adapted ) *adapted = iso->adapted ? 1 : 0;
	if( xform ) for( size_t i=0 ; i<iso->xform.size() ; i++ ) xform[i] = iso->xform[i];
	return EPR_OK;
}

int epr_iso_values( epr_iso *iso , double *out )
{
	return IsoGuard( iso , [&]()
	{
		long long N = 1 ; for( int k=0 ; k<iso->dim ; k++ ) N *= ( 1ll<<iso->depth ) + 1;
		if( !out ) throw std::string( "null output" );
		iso->V.download( out , 2*N , iso->st );
		HIPCHECK( hipStreamSynchronize( iso->st ) );
	} );
}

int epr_iso_level_set_nodes( epr_iso *iso , int all , size_t cap , unsigned *cells , size_t *count )
{
	return IsoGuard( iso , [&]()
	{
		std::vector< unsigned > c;
		if( iso->dim==3 ) epr::iso::Cells< 3 >( *iso , all!=0 , c ) ; else epr::iso::Cells< 4 >( *iso , all!=0 , c );
		const size_t n = c.size()/iso->dim;
		if( count ) *count = n;
		if( cells && cap ) memcpy( cells , c.data() , std::min( cap , n )*iso->dim*sizeof(unsigned) );
	} );
}

int epr_iso_extract( epr_iso *iso , int all , epr_mesh **out ){ return epr_iso_extract_trimmed( iso , all , 0. , out ); }

int epr_iso_extract_trimmed( epr_iso *iso , int all , double trimDensity , epr_mesh **out )
{
	if( !out ) return EPR_ERR_ARG;
	*out = nullptr;
	return IsoGuard( iso , [&](){ *out = iso->dim==3 ? epr::iso::Extract< 3 >( *iso , all!=0 , trimDensity ) : epr::iso::Extract< 4 >( *iso , all!=0 , trimDensity ); } );
}

int epr_iso_read_density( epr_iso *iso , const char *densityPath )
{
	if( !densityPath ) return EPR_ERR_ARG;
	return IsoGuard( iso , [&]()
	{
		std::unique_ptr< epr::iso::DensityTree > T( new epr::iso::DensityTree( epr::iso::ReadDensity( densityPath , iso->dim ) ) );
		iso->dChild.upload( T->child.data() , T->child.size() , iso->st );
		iso->dW.upload( T->w.data() , T->w.size() , iso->st );
		HIPCHECK( hipStreamSynchronize( iso->st ) );
		iso->density = std::move( T );
	} );
}

int epr_density_grid( const char *densityPath , int dim , int depth , const char *gridPath )
{
	if( !densityPath || !gridPath || ( dim!=3 && dim!=4 ) ){ epr::ThreadError() = "bad argument" ; return EPR_ERR_ARG; }
	return epr::ContextFree( [&](){ if( dim==3 ) epr::iso::DensityGrid< 3 >( densityPath , depth , gridPath ) ; else epr::iso::DensityGrid< 4 >( densityPath , depth , gridPath ); } );
}

int epr_iso_component_densities( const epr_iso *iso , size_t cap , double *out , size_t *count )
{
	if( !iso ) return EPR_ERR_ARG;
	if( count ) *count = iso->componentMax.size();
	if( out ) for( size_t i=0 ; i<std::min( cap , iso->componentMax.size() ) ; i++ ) out[i] = iso->componentMax[i];
	return EPR_OK;
}

int epr_iso_level_set( epr_iso *iso , int component , epr_mesh **out )
{
	if( !out || ( component!=0 && component!=1 ) ) return EPR_ERR_ARG;
	*out = nullptr;
	return IsoGuard( iso , [&](){ *out = iso->dim==3 ? epr::iso::LevelSet1< 3 >( *iso , component ) : epr::iso::LevelSet1< 4 >( *iso , component ); } );
}

int epr_iso_write_grid( epr_iso *iso , const char *header )
{
	if( !header ) return EPR_ERR_ARG;
	return IsoGuard( iso , [&](){ epr::iso::WriteGrid( *iso , header ); } );
}

double epr_iso_seconds( const epr_iso *iso , int which ){ return iso && which>=0 && which<5 ? iso->seconds[which] : 0.; }

unsigned long long epr_indexer_hash( int dim , int adapted ){ return epr::iso::IndexerHash( dim , adapted!=0 ); }

int epr_orient( int k , size_t nS , unsigned *simplices )
{
	if( ( k!=2 && k!=3 ) || ( nS && !simplices ) ){ epr::ThreadError() = "bad argument" ; return EPR_ERR_ARG; }
	return epr::ContextFree( [&]()
	{
		if( !nS ) return;
		size_t nV = 0;
		for( size_t i=0 ; i<nS*k ; i++ ) nV = std::max< size_t >( nV , (size_t)simplices[i]+1 );
		epr::DevBuf< unsigned > dS;
		dS.upload( simplices , nS*k , nullptr );
		std::vector< unsigned > S , other;
		std::vector< unsigned char > flipped;
		if( k==2 ) epr::iso::Orient< 1 >( dS.p , nS , nV , nullptr , false , S , flipped , other );
		else       epr::iso::Orient< 2 >( dS.p , nS , nV , nullptr , false , S , flipped , other );
		memcpy( simplices , S.data() , S.size()*sizeof(unsigned) );
	} );
}

void epr_iso_destroy( epr_iso *iso ){ delete iso; }

}
