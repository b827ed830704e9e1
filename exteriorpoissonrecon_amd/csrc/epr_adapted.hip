// Device construction of the adapted hierarchy (the default --aRadius path),
// Include/Hat.adaptive.inl:349-729: per depth, the function flags (active,
// padding), the element flags and the numbering of both.  The sets and orders
// are those of the host construction (AdaptedHierarchy::buildLevels,
// epr_adapted.h); here every step is a grid-parallel pass:
//
//  * active functions: the sample cells' ancestors marked on the function
//    grid, then Dim separable passes of a [-(K+1), K] OR filter (a function F
//    is active iff an ancestor A has F in A + [-K, K+1]^Dim);
//  * their depth-first order: a radix sort of Morton keys (coordinate 0 in
//    the lowest bit of each level);
//  * padding functions and elements in discovery order: the reference visits
//    the active functions in that order and each one's 3^Dim window in window
//    order, so a node's position is the smallest (rank * 3^Dim + window index)
//    over the visits that reach it -- a 64-bit atomicMin per visit, then a
//    radix sort of the keys;
//  * the colour lists of the active functions (grouped by colour, increasing
//    within a colour) for the list-driven sweeps.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>
#include <vector>

#include "epr_adapted_dev.h"

namespace epr
{
namespace
{
template< int Dim > constexpr int Pow3( void ){ int p = 1 ; for( int d=0 ; d<Dim ; d++ ) p *= 3 ; return p; }
template< int Dim > __device__ __forceinline__ void UnflatF( unsigned long long i , int n1 , int *F ){ for( int d=Dim-1 ; d>=0 ; d-- ){ F[d] = (int)( i % (unsigned)n1 ) ; i /= (unsigned)n1; } }
template< int Dim > __device__ __forceinline__ unsigned long long FlatF( const int *F , int n1 ){ unsigned long long i = 0 ; for( int d=0 ; d<Dim ; d++ ) i = i*n1 + F[d] ; return i; }
template< int Dim > __device__ __forceinline__ unsigned long long FlatE( const int *E , int res ){ unsigned long long i = 0 ; for( int d=Dim-1 ; d>=0 ; d-- ) i = i*res + E[d] ; return i; }

// g = 1 at the depth's ancestor (cell >> shift) of every sample cell, function layout (res+1)^Dim
template< int Dim >
__global__ void KMarkAnc( const int *cells , int ncell , int shift , int res , unsigned char *g )
{
	const int c = blockIdx.x*blockDim.x + threadIdx.x;
	if( c>=ncell ) return;
	int A[Dim];
	for( int d=0 ; d<Dim ; d++ ) A[d] = cells[ (size_t)c*Dim + d ] >> shift;
	g[ FlatF< Dim >( A , res+1 ) ] = 1;
}

// out[F] = OR of in[F + t e_axis] over t in [lo, hi] with 0 <= F_axis + t <= res
template< int Dim >
__global__ void KDilate( const unsigned char *in , unsigned char *out , int res , int axis , int lo , int hi , long long N )
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=N ) return;
	long long stride = 1;
	for( int d=Dim-1 ; d>axis ; d-- ) stride *= res+1;
	const int c = (int)( ( i / stride ) % ( res+1 ) );
	const int t0 = max( lo , -c ) , t1 = min( hi , res-c );
	unsigned char v = 0;
	for( int t=t0 ; t<=t1 ; t++ ) v |= in[ i + t*stride ];
	out[i] = v;
}

struct FlagIs
{
	const unsigned char *f;
	unsigned char v;
	__device__ __forceinline__ bool operator()( const unsigned int &i ) const { return f[i]==v; }
};
struct KeySet
{
	const unsigned long long *k;
	__device__ __forceinline__ bool operator()( const unsigned int &i ) const { return k[i]!=~0ULL; }
};

// Morton key of each listed function (d+1 bits per coordinate, coordinate 0 lowest)
template< int Dim >
__global__ void KMortonKeys( const unsigned int *sel , unsigned int n , int res , int bits , unsigned long long *keys )
{
	const unsigned int k = blockIdx.x*blockDim.x + threadIdx.x;
	if( k>=n ) return;
	int F[Dim];
	UnflatF< Dim >( sel[k] , res+1 , F );
	unsigned long long m = 0;
	for( int l=0 ; l<bits ; l++ ) for( int d=0 ; d<Dim ; d++ ) m |= (unsigned long long)( ( F[d]>>l ) & 1 ) << ( l*Dim + d );
	keys[k] = m;
}

// the first visit (rank, window position) that reaches each function outside the active set
template< int Dim >
__global__ void KPadKeys( const unsigned int *order , unsigned int nA , int res , const unsigned char *fm , unsigned long long *pkey )
{
	constexpr int W = Pow3< Dim >();
	const unsigned int r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=nA ) return;
	int F[Dim];
	UnflatF< Dim >( order[r] , res+1 , F );
	for( int w=0 ; w<W ; w++ )
	{
		int G[Dim];
		bool in = true;
		for( int d=0 , t=w , p=W/3 ; d<Dim ; d++ , p/=3 ){ G[d] = F[d] + ( t/p )%3 - 1 ; in = in && G[d]>=0 && G[d]<=res; }
		if( !in ) continue;
		const unsigned long long g = FlatF< Dim >( G , res+1 );
		if( !fm[g] ) atomicMin( pkey + g , (unsigned long long)r*W + w );
	}
}

// the first visit that reaches each element E <= F of an active function F
template< int Dim >
__global__ void KElemKeys( const unsigned int *order , unsigned int nA , int res , unsigned long long *ekey )
{
	constexpr int W = Pow3< Dim >();
	const unsigned int r = blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=nA ) return;
	int F[Dim];
	UnflatF< Dim >( order[r] , res+1 , F );
	for( int b=0 ; b<(1<<Dim) ; b++ )
	{
		int E[Dim] , w = 0;
		bool in = true;
		for( int d=0 ; d<Dim ; d++ ){ const int o = ( ( b>>(Dim-1-d) ) & 1 ) - 1 ; E[d] = F[d] + o ; in = in && E[d]>=0 && E[d]<res ; w = w*3 + o + 1; }
		if( in ) atomicMin( ekey + FlatE< Dim >( E , res ) , (unsigned long long)r*W + w );
	}
}

__global__ void KGatherKeys( const unsigned long long *key , const unsigned int *sel , unsigned int n , unsigned long long *out )
{
	const unsigned int k = blockIdx.x*blockDim.x + threadIdx.x;
	if( k<n ) out[k] = key[ sel[k] ];
}

__global__ void KSetFlags( const unsigned int *list , unsigned int n , unsigned char *flags , unsigned char v )
{
	const unsigned int k = blockIdx.x*blockDim.x + threadIdx.x;
	if( k<n ) flags[ list[k] ] = v;
}

// colour of each listed function (bit d = parity of coordinate d, ExteriorPoisson.h:161-169)
template< int Dim >
__global__ void KColourKeys( const unsigned int *sel , unsigned int n , int res , unsigned int *keys )
{
	const unsigned int k = blockIdx.x*blockDim.x + threadIdx.x;
	if( k>=n ) return;
	int F[Dim];
	UnflatF< Dim >( sel[k] , res+1 , F );
	unsigned int c = 0;
	for( int d=0 ; d<Dim ; d++ ) c |= (unsigned int)( F[d] & 1 ) << d;
	keys[k] = c;
}

// starts[c] = first position of colour c in the colour-sorted keys (untouched when absent)
__global__ void KColourStarts( const unsigned int *keys , unsigned int n , long long *starts )
{
	const unsigned int k = blockIdx.x*blockDim.x + threadIdx.x;
	if( k<n && ( k==0 || keys[k]!=keys[k-1] ) ) starts[ keys[k] ] = k;
}

int KeyBits( unsigned long long maxKey ){ int b = 1 ; while( b<64 && ( maxKey>>b ) ) b++ ; return b; }

// the construction's scratch with its stream
struct Scratch
{
	AdaptedScratch &b;
	hipStream_t st;
	template< typename F > void cub( F f )
	{
		size_t bytes = 0;
		HIPCHECK( f( nullptr , bytes ) );
		b.tmp.ensure( bytes );
		HIPCHECK( f( b.tmp.p , bytes ) );
	}
	unsigned int count( void )
	{
		HIPCHECK( hipMemcpyAsync( b.hCount , b.dn.p , sizeof(unsigned int) , hipMemcpyDeviceToHost , st ) );
		HIPCHECK( hipStreamSynchronize( st ) );
		return *b.hCount;
	}
	// increasing indices i < n with pred(i), into sel; returns the count
	template< typename Pred > unsigned int select( long long n , Pred pred )
	{
		b.sel.ensure( (size_t)n );
		hipcub::CountingInputIterator< unsigned int > it( 0 );
		cub( [&]( void *t , size_t &bytes ){ return hipcub::DeviceSelect::If( t , bytes , it , b.sel.p , b.dn.p , (int)n , pred , st ); } );
		return count();
	}
	// sorts (keys, sel[0..n)) by key into out
	void sortByKey( unsigned int n , int bits , unsigned int *out )
	{
		b.keys2.ensure( n );
		cub( [&]( void *t , size_t &bytes ){ return hipcub::DeviceRadixSort::SortPairs( t , bytes , b.keys.p , b.keys2.p , b.sel.p , out , (int)n , 0 , bits , st ); } );
	}
};
} // namespace

template< int Dim >
void BuildAdaptedDevice( const int *dCells , int ncell , int D , int K , hipStream_t st , const std::vector< AdaptedLevelDev * > &out , AdaptedScratch &scratch )
{
	constexpr int C = 1<<Dim , W = Pow3< Dim >();
	if( (int)out.size()<D+1 ) throw std::string( "BuildAdaptedDevice: too few output levels" );
	Scratch s{ scratch , st };
	scratch.dn.ensure( 1 );
	scratch.dstarts.ensure( C );
	if( !scratch.hCount ) HIPCHECK( hipHostMalloc( (void **)&scratch.hCount , sizeof(unsigned int) ) );
	if( !scratch.hStarts ) HIPCHECK( hipHostMalloc( (void **)&scratch.hStarts , 17*sizeof(long long) ) );
	for( int d=0 ; d<=D ; d++ )
	{
		AdaptedLevelDev &L = *out[d];
		const int res = 1<<d;
		long long N = 1 , Ne = 1;
		for( int k=0 ; k<Dim ; k++ ) N *= res+1 , Ne *= res;
		// ---- active flags: ancestors, then the separable box filter
		s.b.ga.ensure( N ) , s.b.gb.ensure( N );
		HIPCHECK( hipMemsetAsync( s.b.ga.p , 0 , N , st ) );
		KMarkAnc< Dim ><<< Blocks( ncell , 256 ) , 256 , 0 , st >>>( dCells , ncell , D-d , res , s.b.ga.p );
		L.fm.ensure( N );
		for( int a=0 ; a<Dim ; a++ )
		{
			const unsigned char *in = ( a&1 ) ? s.b.gb.p : s.b.ga.p;
			unsigned char *o = a==Dim-1 ? L.fm.p : ( ( a&1 ) ? s.b.ga.p : s.b.gb.p );
			KDilate< Dim ><<< Blocks( N , 256 ) , 256 , 0 , st >>>( in , o , res , a , -(K+1) , K , N );
		}
		HIPCHECK( hipGetLastError() );
		// ---- active functions in depth-first (Morton) order
		const unsigned int nA = s.select( N , FlagIs{ L.fm.p , 1 } );
		s.b.keys.ensure( nA );
		KMortonKeys< Dim ><<< Blocks( nA , 256 ) , 256 , 0 , st >>>( s.b.sel.p , nA , res , d+1 , s.b.keys.p );
		s.b.vals.ensure( nA );
		s.sortByKey( nA , ( d+1 )*Dim , s.b.vals.p );     // vals: the active functions in the hierarchy's order
		// ---- colour lists (from sel: increasing dense index; the radix sort is stable)
		L.act.ensure( nA );
		s.b.ckeys.ensure( nA ) , s.b.ckeys2.ensure( nA );
		KColourKeys< Dim ><<< Blocks( nA , 256 ) , 256 , 0 , st >>>( s.b.sel.p , nA , res , s.b.ckeys.p );
		s.cub( [&]( void *t , size_t &b ){ return hipcub::DeviceRadixSort::SortPairs( t , b , s.b.ckeys.p , s.b.ckeys2.p , s.b.sel.p , L.act.p , (int)nA , 0 , Dim , st ); } );
		{
			const long long *starts = scratch.hStarts;
			HIPCHECK( hipMemsetAsync( s.b.dstarts.p , 0xFF , C*sizeof(long long) , st ) );   // -1: colour absent
			KColourStarts<<< Blocks( nA , 256 ) , 256 , 0 , st >>>( s.b.ckeys2.p , nA , s.b.dstarts.p );
			HIPCHECK( hipMemcpyAsync( scratch.hStarts , s.b.dstarts.p , C*sizeof(long long) , hipMemcpyDeviceToHost , st ) );
			HIPCHECK( hipStreamSynchronize( st ) );
			L.actStart[C] = nA;
			for( int c=C-1 ; c>=0 ; c-- ) L.actStart[c] = starts[c]>=0 ? starts[c] : L.actStart[c+1];
		}
		// ---- padding functions by discovery
		s.b.pkey.ensure( std::max( N , Ne ) );
		HIPCHECK( hipMemsetAsync( s.b.pkey.p , 0xFF , N*sizeof(unsigned long long) , st ) );
		KPadKeys< Dim ><<< Blocks( nA , 128 ) , 128 , 0 , st >>>( s.b.vals.p , nA , res , L.fm.p , s.b.pkey.p );
		const unsigned int nP = s.select( N , KeySet{ s.b.pkey.p } );
		L.nActive = nA , L.nF = (size_t)nA + nP;
		L.forder.ensure( L.nF );
		HIPCHECK( hipMemcpyAsync( L.forder.p , s.b.vals.p , (size_t)nA*sizeof(unsigned int) , hipMemcpyDeviceToDevice , st ) );
		if( nP )
		{
			s.b.keys.ensure( nP );
			KGatherKeys<<< Blocks( nP , 256 ) , 256 , 0 , st >>>( s.b.pkey.p , s.b.sel.p , nP , s.b.keys.p );
			s.sortByKey( nP , KeyBits( (unsigned long long)nA*W ) , L.forder.p + nA );
			KSetFlags<<< Blocks( nP , 256 ) , 256 , 0 , st >>>( L.forder.p + nA , nP , L.fm.p , 2 );
		}
		// ---- elements by discovery
		HIPCHECK( hipMemsetAsync( s.b.pkey.p , 0xFF , Ne*sizeof(unsigned long long) , st ) );
		KElemKeys< Dim ><<< Blocks( nA , 128 ) , 128 , 0 , st >>>( s.b.vals.p , nA , res , s.b.pkey.p );
		const unsigned int nE = s.select( Ne , KeySet{ s.b.pkey.p } );
		L.nE = nE;
		L.eorder.ensure( nE );
		L.em.ensure( Ne );
		HIPCHECK( hipMemsetAsync( L.em.p , 0 , Ne , st ) );
		s.b.keys.ensure( nE );
		KGatherKeys<<< Blocks( nE , 256 ) , 256 , 0 , st >>>( s.b.pkey.p , s.b.sel.p , nE , s.b.keys.p );
		s.sortByKey( nE , KeyBits( (unsigned long long)nA*W ) , L.eorder.p );
		KSetFlags<<< Blocks( nE , 256 ) , 256 , 0 , st >>>( L.eorder.p , nE , L.em.p , 1 );
		HIPCHECK( hipGetLastError() );
	}
	HIPCHECK( hipStreamSynchronize( st ) );
}

template void BuildAdaptedDevice< 3 >( const int * , int , int , int , hipStream_t , const std::vector< AdaptedLevelDev * > & , AdaptedScratch & );
template void BuildAdaptedDevice< 4 >( const int * , int , int , int , hipStream_t , const std::vector< AdaptedLevelDev * > & , AdaptedScratch & );
} // namespace epr
