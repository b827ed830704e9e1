// C ABI and host orchestration of the HIP ExteriorPoissonRecon solve path.
//
// Host work is limited to what the reference does on its input before any
// grid exists: the unit-cube map (ExteriorPoisson.h:100-123) and the ordering
// of samples into occupied cells (OrderedSamples.inl:32-85, the same
// std::sort so cells and their sample order are identical), plus the O(1)
// stencil tables.  Every grid-sized computation runs on the device.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/epr.h"
#include "epr_adapted.h"
#include "epr_adapted_dev.h"
#include "epr_batch.h"
#include "epr_dev.h"
#include "epr_hat.h"
#include "epr_kernels.h"
#include "epr_misha.h"
#include "epr_rand.h"
#include "epr_unet.h"

namespace epr
{
// The unit-cube map (ExteriorPoisson.h:100-123, with the matrix products of
// Misha/Geometry.inl:97-107, 410-427) and the ordered sampler's cells
// (OrderedSamples.inl:32-85: clamped floor binning, lexicographic std::sort,
// grouping) of n world positions.
template< int Dim >
struct SampleCells
{
	std::vector< double > pos;                  // unit-cube positions
	double unitToWorld[Dim+1][Dim+1];
	std::vector< int > cellIdx , cellStart , cellSamples;
	int ncell = 0;

	void bin( size_t n , const double *wpos , double scale , int D , double jitter=0 )
	{
		if( !n ) throw std::string( "Expected some samples: 0" );
		if( n>=( 1ULL<<32 ) ) throw std::string( "more than 2^32 - 1 samples" );
		for( size_t i=0 ; i<n*Dim ; i++ ) if( !std::isfinite( wpos[i] ) ) throw std::string( "non-finite sample position" );
		double bb0[Dim] , bb1[Dim];
		for( int d=0 ; d<Dim ; d++ ) bb0[d] = bb1[d] = wpos[d];
		for( size_t i=1 ; i<n ; i++ ) for( int d=0 ; d<Dim ; d++ ){ bb0[d] = std::min( bb0[d] , wpos[i*Dim+d] ) ; bb1[d] = std::max( bb1[d] , wpos[i*Dim+d] ); }
		double sc = bb1[0]-bb0[0];
		for( int d=1 ; d<Dim ; d++ ) sc = std::max( sc , bb1[d]-bb0[d] );
		double t1[Dim+1][Dim+1] , s[Dim+1][Dim+1] , t2[Dim+1][Dim+1] , ts[Dim+1][Dim+1] , M[Dim+1][Dim+1];
		for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ) t1[i][j] = s[i][j] = t2[i][j] = ( i==j ? 1. : 0. );
		for( int d=0 ; d<Dim ; d++ ) t2[Dim][d] = -( bb0[d] + bb1[d] ) / 2.;
		for( int d=0 ; d<Dim ; d++ ) s[d][d] = 1. / ( sc * scale );
		for( int d=0 ; d<Dim ; d++ ) t1[Dim][d] = 0.5;
		auto Mul = []( const double A[Dim+1][Dim+1] , const double B[Dim+1][Dim+1] , double O[Dim+1][Dim+1] )
		{
			for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ){ O[i][j] = 0 ; for( int k=0 ; k<=Dim ; k++ ) O[i][j] += B[i][k]*A[k][j]; }
		};
		Mul( t1 , s , ts );
		Mul( ts , t2 , M );
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
		// the unit-cube-to-world map: the reference's Gauss-Jordan inverse of the
		// world-to-unit-cube map (ExteriorPoissonRecon.cpp:171), bit for bit
		sample::Mat< Dim+1 > W , U2W;
		memcpy( W.c , M , sizeof(M) );
		U2W = W.inverse();
		memcpy( unitToWorld , U2W.c , sizeof(unitToWorld) );
		// --jitter (ExteriorPoissonRecon.cpp:174-180): a random translation of the
		// unit-cube positions drawn before the seed is set, i.e. from glibc's
		// unseeded sequence (srand(1)); the unit-cube-to-world map absorbs its
		// (Gauss-Jordan) inverse
		if( jitter>0 )
		{
			double T[Dim+1][Dim+1] , U[Dim+1][Dim+1];
			for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ) T[i][j] = ( i==j ? 1. : 0. );
			GlibcRand g( 1 );
			for( int d=0 ; d<Dim ; d++ ) T[Dim][d] = ( g.random01() * 2. - 1. ) * jitter;
			sample::Mat< Dim+1 > Tm;
			memcpy( Tm.c , T , sizeof(T) );
			const sample::Mat< Dim+1 > TiM = Tm.inverse();
			double Ti[Dim+1][Dim+1];
			memcpy( Ti , TiM.c , sizeof(Ti) );
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
			Mul( unitToWorld , Ti , U );
			memcpy( unitToWorld , U , sizeof(U) );
		}
		int res = 1<<D;
		// the cell's multi-index packed into one key, coordinate 0 most significant
		// (13 bits a coordinate: res <= 2^12), so a key comparison is the
		// lexicographic comparison of OrderedSamples.inl and std::sort makes the
		// same comparisons, hence the same permutation, as on the multi-indices
		constexpr int KB = 13;
		struct IS { unsigned long long key ; unsigned int idx ; bool operator<( const IS &o ) const { return key<o.key; } };
		std::vector< IS > v( n );
		for( size_t i=0 ; i<n ; i++ )
		{
			unsigned long long k = 0;
			for( int d=0 ; d<Dim ; d++ ) k = ( k<<KB ) | (unsigned long long)std::max< int >( 0 , std::min< int >( (int)floor( pos[i*Dim+d] * res ) , res-1 ) );
			v[i].key = k , v[i].idx = (unsigned int)i;
		}
		std::sort( v.begin() , v.end() );
		cellIdx.clear() , cellStart.clear() , cellSamples.clear();
		cellSamples.reserve( n );
		size_t start = 0;
		while( start<v.size() )
		{
			size_t end = start;
			while( end<v.size() && v[end].key==v[start].key ) end++;
			cellStart.push_back( (int)cellSamples.size() );
			for( int d=0 ; d<Dim ; d++ ) cellIdx.push_back( (int)( ( v[start].key>>( KB*( Dim-1-d ) ) ) & ( ( 1ULL<<KB )-1 ) ) );
			for( size_t i=start ; i<end ; i++ ) cellSamples.push_back( (int)v[i].idx );
			start = end;
		}
		cellStart.push_back( (int)cellSamples.size() );
		ncell = (int)cellStart.size()-1;
	}
};

// ---- the learned initializer's input: the target field on a coarse level
// mean of the finest per-element field S over each element of level L
// (element index with coordinate 0 fastest, as S)
template< int Dim >
__global__ void __launch_bounds__( 256 ) KCoarseFieldMean( const double *S , int D , int L , double *out )
{
	constexpr int CD = DimC< Dim >::CD;
	const int rl = 1<<L , rf = 1<<D , B = 1<<( D-L );
	long long ne = 1 , cnt = 1;
	for( int d=0 ; d<Dim ; d++ ) ne *= rl , cnt *= B;
	const long long e = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( e>=ne ) return;
	int E[Dim];
	{ long long t = e ; for( int d=0 ; d<Dim ; d++ ){ E[d] = (int)( t % rl ) ; t /= rl; } }
	double s[CD];
	for( int c=0 ; c<CD ; c++ ) s[c] = 0;
	for( long long k=0 ; k<cnt ; k++ )
	{
		long long u = k , fi = 0 , mul = 1;
		for( int d=0 ; d<Dim ; d++ ){ const int o = (int)( u % B ) ; u /= B ; fi += ( (long long)E[d]*B + o )*mul ; mul *= rf; }
		for( int c=0 ; c<CD ; c++ ) s[c] += S[ fi*CD + c ];
	}
	for( int c=0 ; c<CD ; c++ ) out[ e*CD + c ] = s[c] / (double)cnt;
}
// per node of level L (function index, last coordinate fastest): the mean of
// its incident level-L elements' means
template< int Dim >
__global__ void __launch_bounds__( 256 ) KNodeFieldMean( const double *Ec , int L , double *out )
{
	constexpr int CD = DimC< Dim >::CD;
	const int rl = 1<<L , side = rl+1;
	long long N = 1;
	for( int d=0 ; d<Dim ; d++ ) N *= side;
	const long long n = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( n>=N ) return;
	int F[Dim];
	{ long long t = n ; for( int d=Dim-1 ; d>=0 ; d-- ){ F[d] = (int)( t % side ) ; t /= side; } }
	double s[CD];
	for( int c=0 ; c<CD ; c++ ) s[c] = 0;
	int cnt = 0;
	for( int m=0 ; m<(1<<Dim) ; m++ )
	{
		long long e = 0 , mul = 1;
		bool ok = true;
		for( int d=0 ; d<Dim ; d++ ){ const int q = F[d] - 1 + ( ( m>>d ) & 1 ) ; ok = ok && q>=0 && q<rl ; e += (long long)q*mul ; mul *= rl; }
		if( !ok ) continue;
		cnt++;
		for( int c=0 ; c<CD ; c++ ) s[c] += Ec[ e*CD + c ];
	}
	for( int c=0 ; c<CD ; c++ ) out[ n*CD + c ] = cnt ? s[c]/cnt : 0.;
}
// out[0] = sum of squares of v (one workgroup, fixed order)
__global__ void __launch_bounds__( 256 ) KSumSquares( const double *v , long long n , double *out )
{
	__shared__ double sh[256];
	double s = 0;
	for( long long i=threadIdx.x ; i<n ; i+=256 ) s += v[i]*v[i];
	sh[threadIdx.x] = s;
	__syncthreads();
	for( int h=128 ; h>0 ; h>>=1 ){ if( (int)threadIdx.x<h ) sh[threadIdx.x] += sh[threadIdx.x+h] ; __syncthreads(); }
	if( threadIdx.x==0 ) out[0] = sh[0];
}
// the network input: v / rms(v), bf16 rows padded to 16 channels, and as floats
__global__ void __launch_bounds__( 256 ) KFeatureRows( const double *v , long long N , int CD , const double *sumsq , unet::bf16 *rows , float *flt )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( k>=N*unet::CT ) return;
	const long long n = k / unet::CT;
	const int c = (int)( k % unet::CT );
	const double rms = sqrt( sumsq[0] / ( (double)N*CD ) );
	const double f = c<CD ? ( rms>0 ? v[ n*CD + c ]/rms : 0. ) : 0.;
	rows[k] = unet::ToBf16( (float)f );
	if( flt && c<CD ) flt[ n*CD + c ] = (float)f;
}
// x, y of a level from the network output (channels 0, 1); functions outside
// an adapted hierarchy stay zero
__global__ void __launch_bounds__( 256 ) KCoeffsFromNet( const float *yNet , long long N , const unsigned char *fmask , double *x , double *y )
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=N ) return;
	const bool on = !fmask || fmask[i]!=0;
	x[i] = on ? (double)yNet[ i*unet::CT ] : 0.;
	y[i] = on ? (double)yNet[ i*unet::CT + 1 ] : 0.;
}

__global__ void __launch_bounds__( 256 ) KNarrow( const double *a , float *b , long long n ){ const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x ; if( i<n ) b[i] = (float)a[i]; }
__global__ void __launch_bounds__( 256 ) KWiden( const float *a , double *b , long long n ){ const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x ; if( i<n ) b[i] = (double)a[i]; }

// Device scratch of the allocation-free energies (epr_energies_noalloc): the
// per-workgroup partial sums of KEnergy and the four summed terms.  One copy
// per device, held until the process ends; every call on a device shares it,
// so calls that overlap on one device race.  Host code never passes the
// array's name (a host-side stand-in) to a kernel or a copy: it uses the
// device address that hipGetSymbolAddress returns, looked up once per device.
constexpr size_t NoAllocScratchDoubles = size_t( 1 ) << 19;   // 4 MiB
constexpr long long NoAllocMaxBlocks = (long long)( NoAllocScratchDoubles - 4 ) / 4;
static __device__ __align__( 256 ) double scratch[ NoAllocScratchDoubles ];
// A device-side reference to the array.  Without one the device compiler
// drops the (internal-linkage, otherwise device-unused) array from the code
// object, and the symbol the host side registers for hipGetSymbolAddress does
// not exist on the device.  Never launched.
__global__ void KKeepNoAllocScratch( void ){ scratch[0] = 0.; }
inline double *NoAllocScratch( void )
{
	constexpr int MaxDevices = 64;
	static std::once_flag once[MaxDevices];
	static double *ptr[MaxDevices] = {};
	static hipError_t status[MaxDevices] = {};
	int dev = 0;
	HIPCHECK( hipGetDevice( &dev ) );
	if( dev<0 || dev>=MaxDevices ) throw std::string( "device index out of range" );
	std::call_once( once[dev] , [&]{ status[dev] = hipGetSymbolAddress( (void **)&ptr[dev] , HIP_SYMBOL( scratch ) ); } );
	if( status[dev]!=hipSuccess || !ptr[dev] ) throw HipError{ std::string( "hipGetSymbolAddress( scratch ): " ) + hipGetErrorString( status[dev] ) };
	return ptr[dev];
}

// The kernels launched with more than 64 KB of dynamic LDS get their limit
// raised once per device, before any context of the process launches them:
// hipFuncSetAttribute then never runs concurrently with launches of the same
// kernel from the batch API's build and solve threads.
template< int Dim > void SetLdsAttributesDim( int bytes )
{
	auto set = [&]( const void *f ){ HIPCHECK( hipFuncSetAttribute( f , hipFuncAttributeMaxDynamicSharedMemorySize , bytes ) ); };
	set( (const void *)KEstSplatLds< Dim , 0 > ) , set( (const void *)KEstSplatLds< Dim , 1 > ) , set( (const void *)KEstSplatLds< Dim , 2 > );
	set( (const void *)KEstSplatLds< Dim , 3 > ) , set( (const void *)KEstSplatLds< Dim , 4 > );
	set( (const void *)KGsTiny< Dim , true > ) , set( (const void *)KGsTiny< Dim , false > ) , set( (const void *)KGsSmallLds< Dim > );
	set( (const void *)KGsTinyBatch< Dim , true > ) , set( (const void *)KGsTinyBatch< Dim , false > ) , set( (const void *)KGsSmallLdsBatch< Dim > );
	if constexpr( Dim==3 ) set( (const void *)KGsBlock ) , set( (const void *)KGsBlockBatch );
}
inline void SetLdsAttributes( void )
{
	constexpr int MaxDevices = 64;
	static std::once_flag once[MaxDevices];
	static std::exception_ptr failed[MaxDevices];
	int dev = 0;
	HIPCHECK( hipGetDevice( &dev ) );
	if( dev<0 || dev>=MaxDevices ) throw std::string( "device index out of range" );
	std::call_once( once[dev] , [&]{ try { SetLdsAttributesDim< 3 >( 160*1024 ) ; SetLdsAttributesDim< 4 >( 160*1024 ); } catch( ... ){ failed[dev] = std::current_exception(); } } );
	if( failed[dev] ) std::rethrow_exception( failed[dev] );
}

// The energy terms of a level: KEnergy's partial-sum rows (the window terms),
// then KEnergyWedge's (the wedge term), summed in a fixed order into out[4]
// (W, y^T B x, R-form, K-form): EnergySumBlocks contiguous row ranges first
// (KSumRowsPart, their sums after the rows), then those sums.  EnergyRows:
// the rows partials must hold.
// An odd multiplier near n / golden ratio that is coprime with n: k -> (k m) mod n
// is then a permutation of [0, n) whose consecutive values lie far apart.
inline unsigned int CoprimeStride( unsigned int n )
{
	if( n<=2 ) return 1;
	unsigned int m = (unsigned int)( 0.6180339887*n ) | 1u;
	while( std::gcd( m , n )!=1 ) m += 2;
	return m;
}

constexpr int EnergySumBlocks = 256;
template< int Dim > long long EnergyRows( long long N , int res )
{
	constexpr int C = DimC< Dim >::C , NP = Dim*( Dim-1 )/2;
	long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= res;
	return (long long)Blocks( N*C , 256 ) + (long long)Blocks( ne*NP , 256 ) + EnergySumBlocks;
}
template< int Dim > void LaunchEnergy( const LevelView< Dim > &v , double *partials , double *out , hipStream_t st )
{
	constexpr int C = DimC< Dim >::C , NP = Dim*( Dim-1 )/2;
	long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= v.res;
	static const bool perElement = !( getenv( "EPR_WEDGE_ELEM" ) && atoi( getenv( "EPR_WEDGE_ELEM" ) )==0 );   // 0: one thread per (element, pair) (for timing)
	const unsigned nbW = Blocks( v.N*C , 256 ) , nbE = perElement ? Blocks( ne , 256 ) : Blocks( ne*NP , 256 );
	if( v.N*DimC< Dim >::W<( 1LL<<31 ) ) KEnergy< Dim , int ><<< nbW , 256 , 0 , st >>>( v , partials );
	else KEnergy< Dim ><<< nbW , 256 , 0 , st >>>( v , partials );
	if( perElement )
	{
		if( ne<( 1LL<<31 ) && v.N<( 1LL<<31 ) ) KEnergyWedgeElem< Dim , int ><<< nbE , 256 , 0 , st >>>( v , partials + 4*(size_t)nbW );
		else KEnergyWedgeElem< Dim ><<< nbE , 256 , 0 , st >>>( v , partials + 4*(size_t)nbW );
	}
	else if( ne*NP<( 1LL<<31 ) && v.N<( 1LL<<31 ) ) KEnergyWedge< Dim , int ><<< nbE , 256 , 0 , st >>>( v , partials + 4*(size_t)nbW );
	else KEnergyWedge< Dim ><<< nbE , 256 , 0 , st >>>( v , partials + 4*(size_t)nbW );
	double *sums = partials + 4*( (size_t)nbW + nbE );
	KSumRowsPart<4><<< EnergySumBlocks , 256 , 0 , st >>>( partials , (long long)nbW + nbE , sums );
	KSumRows<4><<< 1 , 256 , 0 , st >>>( sums , EnergySumBlocks , out );
}

// ---- tuned sweep plans (epr_autotune, epr_autotune_batch)
// Per device, dimension, level resolution, single/batched solve and precision:
// the sweep kernel of the level (kind, one of Ctx::GS_*; -1 keeps the built-in
// rule) and the workgroup size of its colour launches (threads: 64, 128 or
// 256, i.e. 64/2^Dim .. 256/2^Dim nodes per workgroup; 0 keeps 256).  Entries
// are written by the autotuners from timings on the device, or restored from
// their text form; contexts consult the table whenever they plan a level.
struct TunePick { int kind = -1 , threads = 0; double us = 0; };
struct TuneTable
{
	using Key = std::tuple< int , int , int , int , int >;   // device, dim, res, batched, fp32
	std::mutex m;
	std::map< Key , TunePick > picks;
	static TuneTable &Get( void ){ static TuneTable *t = new TuneTable ; return *t; }
	static int Device( void ){ int d = 0 ; HIPCHECK( hipGetDevice( &d ) ) ; return d; }
	TunePick find( int dim , int res , bool batched , bool fp32 )
	{
		const Key k{ Device() , dim , res , batched ? 1 : 0 , fp32 ? 1 : 0 };
		std::lock_guard< std::mutex > lock( m );
		auto it = picks.find( k );
		return it==picks.end() ? TunePick() : it->second;
	}
	void set( int dim , int res , bool batched , bool fp32 , TunePick p )
	{
		const Key k{ Device() , dim , res , batched ? 1 : 0 , fp32 ? 1 : 0 };
		std::lock_guard< std::mutex > lock( m );
		picks[k] = p;
	}
	// one line per entry of the current device: "dim res batched fp32 kind threads us"
	std::string text( void )
	{
		const int dev = Device();
		std::lock_guard< std::mutex > lock( m );
		std::string s;
		char line[160];
		for( const auto &e : picks ) if( std::get< 0 >( e.first )==dev )
		{
			snprintf( line , sizeof(line) , "%d %d %d %d %d %d %.3f\n" , std::get< 1 >( e.first ) , std::get< 2 >( e.first ) , std::get< 3 >( e.first ) , std::get< 4 >( e.first ) , e.second.kind , e.second.threads , e.second.us );
			s += line;
		}
		return s;
	}
	// replaces the current device's entries by those of a text of text()'s form (empty: none)
	void parse( const char *txt )
	{
		const int dev = Device();
		std::map< Key , TunePick > add;
		if( txt ) for( const char *p=txt ; *p ; )
		{
			const char *e = strchr( p , '\n' );
			std::string line( p , e ? (size_t)( e-p ) : strlen( p ) );
			p = e ? e+1 : p+line.size();
			if( line.find_first_not_of( " \t\r" )==std::string::npos ) continue;
			int dim , res , b , f , kind , th;
			double us = 0;
			if( sscanf( line.c_str() , "%d %d %d %d %d %d %lf" , &dim , &res , &b , &f , &kind , &th , &us )<6 ) throw std::string( "bad tuning line: " + line );
			if( ( dim!=3 && dim!=4 ) || res<1 || ( b!=0 && b!=1 ) || ( f!=0 && f!=1 ) || kind<-1 || kind>5 || ( th!=0 && th!=64 && th!=128 && th!=256 ) ) throw std::string( "bad tuning line: " + line );
			add[ Key{ dev , dim , res , b , f } ] = TunePick{ kind , th , us };
		}
		std::lock_guard< std::mutex > lock( m );
		for( auto it=picks.begin() ; it!=picks.end() ; ) if( std::get< 0 >( it->first )==dev ) it = picks.erase( it ) ; else ++it;
		for( const auto &e : add ) picks[ e.first ] = e.second;
	}
};

struct CtxBase
{
	virtual ~CtxBase(){}
	virtual void setSamples( size_t n , const double *pos , const double *skew ) = 0;
	virtual void build( void ) = 0;
	virtual void init( int level , const double *x0 , const double *y0 ) = 0;
	virtual void gs( int level , int iters ) = 0;
	virtual void gsColour( int level , int colour ) = 0;
	virtual void timeGs( int level , int iters , double *ms , int *launches ) = 0;
	virtual void timeGsF32( int level , int iters , double *ms , int *launches ) = 0;
	virtual void prolongAdd( int level ) = 0;
	virtual void energies( int level , double out[2] ) = 0;
	virtual void energiesNoAlloc( int level , double out[2] ) = 0;
	virtual void solve( double out[2] ) = 0;
	virtual void learnedFeatures( int level , float *out ) = 0;
	virtual void learnedInit( int level , unet::Net *net , double info[3] ) = 0;
	virtual void solveLearned( unet::Net *net , double out[2] ) = 0;
	virtual void solveSchedule( unet::Net *net , const int *iters , double out[2] ) = 0;
	virtual size_t numFunctions( int level ) const = 0;
	virtual void getCoeffs( int level , double *x , double *y ) = 0;
	virtual void setCoeffs( int level , const double *x , const double *y ) = 0;
	virtual void qfits( int level , double *q ) = 0;
	virtual void apply( int level , char which , const double *x , double *y ) = 0;
	virtual size_t getArray( const std::string &name , double *out , size_t cap ) = 0;
	virtual void writeTree( const char *path , unsigned long long hash ) = 0;
	virtual void writeGrid( const char *path ) = 0;
	virtual void writeDensity( const char *path ) = 0;
	virtual std::string autotune( int flags ) = 0;
	virtual void timeOp( const std::string &op , int level , int reps , double *ms , double *bytes ) = 0;
	double seconds[2] = { 0 , 0 };
};

// one-workgroup sweep kernels are tried on levels of at most this many nodes
constexpr long long TuneOneGroupMax = 5000;
// the workgroup sizes tried for colour launches
constexpr int TuneThreads[3] = { 64 , 128 , 256 };
// times f() on stream st in milliseconds
template< typename F > double TimeOnStream( hipStream_t st , F f )
{
	hipEvent_t e0 = nullptr , e1 = nullptr;
	HIPCHECK( hipEventCreate( &e0 ) );
	float t = 0;
	try
	{
		HIPCHECK( hipEventCreate( &e1 ) );
		HIPCHECK( hipEventRecord( e0 , st ) );
		f();
		HIPCHECK( hipEventRecord( e1 , st ) );
		HIPCHECK( hipEventSynchronize( e1 ) );
		HIPCHECK( hipEventElapsedTime( &t , e0 , e1 ) );
	}
	catch( ... ){ (void)hipEventDestroy( e0 ) ; if( e1 ) (void)hipEventDestroy( e1 ) ; throw; }
	(void)hipEventDestroy( e0 ) , (void)hipEventDestroy( e1 );
	return t;
}
static const char *GsKindName( int k )
{
	static const char *names[] = { "tiny" , "small_lds" , "small" , "phases" , "wave" , "block" };
	return k>=0 && k<6 ? names[k] : "default";
}

template< int Dim >
struct Ctx : public CtxBase
{
	static constexpr int W = DimC< Dim >::W , C = DimC< Dim >::C , CD = DimC< Dim >::CD , NV = DimC< Dim >::NDNI , CoDim = 2;
	// levels up to this many nodes (and not tiny) run all their iterations in one
	// workgroup; larger ones use graph-replayed colour launches (EPR_SMALL_LEVEL
	// overrides, for tuning).  Measured on MI355X: the one-workgroup kernel wins
	// for 125 nodes (3D level 2), colour launches for 729 (3D) and 625 (4D).
	long long SmallLevel = getenv( "EPR_SMALL_LEVEL" ) ? atoll( getenv( "EPR_SMALL_LEVEL" ) ) : ( Dim==3 ? 200 : 100 );
	static constexpr size_t LdsLimit = 160*1024;   // gfx950 LDS per workgroup
	// levels with at most this many elements use KGsTiny (EPR_TINY_ELEMENTS overrides, for tuning)
	long long tinyElements = getenv( "EPR_TINY_ELEMENTS" ) ? atoll( getenv( "EPR_TINY_ELEMENTS" ) ) : ( Dim==3 ? 8 : 16 );

	epr_params prm;
	int D;
	hipStream_t stream = nullptr;

	// samples (host, unit cube) and their cell order
	size_t n = 0;
	std::vector< double > pos , skewPacked;
	double unitToWorld[Dim+1][Dim+1];
	std::vector< int > cellIdx , cellStart , cellSamples;
	int ncell = 0;
	bool built = false , solved = false;

	// One depth: the dense storage and, for the adapted hierarchy (empty for
	// the dense grid), the AdaptedLevelDev sets -- function flags (1 active,
	// 2 padding), element flags, the numbering, the active nodes by colour --
	// which the device builder refills in place on every build.
	struct Level : public AdaptedLevelDev
	{
		int res;
		long long N;
		long long stride[Dim];
		ColLayout< Dim > lay;              // colour-grouped column order of R and B
		DevBuf< double > x , y , R , B , K;
		hipGraphExec_t graph = nullptr;    // captured Gauss-Seidel launches of this level
		int graphIters = 0 , graphThreads = 0;
		// --fp32: float copies of x, y, R, B, K (R, B, K narrowed once per build)
		// and the captured float launches
		DevBuf< float > fx , fy , fR , fB , fK;
		bool narrowed = false;
		bool fusedR = false , fusedB = false;   // fR, fB written by this build's assembly
		hipGraphExec_t graph32 = nullptr;
		int graph32Iters = 0 , graph32Threads = 0;
		const float *graph32P[5] = { nullptr , nullptr , nullptr , nullptr , nullptr };
		const double *graphX = nullptr , *graphR = nullptr;
		const unsigned char *graphF = nullptr;
		const unsigned int *graphAct = nullptr;
		long long graphActStart[(1<<Dim)+1] = {};
		Level( void ){}
		Level( Level &&o ) noexcept : AdaptedLevelDev( std::move( o ) ) , res( o.res ) , N( o.N ) , lay( o.lay ) , x( std::move( o.x ) ) , y( std::move( o.y ) ) , R( std::move( o.R ) ) , B( std::move( o.B ) ) , K( std::move( o.K ) ) , graph( o.graph ) , graphIters( o.graphIters ) , graphThreads( o.graphThreads ) ,
		                              fx( std::move( o.fx ) ) , fy( std::move( o.fy ) ) , fR( std::move( o.fR ) ) , fB( std::move( o.fB ) ) , fK( std::move( o.fK ) ) , narrowed( o.narrowed ) , graph32( o.graph32 ) , graph32Iters( o.graph32Iters ) , graph32Threads( o.graph32Threads )
		{
			for( int d=0 ; d<Dim ; d++ ) stride[d] = o.stride[d];
			for( int k=0 ; k<5 ; k++ ) graph32P[k] = o.graph32P[k];
			graphX = o.graphX , graphR = o.graphR , graphF = o.graphF , graphAct = o.graphAct;
			for( int c=0 ; c<=(1<<Dim) ; c++ ) graphActStart[c] = o.graphActStart[c];
			o.graph = nullptr , o.graph32 = nullptr;
		}
		~Level(){ if( graph ) (void)hipGraphExecDestroy( graph ) ; if( graph32 ) (void)hipGraphExecDestroy( graph32 ); }
		// drops the captured launches when the buffers they point at move or the
		// active lists they cover change
		void checkGraph( void )
		{
			bool same = graphX==x.p && graphR==R.p && graphF==fm.p && graphAct==act.p;
			for( int c=0 ; c<=(1<<Dim) ; c++ ) same = same && graphActStart[c]==actStart[c];
			if( graph && !same ){ (void)hipGraphExecDestroy( graph ) ; graph = nullptr; }   // the caller has drained the stream
		}
	};
	std::vector< Level > levels;

	// build-phase device data
	DevBuf< double > dGalT , dPosS , dRWS , dPos , dSkew , dCellP , dCellData , dMeasure , dNoise , dDepth , dRW , dSkewScaled , dS , dS2 , dFS , dScale , dPartials , dTmp , dTmp2 , dDepthScale;
	DevBuf< int > dCellIdx , dCellStart , dCellSamples , dCellMap , dNodeFlag;
	DevBuf< unsigned int > dNodeList , dNodeCount;
	std::vector< std::unique_ptr< DevBuf< double > > > estData , skData;
	std::vector< std::unique_ptr< DevBuf< unsigned char > > > skKids , estKids;
	double c_ = 0;
	long long cElements = 1;   // elements the target-field mean of c_ runs over
	// --aRadius >= 0: the adapted hierarchy, built on the device with the other
	// build-phase data (levels[].fm/em/forder/eorder/act); the host copy with
	// the tree's shape is made only to write a .tree
	bool adapted( void ) const { return prm.aRadius!=-1; }
	std::unique_ptr< AdaptedHierarchy< Dim > > ah;
	AdaptedScratch adScratch;   // the device builder's temporaries, kept between builds
	template< typename T > std::vector< T > download( const DevBuf< T > &b , size_t n )
	{
		std::vector< T > h( n );
		b.download( h.data() , n , stream );
		HIPCHECK( hipStreamSynchronize( stream ) );
		return h;
	}
	const AdaptedHierarchy< Dim > &hostHierarchy( void )
	{
		if( !ah )
		{
			ah.reset( new AdaptedHierarchy< Dim >() );
			ah->D = D , ah->K = prm.aRadius;
			ah->levels.resize( D+1 );
			for( int l=0 ; l<=D ; l++ )
			{
				const Level &L = levels[l];
				auto &A = ah->levels[l];
				A.res = L.res , A.nActive = L.nActive;
				A.fmask = download( L.fm , (size_t)L.N );
				long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= L.res;
				A.emask = download( L.em , (size_t)ne );
				std::vector< unsigned int > fo = download( L.forder , L.nF ) , eo = download( L.eorder , L.nE );
				A.fOrder.assign( fo.begin() , fo.end() ) , A.eOrder.assign( eo.begin() , eo.end() );
			}
			ah->buildTree( cellIdx.data() , (size_t)ncell );
		}
		return *ah;
	}

	explicit Ctx( const epr_params &p ) : prm( p ) , D( p.depth )
	{
		if( D<0 || D>12 ) throw std::string( "depth must be in [0,12]" );
		if( prm.tRadius<0 || prm.tRadius>4 ) throw std::string( "Kernel radius must be in range [0,4]: " + std::to_string( prm.tRadius ) );
		if( prm.dRadius<0 || prm.dRadius>4 ) throw std::string( "Only radii [0,4] supported: " + std::to_string( prm.dRadius ) );
		if( prm.aRadius!=-1 && ( prm.aRadius<0 || prm.aRadius>5 ) ) throw std::string( "Only kernel radius in the range [0,5] supported: " + std::to_string( prm.aRadius ) );
		SetLdsAttributes();
		HIPCHECK( hipStreamCreate( &stream ) );
		levels.resize( D+1 );
		for( int l=0 ; l<=D ; l++ )
		{
			Level &L = levels[l];
			L.res = 1<<l;
			L.lay.set( L.res );
			L.N = 1;
			for( int d=0 ; d<Dim ; d++ ) L.N *= L.res+1;
			long long s = 1;
			for( int d=Dim-1 ; d>=0 ; d-- ){ L.stride[d] = s ; s *= L.res+1; }
		}
	}
	// --fp32: the levels whose colour launches sweep in FP32 (all iterations but
	// the last): dense levels of at least Fp32MinNodes functions.  The smaller
	// levels are latency-bound (FP32 gains nothing there) and run the most
	// iterations, so they stay FP64 like the one-workgroup levels.
	static constexpr long long Fp32MinNodes = 32768;
	bool floatLevel( int l ) const { return prm.fp32Sweeps && !adapted() && levels[l].N>=Fp32MinNodes; }
	// the levels' executable graphs are destroyed after this body: drain the stream first
	~Ctx(){ if( stream ){ (void)hipStreamSynchronize( stream ) ; (void)hipStreamDestroy( stream ); } }

	size_t numFunctions( int l ) const { return (size_t)levels[l].N; }
	int minD( void ) const { return std::min( std::max( prm.minSolveDepth , 0 ) , D ); }
	int maxD( void ) const { return prm.maxSolveDepth<0 ? D : std::min( prm.maxSolveDepth , D ); }
	void checkLevel( int l ) const { if( l<0 || l>D ) throw std::string( "level out of range" ); }

	LevelView< Dim > view( int l )
	{
		Level &L = levels[l];
		LevelView< Dim > v;
		v.res = L.res , v.N = L.N , v.x = L.x.p , v.y = L.y.p , v.R = L.R.p , v.B = L.B.p , v.K = L.K.p;
		v.sW = std::pow( (double)L.res , 4.-Dim ) , v.dW = prm.dWeight;
		for( int d=0 ; d<Dim ; d++ ) v.stride[d] = L.stride[d];
		v.lay = L.lay;
		v.fmask = L.fm.p;
		v.separate = prm.separate;
		return v;
	}

	// ---------------- samples ----------------
	void setSamples( size_t _n , const double *wpos , const double *wskew )
	{
		SampleCells< Dim > sc;
		sc.bin( _n , wpos , prm.scale , D , prm.jitter );
		n = _n;
		skewPacked.resize( n*CD );
		for( size_t s=0 ; s<n ; s++ ) for( int i=0 , idx=0 ; i<Dim ; i++ ) for( int j=0 ; j<i ; j++ , idx++ )
			skewPacked[s*CD+idx] = ( wskew[ s*Dim*Dim + i*Dim + j ] - wskew[ s*Dim*Dim + j*Dim + i ] ) / 2.;
		pos.swap( sc.pos );
		memcpy( unitToWorld , sc.unitToWorld , sizeof(unitToWorld) );
		cellIdx.swap( sc.cellIdx ) , cellStart.swap( sc.cellStart ) , cellSamples.swap( sc.cellSamples );
		ncell = sc.ncell;
		built = solved = false;
	}

	std::vector< long long > treeSizes( int nv , bool kids )
	{
		std::vector< long long > s( D+2 , 0 );
		for( int t=1 ; t<=D+1 ; t++ )
		{
			long long c = 1;
			if( kids ){ if( t>D ) continue ; for( int d=0 ; d<Dim ; d++ ) c *= 1LL<<t; }
			else for( int d=0 ; d<Dim ; d++ ) c *= ( 1LL<<(t-1) ) + 1;
			s[t] = c*nv;
		}
		return s;
	}

	// T = P^T M of a fine level into dGalT, all parity classes in one launch
	void galerkinT( const double *Mf , const Level &Fl , float *fMf=nullptr )
	{
		ColourSet< Dim > cs0 ; cs0.set( Fl.res , 0 );   // class 0 (all coordinates even) is the largest
		KGalerkinT< Dim ><<< dim3( Blocks( cs0.total , 64 ) , C ) , 64*( W/27 ) , 0 , stream >>>( Mf , dGalT.p , Fl.N , Fl.res/2 , Fl.lay , fMf );
	}
	// the coarser level's M from T (galerkinT of the level above it)
	void galerkinC( double *Mc , const Level &Cl , const Level &Fl )
	{
		if( Fl.N>=( 1LL<<31 ) ) throw std::string( "Galerkin restriction: more than 2^31 fine functions" );
		// small coarse levels: one thread per entry (more threads in flight)
		const bool perEntry = Cl.N*W<( 1LL<<20 );
		KGalerkinC< Dim ><<< perEntry ? Blocks( Cl.N*W , 256 ) : Blocks( Cl.N , 64 ) , 256 , 0 , stream >>>( dGalT.p , Mc , Cl.res , Cl.lay , Fl.lay , (int)Fl.N , Cl.fm.p , perEntry ? 1 : 0 );
	}

	template< int KR > void launchEstSplat( TreeGrids< Dim > T )
	{
		int taps = 1 ; for( int d=0 ; d<Dim ; d++ ) taps *= 2*KR+2;
		// depths whose corner values fit in LDS are summed per workgroup first:
		// those that fit twice into a CU's LDS in one launch (blockIdx.y = depth
		// - tLo, every workgroup sized for the largest), a larger one on its own
		// (so that it does not limit the smaller depths' workgroups to one a CU)
		const long long items = (long long)ncell*taps;
		const unsigned blocks = (unsigned)std::max< long long >( 1 , std::min< long long >( 256 , items/4096 ) );
		int t = 1 , tLo = 1;
		size_t lds = 0;
		auto flush = [&]( int tEnd )
		{
			if( tEnd>tLo ) KEstSplatLds< Dim , KR ><<< dim3( blocks , tEnd-tLo ) , 256 , lds , stream >>>( dCellP.p , dCellData.p , ncell , T , dDepthScale.p , tLo );
			tLo = tEnd , lds = 0;
		};
		for( ; t<=D+1 ; t++ )
		{
			long long corners = 1 ; for( int d=0 ; d<Dim ; d++ ) corners *= ( 1LL<<(t-1) ) + 1;
			const size_t b = (size_t)corners*NV*sizeof(double);
			if( b>LdsLimit ) break;
			if( b>LdsLimit/2 ){ flush( t ) ; lds = b ; flush( t+1 ); }
			else lds = b;
		}
		flush( t );
		if( t<=D+1 )
		{
			long long tot = (long long)ncell*(D+2-t)*taps;
			if( segSplats( 1 ) )
			{
				mortonOrder();
				static const bool wide = !( getenv( "EPR_EST_WIDE" ) && atoi( getenv( "EPR_EST_WIDE" ) )==0 );   // 0: per-lane atomics (for timing)
				if( wide && taps%64==0 ) KEstSplatSeg< Dim , KR , true ><<< Blocks( tot , 256 ) , 256 , 0 , stream >>>( dCellP.p , dCellData.p , dMKey[1].p , dMOrd[1].p , ncell , T , dDepthScale.p , t );
				else KEstSplatSeg< Dim , KR , false ><<< Blocks( tot , 256 ) , 256 , 0 , stream >>>( dCellP.p , dCellData.p , dMKey[1].p , dMOrd[1].p , ncell , T , dDepthScale.p , t );
			}
			else KEstSplat< Dim , KR ><<< Blocks( tot , 256 ) , 256 , 0 , stream >>>( dCellP.p , dCellData.p , ncell , T , dDepthScale.p , t , CoprimeStride( (unsigned int)ncell ) );
		}
	}
	// the splats over Morton runs of cells (EPR_EST_SEG bit 0: the estimator
	// splat, bit 1: the skew splat; a cleared bit: one atomic per cell or sample
	// and value, for timing)
	bool segSplats( int which ) const
	{
		static const int seg = getenv( "EPR_EST_SEG" ) ? atoi( getenv( "EPR_EST_SEG" ) ) : 3;
		return ( seg & which ) && Dim*D+4<=64 && ( which!=2 || n<( 1LL<<30 ) );
	}
	// the occupied cells sorted by Morton key: keys in dMKey[1], cell indices in
	// dMOrd[1] (once per build)
	DevBuf< unsigned long long > dMKey[2];
	DevBuf< int > dMOrd[2];
	DevBuf< unsigned char > dSortTmp;
	// the skew splat's (sample, depth) entries, sorted by (depth, depth-t cell) into [1]
	DevBuf< unsigned long long > dSKey[2];
	DevBuf< int > dSVal[2];
	bool mortonDone = false;
	void mortonOrder( void )
	{
		if( mortonDone ) return;
		mortonDone = true;
		for( int k=0 ; k<2 ; k++ ) dMKey[k].ensure( (size_t)ncell ) , dMOrd[k].ensure( (size_t)ncell );
		KMortonKeys< Dim ><<< Blocks( ncell , 256 ) , 256 , 0 , stream >>>( dCellIdx.p , ncell , D , dMKey[0].p , dMOrd[0].p );
		HIPCHECK( hipGetLastError() );
		size_t tb = 0;
		HIPCHECK( rocprim::radix_sort_pairs( nullptr , tb , dMKey[0].p , dMKey[1].p , dMOrd[0].p , dMOrd[1].p , (unsigned int)ncell , 0 , Dim*D , stream ) );
		dSortTmp.ensure( std::max< size_t >( tb , 1 ) );
		HIPCHECK( rocprim::radix_sort_pairs( (void *)dSortTmp.p , tb , dMKey[0].p , dMKey[1].p , dMOrd[0].p , dMOrd[1].p , (unsigned int)ncell , 0 , Dim*D , stream ) );
	}
	template< int KR > void launchSkewSplat( TreeGrids< Dim > T )
	{
		int taps = 1 ; for( int d=0 ; d<Dim ; d++ ) taps *= 2*KR+2;
		if( segSplats( 2 ) )
		{
			const long long nent = 2*(long long)n;
			for( int k=0 ; k<2 ; k++ ) dSKey[k].ensure( (size_t)nent ) , dSVal[k].ensure( (size_t)nent );
			KSkewEntries< Dim ><<< Blocks( (long long)n , 256 ) , 256 , 0 , stream >>>( dPos.p , dDepth.p , (int)n , D , dSKey[0].p , dSVal[0].p );
			HIPCHECK( hipGetLastError() );
			size_t tb = 0;
			HIPCHECK( rocprim::radix_sort_pairs( nullptr , tb , dSKey[0].p , dSKey[1].p , dSVal[0].p , dSVal[1].p , (unsigned int)nent , 0 , Dim*D+4 , stream ) );
			dSortTmp.ensure( std::max< size_t >( tb , 1 ) );
			HIPCHECK( rocprim::radix_sort_pairs( (void *)dSortTmp.p , tb , dSKey[0].p , dSKey[1].p , dSVal[0].p , dSVal[1].p , (unsigned int)nent , 0 , Dim*D+4 , stream ) );
			static const bool wide = !( getenv( "EPR_EST_WIDE" ) && atoi( getenv( "EPR_EST_WIDE" ) )==0 );
			if( wide && taps%64==0 ) KSkewSplatRuns< Dim , KR , true ><<< Blocks( nent*taps , 256 ) , 256 , 0 , stream >>>( dPos.p , dSkewScaled.p , dDepth.p , dSKey[1].p , dSVal[1].p , nent , T );
			else KSkewSplatRuns< Dim , KR , false ><<< Blocks( nent*taps , 256 ) , 256 , 0 , stream >>>( dPos.p , dSkewScaled.p , dDepth.p , dSKey[1].p , dSVal[1].p , nent , T );
		}
		else
		{
			const long long tot = (long long)n*taps;
			KSkewSplat< Dim , KR ><<< Blocks( tot , 256 ) , 256 , 0 , stream >>>( dPos.p , dSkewScaled.p , dDepth.p , (int)n , T , CoprimeStride( (unsigned int)n ) );
		}
	}

	// ---------------- build ----------------
	void build( void )
	{
		if( !n ) throw std::string( "no samples" );
		for( Level &L : levels ) L.narrowed = false , L.fusedR = L.fusedB = false;   // the float copies of R, B follow the new operators
		mortonDone = false;   // the cells may have changed
		auto t0 = std::chrono::steady_clock::now();
		const int res = 1<<D;
		dPos.upload( pos.data() , pos.size() , stream );
		dSkew.upload( skewPacked.data() , skewPacked.size() , stream );
		dCellIdx.upload( cellIdx.data() , cellIdx.size() , stream );
		dCellStart.upload( cellStart.data() , cellStart.size() , stream );
		dCellSamples.upload( cellSamples.data() , cellSamples.size() , stream );

		// ---- the trees of the estimator and the skew splat, and the coefficients: zeroed in one launch
		{
			const std::vector< long long > ez = treeSizes( NV , false ) , sz = treeSizes( CD , false ) , kz = treeSizes( 1 , true );
			estData.resize( D+2 ) , estKids.resize( D+1 ) , skData.resize( D+2 ) , skKids.resize( D+2 );
			std::vector< std::pair< void * , size_t > > spans;
			auto dbl = [&]( std::unique_ptr< DevBuf< double > > &b , long long m ){ if( !b ) b.reset( new DevBuf< double >() ) ; b->ensure( m ) ; spans.push_back( { (void *)b->p , (size_t)m*sizeof(double) } ); };
			auto u8 = [&]( std::unique_ptr< DevBuf< unsigned char > > &b , long long m ){ if( !b ) b.reset( new DevBuf< unsigned char >() ) ; b->ensure( m ) ; spans.push_back( { (void *)b->p , (size_t)m } ); };
			for( int t=1 ; t<=D+1 ; t++ ){ dbl( estData[t] , ez[t] ) , dbl( skData[t] , sz[t] ) ; if( t<=D ) u8( estKids[t] , kz[t] ) , u8( skKids[t] , kz[t] ); }
			for( int l=0 ; l<=D ; l++ ){ Level &L = levels[l] ; L.x.ensure( L.N ) , L.y.ensure( L.N ) ; spans.push_back( { (void *)L.x.p , (size_t)L.N*sizeof(double) } ) , spans.push_back( { (void *)L.y.p , (size_t)L.N*sizeof(double) } ); }
			zeroSpans( spans );
		}

		// ---- the adapted hierarchy (Hat.adaptive.inl:349-729) on every level
		ah.reset();
		if( adapted() )
		{
			AdaptedHierarchy< Dim >::CheckArgs( D , prm.aRadius );
			std::vector< AdaptedLevelDev * > ad;
			for( int l=0 ; l<=D ; l++ ) ad.push_back( &levels[l] );
			BuildAdaptedDevice< Dim >( dCellIdx.p , ncell , D , prm.aRadius , stream , ad , adScratch );
		}
		else for( int l=0 ; l<=D ; l++ )
		{
			Level &L = levels[l];
			L.fm.alloc( 0 ) , L.em.alloc( 0 ) , L.forder.alloc( 0 ) , L.eorder.alloc( 0 ) , L.act.alloc( 0 );
			L.nF = L.nActive = L.nE = 0;
		}

		// ---- tree estimator (GridSamples.inl:740-771)
		{
			const int K = prm.dRadius;
			double scale;
			{
				std::vector< double > v( 2*K+2 );
				// runtime B-spline (GridSamples.inl:50-65)
				std::function< void( int , double , double* ) > Ev = [&]( int deg , double x , double *val )
				{
					if( deg==0 ){ val[0] = 1. ; return; }
					const double Scale = 1./deg;
					Ev( deg-1 , x , val+1 );
					val[0] = val[1] * (1.-x) * Scale;
					for( int i=1 ; i<deg ; i++ ){ double x1 = ( x-i+deg ) , x2 = ( -x+i+1 ) ; val[i] = ( val[i]*x1 + val[i+1]*x2 ) * Scale; }
					val[deg] *= x * Scale;
				};
				Ev( 2*K+1 , 0.5 , &v[0] );
				scale = std::pow( 1./v[K+1] , CoDim );
			}
			std::vector< double > scales( D+1 );
			for( int d=0 ; d<=D ; d++ ) scales[d] = scale / std::pow( std::pow( 0.5 , d ) , Dim-CoDim );
			dDepthScale.upload( scales.data() , scales.size() , stream );
			dCellP.ensure( (size_t)ncell*Dim ) , dCellData.ensure( (size_t)ncell*NV );
			KCellStats< Dim ><<< Blocks( ncell , 128 ) , 128 , 0 , stream >>>( dPos.p , dCellStart.p , dCellSamples.p , ncell , dCellP.p , dCellData.p );
			std::vector< long long > sz = treeSizes( NV , false );
			TreeGrids< Dim > T;
			memset( &T , 0 , sizeof(T) );
			T.maxDepth = D;
			for( int t=1 ; t<=D+1 ; t++ ) T.data[t] = estData[t]->p;
			switch( K ){ case 0: launchEstSplat<0>( T ) ; break ; case 1: launchEstSplat<1>( T ) ; break ; case 2: launchEstSplat<2>( T ) ; break ; case 3: launchEstSplat<3>( T ) ; break ; default: launchEstSplat<4>( T ); }
			HIPCHECK( hipGetLastError() );
			// the estimator tree's shape (which nodes have children), for its file
			{
				std::vector< long long > kz = treeSizes( 1 , true );
				TreeGrids< Dim > TK;
				memset( &TK , 0 , sizeof(TK) );
				TK.maxDepth = D;
				for( int t=1 ; t<=D ; t++ ) TK.kids[t] = estKids[t]->p;
				if( D>=1 ) KTouch< Dim ><<< Blocks( (long long)ncell*D , 256 ) , 256 , 0 , stream >>>( dCellIdx.p , ncell , TK , K );
				HIPCHECK( hipGetLastError() );
			}
			dMeasure.ensure( n ) , dNoise.ensure( n ) , dDepth.ensure( n ) , dRW.ensure( n ) , dSkewScaled.ensure( n*CD );
			KEstQuery< Dim ><<< Blocks( (long long)n , 128 ) , 128 , 0 , stream >>>( dPos.p , dSkew.p , (int)n , T , prm.nSigma , dMeasure.p , dNoise.p , dDepth.p , dRW.p , dSkewScaled.p );
			HIPCHECK( hipGetLastError() );
		}
		// ---- skew splat and the target field (ExteriorPoisson.h:175-264, 290-309)
		{
			const int K = prm.tRadius;
			std::vector< long long > sz = treeSizes( CD , false ) , kz = treeSizes( 1 , true );
			TreeGrids< Dim > T;
			memset( &T , 0 , sizeof(T) );
			T.maxDepth = D;
			for( int t=1 ; t<=D+1 ; t++ ){ T.data[t] = skData[t]->p ; if( t<=D ) T.kids[t] = skKids[t]->p; }
			if( D>=1 ) KTouch< Dim ><<< Blocks( (long long)ncell*D , 256 ) , 256 , 0 , stream >>>( dCellIdx.p , ncell , T , K );
			switch( K ){ case 0: launchSkewSplat<0>( T ) ; break ; case 1: launchSkewSplat<1>( T ) ; break ; case 2: launchSkewSplat<2>( T ) ; break ; case 3: launchSkewSplat<3>( T ) ; break ; default: launchSkewSplat<4>( T ); }
			const double scl = std::pow( std::pow( 0.5 , D ) , 1.5 * CoDim );
			{
				Spans sp;
				unsigned long long mx = 1;
				for( int t=1 ; t<=D+1 ; t++ ) sp.p[t-1] = T.data[t] , sp.n[t-1] = (unsigned long long)sz[t] , mx = std::max( mx , sp.n[t-1] );
				KScaleSpans<<< dim3( (unsigned)std::min< unsigned long long >( ( mx+255 )/256 , 1024 ) , D+1 ) , 256 , 0 , stream >>>( sp , scl );
			}
			long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= res;
			dS.ensure( (size_t)ne*CD ) , dS2.ensure( (size_t)ne );
			KField< Dim ><<< Blocks( ne , 256 ) , 256 , 0 , stream >>>( T , dS.p , dS2.p );
			// adapted: c sums over the indexed elements only (SystemEnergy.h:455-461)
			if( adapted() ) KMaskZero<<< Blocks( ne , 256 ) , 256 , 0 , stream >>>( dS2.p , levels[D].em.p , ne );
			HIPCHECK( hipGetLastError() );
			// c = sum |S(E)|^2 / #elements (SystemEnergy.h:455-461)
			const int nb = 1024;
			dPartials.ensure( (size_t)nb*8 );
			dTmp.ensure( 8 );
			KPartialSum<<< nb , 256 , 0 , stream >>>( dS2.p , ne , dPartials.p );
			KSumRows<1><<< 1 , 256 , 0 , stream >>>( dPartials.p , nb , dTmp.p );
			// read back with the build's final synchronisation (nothing later in the build writes dTmp)
			cElements = ne;
		}
		// ---- finest R and B, then the Galerkin chain
		{
			Level &F = levels[D];
			F.R.ensure( (size_t)F.N*W ) , F.B.ensure( (size_t)F.N*W );
			// the FP32 sweeps' copies of the finest R and B written by the assembly
			// itself where it can (instead of a narrowing pass that reads them back)
			static const bool fuse = !( getenv( "EPR_FUSE_NARROW" ) && atoi( getenv( "EPR_FUSE_NARROW" ) )==0 );   // 0: a narrowing pass (for timing)
			const bool f32 = fuse && floatLevel( D );
			if( floatLevel( D ) ) F.fR.ensure( (size_t)F.N*W ) , F.fB.ensure( (size_t)F.N*W );
			// B
			static const bool bcols = !( getenv( "EPR_BCOLS" ) && atoi( getenv( "EPR_BCOLS" ) )==0 );   // 0: one thread per node in grid order (for timing)
			if( bcols && Dim==3 && !F.em.p && !F.fm.p )
			{
				KAssembleBCols< Dim ><<< Blocks( F.N , 256 ) , 256 , 0 , stream >>>( dS.p , res , std::pow( (double)res , 2.-Dim ) , F.B.p , F.lay , f32 ? F.fB.p : nullptr );
				F.fusedB = f32;
			}
			// (4D: a float copy written here, 81 strided stores a thread, made the
			// FP32 build slower, 51.5 -> 53.4 ms, than the narrowing pass)
			else KAssembleBNode< Dim ><<< Blocks( F.N , 256 ) , 256 , 0 , stream >>>( dS.p , res , std::pow( (double)res , 2.-Dim ) , F.B.p , F.lay , F.em.p , F.fm.p , nullptr );
			// R
			if( prm.sWeight>0 )
			{
				long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= res;
				// element -> occupied cell (-1: none), the element index with coordinate 0 fastest
				dCellMap.ensure( (size_t)ne );
				HIPCHECK( hipMemsetAsync( dCellMap.p , 0xFF , (size_t)ne*sizeof(int) , stream ) );
				KCellMap< Dim ><<< Blocks( ncell , 256 ) , 256 , 0 , stream >>>( dCellIdx.p , ncell , res , dCellMap.p );
				const int nb = 1024;
				dScale.ensure( 2 );
				KPartialSum<<< nb , 256 , 0 , stream >>>( dMeasure.p , (long long)n , dPartials.p );
				KMeanMeasure<<< 1 , 256 , 0 , stream >>>( dPartials.p , nb , (long long)n , prm.sWeight , dScale.p );
				// R is zero but in the columns of nodes next to occupied cells
				dNodeFlag.ensure( (size_t)F.N ) , dNodeList.ensure( std::min< size_t >( (size_t)F.N , (size_t)ncell<<Dim ) ) , dNodeCount.ensure( 1 );
				if( f32 ) zeroSpans( { { (void *)F.R.p , (size_t)F.N*W*sizeof(double) } , { (void *)F.fR.p , (size_t)F.N*W*sizeof(float) } , { (void *)dNodeFlag.p , (size_t)F.N*sizeof(int) } , { (void *)dNodeCount.p , sizeof(unsigned int) } } );
				else zeroSpans( { { (void *)F.R.p , (size_t)F.N*W*sizeof(double) } , { (void *)dNodeFlag.p , (size_t)F.N*sizeof(int) } , { (void *)dNodeCount.p , sizeof(unsigned int) } } );
				KOccupiedNodes< Dim ><<< Blocks( (long long)ncell<<Dim , 256 ) , 256 , 0 , stream >>>( dCellIdx.p , ncell , res , dNodeFlag.p , dNodeList.p , dNodeCount.p );
				dPosS.ensure( n*Dim ) , dRWS.ensure( n );
				KCellOrder< Dim ><<< Blocks( (long long)n , 256 ) , 256 , 0 , stream >>>( dPos.p , dRW.p , dCellSamples.p , (int)n , dPosS.p , dRWS.p );
				KAssembleR< Dim ><<< Blocks( std::min< long long >( F.N , (long long)ncell<<Dim )*W , 256 ) , 256 , 0 , stream >>>( dPosS.p , dCellMap.p , dCellStart.p , dRWS.p , res , dScale.p , F.R.p , F.lay , dNodeList.p , dNodeCount.p , f32 ? F.fR.p : nullptr );
				F.fusedR = f32;
			}
			else F.R.zero( stream );
			HIPCHECK( hipGetLastError() );
			if( D>0 ) dGalT.ensure( (size_t)levels[D].N*W );
			for( int l=D-1 ; l>=0 ; l-- )
			{
				Level &Cl = levels[l] , &Fl = levels[l+1];
				Cl.R.ensure( (size_t)Cl.N*W ) , Cl.B.ensure( (size_t)Cl.N*W );
				// the FP32 sweeps' copies of a float level's R and B written as
				// the restriction reads them (those its assembly has not written)
				const bool f32 = fuse && floatLevel( l+1 );
				if( f32 ) Fl.fR.ensure( (size_t)Fl.N*W ) , Fl.fB.ensure( (size_t)Fl.N*W );
				galerkinT( Fl.R.p , Fl , f32 && !Fl.fusedR ? Fl.fR.p : nullptr ) , galerkinC( Cl.R.p , Cl , Fl );
				galerkinT( Fl.B.p , Fl , f32 && !Fl.fusedB ? Fl.fB.p : nullptr ) , galerkinC( Cl.B.p , Cl , Fl );
				if( f32 ) Fl.fusedR = Fl.fusedB = true;
			}
			HIPCHECK( hipGetLastError() );
		}
		// ---- stencil case rows and coefficient storage (a captured graph that
		// is dropped may still run from an earlier call: drain the stream first)
		HIPCHECK( hipStreamSynchronize( stream ) );
		for( int l=0 ; l<=D ; l++ )
		{
			Level &L = levels[l];
			if( !L.K.n ){ std::vector< double > k = hat::StiffnessCaseRows< Dim >( L.res ) ; L.K.upload( k.data() , k.size() , stream ); }   // depends on the resolution only
			L.checkGraph();
		}
		// --fp32: the float copies of the operators of the levels swept by colour launches
		for( int l=0 ; l<=D ; l++ ) if( floatLevel( l ) && ( gsPlan( l ).kind==GS_PHASES || gsPlan( l , true ).kind==GS_PHASES ) ) narrowOperators( l );
		{
			double csum;
			dTmp.download( &csum , 1 , stream );
			HIPCHECK( hipStreamSynchronize( stream ) );
			c_ = csum / (double)cElements;
		}
		built = true , solved = false;
		seconds[0] = std::chrono::duration< double >( std::chrono::steady_clock::now()-t0 ).count();
	}

	void needBuilt( void ) const { if( !built ) throw std::string( "epr_build has not run" ); }

	// zero-fills of many device ranges, 64 per KZeroSpans launch
	void zeroSpans( const std::vector< std::pair< void * , size_t > > &spans )
	{
		for( size_t i=0 ; i<spans.size() ; i+=64 )
		{
			Spans sp;
			const unsigned cnt = (unsigned)std::min< size_t >( 64 , spans.size()-i );
			unsigned long long mx = 16;
			for( unsigned j=0 ; j<cnt ; j++ ) sp.p[j] = spans[i+j].first , sp.n[j] = spans[i+j].second , mx = std::max( mx , sp.n[j] );
			KZeroSpans<<< dim3( (unsigned)std::min< unsigned long long >( ( mx/16+255 )/256 , 512 ) , cnt ) , 256 , 0 , stream >>>( sp );
		}
		HIPCHECK( hipGetLastError() );
	}

	// ---------------- solve pieces ----------------
	// How a level's sweeps run: all iterations in one workgroup (KGsTiny: one
	// wavefront per element, a 16-lane row per node, all nodes of a colour at
	// once; KGsSmallLds / KGsSmall: a node group per 2^Dim lanes), or one launch
	// per colour phase.
	enum { GS_TINY , GS_SMALL_LDS , GS_SMALL , GS_PHASES , GS_WAVE , GS_BLOCK };
	struct GsPlan { int kind , threads , rb ; size_t lds; };
	// the plan of level l with a given kind of sweep kernel; kind -1 when that
	// kind cannot run the level
	GsPlan planKind( int l , int kind ) const
	{
		const long long N = levels[l].N;
		const int r = levels[l].res;
		long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= r;
		long long maxColour = 0;
		for( int c=0 ; c<C ; c++ ){ ColourSet< Dim > cs ; cs.set( r , c ) ; maxColour = std::max( maxColour , cs.total ); }
		switch( kind )
		{
		// one wavefront, registers: 3D levels of resolution 1 and 2 (in 4D each
		// row lane's 120 cubic-table coefficients do not fit in registers: 512
		// VGPRs and 1.7 K spills measured for the resolution-1 form)
		case GS_WAVE: if( Dim==3 && r<=2 && N==(long long)( r+1 )*( r+1 )*( r+1 ) ) return { GS_WAVE , 64 , 0 , 0 }; break;
		// eight wavefronts, one LDS exchange per colour: the 3D resolution-4 level
		case GS_BLOCK: if( Dim==3 && r==4 && N==125 ) return { GS_BLOCK , 512 , 0 , (size_t)BlockLdsBytes }; break;
		case GS_TINY:
		{
			const int maxWaves = TinyMaxThreads< Dim >()/64;
			const int tinyWaves = (int)std::min< long long >( maxWaves , std::max( ne , ( maxColour+3 )/4 ) );
			const bool tinyRB = TinyLds< Dim >::Bytes( N , ne , tinyWaves , true ) <= LdsLimit;
			if( maxColour<=4*maxWaves && TinyLds< Dim >::Bytes( N , ne , tinyWaves , tinyRB )<=LdsLimit )
				return { GS_TINY , 64*tinyWaves , tinyRB ? 1 : 0 , TinyLds< Dim >::Bytes( N , ne , tinyWaves , tinyRB ) };
			break;
		}
		case GS_SMALL_LDS:
		{
			// one workgroup sized to the largest colour, coefficients in LDS (and R, B when they fit)
			const int threads = (int)std::min< long long >( Dim==3 ? 1024 : 512 , ( maxColour*C + 63 )/64*64 );
			size_t lds = ( 2*N + W*W )*sizeof(double);
			const int rb = lds + 2*N*W*sizeof(double) <= LdsLimit ? 1 : 0;
			if( rb ) lds += 2*N*W*sizeof(double);
			if( lds<=LdsLimit ) return { GS_SMALL_LDS , threads , rb , lds };
			break;
		}
		case GS_SMALL: return { GS_SMALL , 512 , 0 , 0 };
		case GS_PHASES: return { GS_PHASES , 256 , 0 , 0 };
		}
		return { -1 , 0 , 0 , 0 };
	}
	// The plan of level l: the tuned kind and colour-launch workgroup (single
	// or batched solve) where the tuning table has one, else the built-in rule
	// (measured on MI355X): one wavefront for 3D resolutions 1 and 2, eight for
	// 3D resolution 4, KGsTiny up to tinyElements elements, one workgroup up to
	// SmallLevel nodes, colour launches of 256 threads above.
	GsPlan gsPlan( int l , bool batched=false ) const
	{
		const TunePick t = TuneTable::Get().find( Dim , levels[l].res , batched , false );
		if( t.kind>=0 )
		{
			GsPlan p = planKind( l , t.kind );
			if( p.kind>=0 ){ if( p.kind==GS_PHASES && t.threads ) p.threads = t.threads ; return p; }
		}
		GsPlan p = { -1 , 0 , 0 , 0 };
		if( !getenv( "EPR_NO_WAVE" ) ) p = planKind( l , GS_WAVE );
		if( p.kind<0 && !getenv( "EPR_NO_WAVE" ) ) p = planKind( l , GS_BLOCK );
		long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= levels[l].res;
		if( p.kind<0 && ne<=tinyElements ) p = planKind( l , GS_TINY );
		if( p.kind<0 && levels[l].N<=SmallLevel ){ p = planKind( l , GS_SMALL_LDS ) ; if( p.kind<0 ) p = planKind( l , GS_SMALL ); }
		if( p.kind<0 ) p = planKind( l , GS_PHASES );
		if( p.kind==GS_PHASES && t.threads ) p.threads = t.threads;
		return p;
	}
	// workgroup size of level l's FP32 colour launches
	int floatThreads( int l , bool batched=false ) const
	{
		const TunePick t = TuneTable::Get().find( Dim , levels[l].res , batched , true );
		return t.threads ? t.threads : 256;
	}
	// one colour launch of the sweep with `threads` per workgroup; 32-bit
	// indices when every index of the level (node, column entry) is below 2^31
	template< typename Real >
	void launchPhase( const LevelView< Dim , Real > &v , const ColourSet< Dim > &cs , long long colBase , int threads=256 )
	{
		// 32-bit byte offsets while every coefficient offset of the level fits
		if( v.N*(long long)sizeof(Real) < ( 1LL<<32 ) && !v.fmask )
		{
			ColourDiv< Dim > cd ; cd.set( cs );
			KGsPhase32< Dim , Real ><<< Blocks( cs.total*C , threads ) , threads , 0 , stream >>>( v , cd , colBase );
		}
		else if( v.N*W < ( 1LL<<31 ) ) KGsPhase< Dim , Real , int ><<< Blocks( cs.total*C , threads ) , threads , 0 , stream >>>( v , cs , colBase );
		else KGsPhase< Dim , Real , long long ><<< Blocks( cs.total*C , threads ) , threads , 0 , stream >>>( v , cs , colBase );
	}
	void gs( int l , int iters )
	{
		needBuilt() ; checkLevel( l );
		if( iters<=0 ) return;
		LevelView< Dim > v = view( l );
		const GsPlan pl = gsPlan( l );
		if( pl.kind==GS_WAVE )
		{
			if constexpr( Dim==3 )
			{
				if( levels[l].res==1 ) KGsWave< 1 ><<< 1 , 64 , 0 , stream >>>( v , iters );
				else KGsWave< 2 ><<< 1 , 64 , 0 , stream >>>( v , iters );
			}
		}
		else if( pl.kind==GS_BLOCK ){ if constexpr( Dim==3 ) KGsBlock<<< 1 , 512 , pl.lds , stream >>>( v , iters ); }
		else if( pl.kind==GS_TINY )
		{
			if( pl.rb ) KGsTiny< Dim , true  ><<< 1 , pl.threads , pl.lds , stream >>>( v , iters );
			else        KGsTiny< Dim , false ><<< 1 , pl.threads , pl.lds , stream >>>( v , iters );
		}
		else if( pl.kind==GS_SMALL_LDS )
		{
			KGsSmallLds< Dim ><<< 1 , pl.threads , pl.lds , stream >>>( v , iters , pl.rb );
		}
		else if( pl.kind==GS_SMALL ) KGsSmall< Dim ><<< 1 , 512 , 0 , stream >>>( v , iters );
		else
		{
			Level &L = levels[l];
			if( floatLevel( l ) && iters>1 )
			{
				// --fp32: all iterations but the last in FP32, the last in FP64 from
				// the widened iterate (the refinement), energies FP64
				gsFloat( l , iters-1 );
				iters = 1;
			}
			// the iters*2^Dim colour launches of a level are captured once into a
			// HIP graph and replayed (the level's buffers are stable between builds
			// of the same size)
			if( !L.graph || L.graphIters!=iters || L.graphThreads!=pl.threads )
			{
				// an executable graph is destroyed only once its last launch has ended
				if( L.graph ){ HIPCHECK( hipStreamSynchronize( stream ) ) ; (void)hipGraphExecDestroy( L.graph ) ; L.graph = nullptr; }
				hipGraph_t g;
				HIPCHECK( hipStreamBeginCapture( stream , hipStreamCaptureModeThreadLocal ) );
				for( int it=0 ; it<iters ; it++ ) for( int c=0 ; c<C ; c++ )
				{
					if( L.fm.p )   // adapted: the active nodes of the colour only
					{
						const long long cnt = L.actStart[c+1] - L.actStart[c];
						if( cnt ) KGsPhaseList< Dim ><<< Blocks( cnt*C , pl.threads ) , pl.threads , 0 , stream >>>( v , L.act.p + L.actStart[c] , cnt );
					}
					else
					{
						ColourSet< Dim > cs ; cs.set( L.res , c );
						launchPhase( v , cs , v.lay.cstart[c] , pl.threads );
					}
				}
				HIPCHECK( hipStreamEndCapture( stream , &g ) );
				HIPCHECK( hipGraphInstantiate( &L.graph , g , nullptr , nullptr , 0 ) );
				HIPCHECK( hipGraphDestroy( g ) );
				L.graphIters = iters , L.graphThreads = pl.threads;
				L.graphX = L.x.p , L.graphR = L.R.p , L.graphF = L.fm.p , L.graphAct = L.act.p;
				for( int c=0 ; c<=C ; c++ ) L.graphActStart[c] = L.actStart[c];
			}
			HIPCHECK( hipGraphLaunch( L.graph , stream ) );
		}
		HIPCHECK( hipGetLastError() );
	}

	// one colour of one Gauss-Seidel iteration (the body of the colour loop of
	// PairedGaussSeidelSolve, ExteriorPoissonSolvers.h:163-185), on levels
	// swept by colour launches
	void gsColour( int l , int c )
	{
		needBuilt() ; checkLevel( l );
		if( c<0 || c>=C ) throw std::string( "colour out of range" );
		const GsPlan pl = gsPlan( l );
		if( pl.kind!=GS_PHASES ) throw std::string( "single colour phases run on the levels swept by colour launches only" );
		LevelView< Dim > v = view( l );
		Level &L = levels[l];
		if( L.fm.p )
		{
			const long long cnt = L.actStart[c+1] - L.actStart[c];
			if( cnt ) KGsPhaseList< Dim ><<< Blocks( cnt*C , pl.threads ) , pl.threads , 0 , stream >>>( v , L.act.p + L.actStart[c] , cnt );
		}
		else
		{
			ColourSet< Dim > cs ; cs.set( L.res , c );
			launchPhase( v , cs , v.lay.cstart[c] , pl.threads );
		}
		HIPCHECK( hipGetLastError() );
		HIPCHECK( hipStreamSynchronize( stream ) );
	}

	// the float copies of level l's R, B and K (once per build), and room for x, y
	void narrowOperators( int l )
	{
		Level &L = levels[l];
		const long long N = L.N;
		if( L.narrowed ) return;
		L.fx.ensure( N ) , L.fy.ensure( N ) , L.fR.ensure( (size_t)N*W ) , L.fB.ensure( (size_t)N*W ) , L.fK.ensure( (size_t)W*W );
		if( !L.fusedR ) KNarrow<<< Blocks( N*W , 256 ) , 256 , 0 , stream >>>( L.R.p , L.fR.p , N*W );
		if( !L.fusedB ) KNarrow<<< Blocks( N*W , 256 ) , 256 , 0 , stream >>>( L.B.p , L.fB.p , N*W );
		KNarrow<<< Blocks( (long long)W*W , 256 ) , 256 , 0 , stream >>>( L.K.p , L.fK.p , (long long)W*W );
		HIPCHECK( hipGetLastError() );
		L.narrowed = true;
	}
	// iters FP32 iterations of a dense level swept by colour launches: x, y
	// (and, once per build, R, B, K) rounded to float, the float colour
	// launches (captured into a graph and replayed), x, y widened back
	void gsFloat( int l , int iters )
	{
		Level &L = levels[l];
		const long long N = L.N;
		narrowOperators( l );
		KNarrow<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( L.x.p , L.fx.p , N );
		KNarrow<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( L.y.p , L.fy.p , N );
		// the captured launches stay valid across builds while the float buffers stay put
		const float *bufs[5] = { L.fx.p , L.fy.p , L.fR.p , L.fB.p , L.fK.p };
		bool moved = false;
		for( int k=0 ; k<5 ; k++ ) moved |= bufs[k]!=L.graph32P[k];
		const int threads = floatThreads( l );
		if( !L.graph32 || L.graph32Iters!=iters || moved || L.graph32Threads!=threads )
		{
			if( L.graph32 ){ HIPCHECK( hipStreamSynchronize( stream ) ) ; (void)hipGraphExecDestroy( L.graph32 ) ; L.graph32 = nullptr; }
			const LevelView< Dim > v = view( l );
			LevelView< Dim , float > f;
			f.res = v.res , f.N = v.N , f.x = L.fx.p , f.y = L.fy.p , f.R = L.fR.p , f.B = L.fB.p , f.K = L.fK.p;
			f.sW = (float)v.sW , f.dW = (float)v.dW , f.lay = v.lay , f.fmask = nullptr , f.separate = v.separate;
			for( int d=0 ; d<Dim ; d++ ) f.stride[d] = v.stride[d];
			hipGraph_t g;
			HIPCHECK( hipStreamBeginCapture( stream , hipStreamCaptureModeThreadLocal ) );
			for( int it=0 ; it<iters ; it++ ) for( int c=0 ; c<C ; c++ )
			{
				ColourSet< Dim > cs ; cs.set( L.res , c );
				launchPhase( f , cs , f.lay.cstart[c] , threads );
			}
			HIPCHECK( hipStreamEndCapture( stream , &g ) );
			HIPCHECK( hipGraphInstantiate( &L.graph32 , g , nullptr , nullptr , 0 ) );
			HIPCHECK( hipGraphDestroy( g ) );
			L.graph32Iters = iters , L.graph32Threads = threads;
			for( int k=0 ; k<5 ; k++ ) L.graph32P[k] = bufs[k];
		}
		HIPCHECK( hipGraphLaunch( L.graph32 , stream ) );
		KWiden<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( L.fx.p , L.x.p , N );
		KWiden<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( L.fy.p , L.y.p , N );
		HIPCHECK( hipGetLastError() );
	}

	// The FP32 form of the colour launches (measurement): the level's x, y, R,
	// B and K rounded to float, iters iterations of KGsPhase< Dim , float >
	// timed with HIP events, the float x, y written back to the level.
	DevBuf< float > fX , fY , fR , fB , fK;
	void timeGsF32( int l , int iters , double *ms , int *launches )
	{
		needBuilt() ; checkLevel( l );
		Level &L = levels[l];
		if( L.fm.p ) throw std::string( "the FP32 sweep measurement runs on dense levels" );
		if( gsPlan( l ).kind!=GS_PHASES ) throw std::string( "the FP32 sweep measurement runs on levels swept by colour launches" );
		const long long N = L.N;
		fX.ensure( N ) , fY.ensure( N ) , fR.ensure( (size_t)N*W ) , fB.ensure( (size_t)N*W ) , fK.ensure( (size_t)W*W );
		KNarrow<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( L.x.p , fX.p , N );
		KNarrow<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( L.y.p , fY.p , N );
		KNarrow<<< Blocks( N*W , 256 ) , 256 , 0 , stream >>>( L.R.p , fR.p , N*W );
		KNarrow<<< Blocks( N*W , 256 ) , 256 , 0 , stream >>>( L.B.p , fB.p , N*W );
		KNarrow<<< Blocks( (long long)W*W , 256 ) , 256 , 0 , stream >>>( L.K.p , fK.p , (long long)W*W );
		const LevelView< Dim > v = view( l );
		LevelView< Dim , float > f;
		f.res = v.res , f.N = v.N , f.x = fX.p , f.y = fY.p , f.R = fR.p , f.B = fB.p , f.K = fK.p;
		f.sW = (float)v.sW , f.dW = (float)v.dW , f.lay = v.lay , f.fmask = nullptr , f.separate = v.separate;
		for( int d=0 ; d<Dim ; d++ ) f.stride[d] = v.stride[d];
		const int threads = floatThreads( l );
		hipEvent_t e0 , e1;
		HIPCHECK( hipEventCreate( &e0 ) ) ; HIPCHECK( hipEventCreate( &e1 ) );
		HIPCHECK( hipEventRecord( e0 , stream ) );
		for( int it=0 ; it<iters ; it++ ) for( int c=0 ; c<C ; c++ )
		{
			ColourSet< Dim > cs ; cs.set( L.res , c );
			launchPhase( f , cs , f.lay.cstart[c] , threads );
		}
		HIPCHECK( hipEventRecord( e1 , stream ) );
		HIPCHECK( hipEventSynchronize( e1 ) );
		float t = 0;
		HIPCHECK( hipEventElapsedTime( &t , e0 , e1 ) );
		HIPCHECK( hipEventDestroy( e0 ) ) ; HIPCHECK( hipEventDestroy( e1 ) );
		KWiden<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( fX.p , L.x.p , N );
		KWiden<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( fY.p , L.y.p , N );
		HIPCHECK( hipGetLastError() );
		HIPCHECK( hipStreamSynchronize( stream ) );
		if( ms ) *ms = t;
		if( launches ) *launches = iters*C;
	}

	void timeGs( int l , int iters , double *ms , int *launches )
	{
		needBuilt() ; checkLevel( l );
		hipEvent_t e0 , e1;
		HIPCHECK( hipEventCreate( &e0 ) ) ; HIPCHECK( hipEventCreate( &e1 ) );
		HIPCHECK( hipEventRecord( e0 , stream ) );
		gs( l , iters );
		HIPCHECK( hipEventRecord( e1 , stream ) );
		HIPCHECK( hipEventSynchronize( e1 ) );
		float t = 0;
		HIPCHECK( hipEventElapsedTime( &t , e0 , e1 ) );
		HIPCHECK( hipEventDestroy( e0 ) ) ; HIPCHECK( hipEventDestroy( e1 ) );
		if( ms ) *ms = t;
		if( launches ) *launches = gsPlan( l ).kind==GS_PHASES ? iters*C : 1;
	}

	// P x (mode 0), out += P x (1) or out -= P x (2).  Adapted levels use the
	// hierarchy's prolongation (rows of indexed fine functions only) unless
	// regular is set (the regular prolongation of regularCoefficients,
	// Hat.adaptive.inl:155-176).
	void prolong( int l , const double *in , double *out , int mode , bool regular=false )
	{
		long long nf = levels[l+1].N;
		KProlong< Dim ><<< Blocks( nf , 256 ) , 256 , 0 , stream >>>( in , out , levels[l].res , mode , regular ? nullptr : levels[l+1].fm.p );
	}
	// the same for x and y in one launch (32-bit indices while the fine level
	// has fewer than 2^31 functions, else one launch each)
	void prolong2( int l , const double *inX , const double *inY , double *outX , double *outY , int mode , bool regular=false )
	{
		long long nf = levels[l+1].N;
		if( nf>=( 1LL<<31 ) ){ prolong( l , inX , outX , mode , regular ) ; prolong( l , inY , outY , mode , regular ) ; return; }
		FastDiv n1 ; n1.set( (unsigned int)( 2*levels[l].res+1 ) );
		KProlong2< Dim ><<< Blocks( nf , 256 ) , 256 , 0 , stream >>>( inX , inY , outX , outY , levels[l].res , mode , regular ? nullptr : levels[l+1].fm.p , n1 );
	}

	void prolongAdd( int l )
	{
		needBuilt();
		if( l<0 || l>=D ) throw std::string( "level out of range" );
		prolong2( l , levels[l].x.p , levels[l].y.p , levels[l+1].x.p , levels[l+1].y.p , 1 );
		HIPCHECK( hipGetLastError() );
		HIPCHECK( hipStreamSynchronize( stream ) );
	}

	// Measurement of the HBM-bound build and transfer kernels: reps launches of
	// one of them at level l after a warm-up, on scratch outputs (the level's
	// data is left as it is); ms per launch and the launch's algorithmic HBM
	// bytes (each input and output value once):
	//   "prolong"     KProlong2, level l -> l+1, x and y: out += P in, 2 (N_l + 2 N_{l+1}) doubles
	//   "galerkin_t"  KGalerkinT, T = P^T R of level l (> 0): 2 N_l 3^Dim doubles
	//   "galerkin_c"  KGalerkinC, R_{l-1} = T P from level l's T: (N_l + N_{l-1}) 3^Dim doubles
	//   "energy"      KEnergy + KSumRows at level l: x, y, R, B (2 N_l + 2 N_l 3^Dim doubles)
	void timeOp( const std::string &op , int l , int reps , double *ms , double *bytes )
	{
		needBuilt() ; checkLevel( l );
		if( reps<=0 ) throw std::string( "reps must be positive" );
		Level &L = levels[l];
		std::function< void( void ) > f;
		double b = 0;
		if( op=="prolong" )
		{
			if( l>=D ) throw std::string( "prolong: level must be below the depth" );
			Level &Fl = levels[l+1];
			dTmp.ensure( (size_t)Fl.N ) , dTmp2.ensure( (size_t)Fl.N );
			HIPCHECK( hipMemsetAsync( dTmp.p , 0 , (size_t)Fl.N*sizeof(double) , stream ) );
			HIPCHECK( hipMemsetAsync( dTmp2.p , 0 , (size_t)Fl.N*sizeof(double) , stream ) );
			f = [&]{ prolong2( l , L.x.p , L.y.p , dTmp.p , dTmp2.p , 1 ); };
			b = 16.*( L.N + 2*Fl.N );
		}
		else if( op=="galerkin_t" )
		{
			if( l<1 ) throw std::string( "galerkin_t: level must be above 0" );
			dGalT.ensure( (size_t)L.N*W );
			f = [&]{ galerkinT( L.R.p , L ); };
			b = 16.*L.N*W;
		}
		else if( op=="galerkin_c" )
		{
			if( l<1 ) throw std::string( "galerkin_c: level must be above 0" );
			Level &Cl = levels[l-1];
			dGalT.ensure( (size_t)L.N*W );
			galerkinT( L.R.p , L );
			dTmp2.ensure( (size_t)Cl.N*W );
			f = [&]{ galerkinC( dTmp2.p , Cl , L ); };
			b = 8.*( L.N + Cl.N )*W;
		}
		else if( op=="energy" )
		{
			dPartials.ensure( (size_t)EnergyRows< Dim >( L.N , L.res )*4 ) , dTmp.ensure( 8 );
			const LevelView< Dim > v = view( l );
			f = [&]{ LaunchEnergy< Dim >( v , dPartials.p , dTmp.p , stream ); };
			b = 8.*( 2*L.N + 2*L.N*W );
		}
		else throw std::string( "unknown operation: " + op );
		f();
		HIPCHECK( hipGetLastError() );
		const double t = TimeOnStream( stream , [&]{ for( int r=0 ; r<reps ; r++ ) f(); } );
		HIPCHECK( hipGetLastError() );
		if( ms ) *ms = t/reps;
		if( bytes ) *bytes = b;
	}

	// (W, yBx, R-form, K-form) at a level
	void energyTerms( int l , double t[4] )
	{
		LevelView< Dim > v = view( l );
		dPartials.ensure( (size_t)EnergyRows< Dim >( levels[l].N , levels[l].res )*4 );
		dTmp.ensure( 8 );
		LaunchEnergy< Dim >( v , dPartials.p , dTmp.p , stream );
		dTmp.download( t , 4 , stream );
		HIPCHECK( hipStreamSynchronize( stream ) );
	}

	void energies( int l , double out[2] )
	{
		needBuilt() ; checkLevel( l );
		double t[4];
		energyTerms( l , t );
		out[0] = t[0] - 2.*t[1] + c_;
		out[1] = t[2];
		if( prm.dWeight ) out[1] += t[3]*prm.dWeight;
	}

	// energies() with the partial sums in the device's static scratch instead
	// of the context's buffers: no allocation, the same kernels and summation
	// order (so the same bits)
	void energiesNoAlloc( int l , double out[2] )
	{
		needBuilt() ; checkLevel( l );
		LevelView< Dim > v = view( l );
		const long long nb = EnergyRows< Dim >( levels[l].N , levels[l].res );
		if( nb>NoAllocMaxBlocks )
			throw std::string( "level too large for epr_energies_noalloc: " + std::to_string( nb ) + " partial sums, the scratch holds " + std::to_string( NoAllocMaxBlocks ) );
		double *sp = NoAllocScratch();
		LaunchEnergy< Dim >( v , sp + 4 , sp , stream );
		HIPCHECK( hipGetLastError() );
		double t[4];
		HIPCHECK( hipMemcpyAsync( t , sp , 4*sizeof(double) , hipMemcpyDeviceToHost , stream ) );
		HIPCHECK( hipStreamSynchronize( stream ) );
		out[0] = t[0] - 2.*t[1] + c_;
		out[1] = t[2];
		if( prm.dWeight ) out[1] += t[3]*prm.dWeight;
	}

	void init( int l , const double *x0 , const double *y0 )
	{
		needBuilt() ; checkLevel( l );
		Level &L = levels[l];
		if( x0 && y0 ){ L.x.upload( x0 , L.N , stream ) ; L.y.upload( y0 , L.N , stream ) ; HIPCHECK( hipStreamSynchronize( stream ) ) ; return; }
		// ExteriorPoisson.h:333-341; the levels above start at zero (the cascade adds into them)
		{
			std::vector< std::pair< void * , size_t > > spans;
			for( int k=l+1 ; k<=D ; k++ ) spans.push_back( { (void *)levels[k].x.p , (size_t)levels[k].N*sizeof(double) } ) , spans.push_back( { (void *)levels[k].y.p , (size_t)levels[k].N*sizeof(double) } );
			if( !spans.empty() ) zeroSpans( spans );
		}
		std::vector< double > hx( L.N , 0. ) , hy( L.N , 0. );
		GlibcRand g( (unsigned int)prm.seed );
		if( adapted() ){ for( unsigned int f : download( L.forder , L.nF ) ) hx[f] = g.random01() , hy[f] = g.random01(); }   // the hierarchy's numbering
		else for( long long i=0 ; i<L.N ; i++ ) hx[i] = g.random01() , hy[i] = g.random01();
		L.x.upload( hx.data() , L.N , stream ) , L.y.upload( hy.data() , L.N , stream );
		double t[4];
		energyTerms( l , t );
		double s = std::sqrt( std::sqrt( c_ / t[0] ) );
		dTmp2.ensure( 1 );
		dTmp2.upload( &s , 1 , stream );
		KScale2<<< Blocks( L.N , 256 ) , 256 , 0 , stream >>>( L.x.p , L.y.p , L.N , dTmp2.p );
		HIPCHECK( hipStreamSynchronize( stream ) );
	}

	// ---------------- autotuning ----------------
	int scheduleIters( int l ) const { return (int)std::ceil( std::pow( prm.iMult , D-l ) * prm.gsIters ); }
	// Times every sweep kernel that can run each level of the solve on this
	// context's own data (x and y restored after every trial) and keeps the
	// fastest in the tuning table (single-context plans): the one-workgroup
	// kernels on levels of at most TuneOneGroupMax nodes, colour launches of
	// 64, 128 and 256 threads; flags bit 0 also times the FP32 colour launches'
	// workgroup sizes.  A trial runs the level's schedule iterations, at most
	// 128 and at most 2^24 / N of them.  Returns one line per trial: "level res
	// nodes precision kind threads us_per_iteration", the kept ones marked '*'.
	std::string autotune( int flags )
	{
		needBuilt();
		std::string rep;
		char line[200];
		const int fp32 = prm.fp32Sweeps;
		prm.fp32Sweeps = 0;   // the FP64 form of a level alone
		DevBuf< double > sx , sy;
		try
		{
			for( int l=minD() ; l<=maxD() ; l++ )
			{
				Level &L = levels[l];
				const int its = (int)std::max< long long >( 1 , std::min< long long >( { (long long)scheduleIters( l ) , 128LL , ( 1LL<<24 )/L.N } ) );
				sx.ensure( L.N ) , sy.ensure( L.N );
				HIPCHECK( hipMemcpyAsync( sx.p , L.x.p , L.N*sizeof(double) , hipMemcpyDeviceToDevice , stream ) );
				HIPCHECK( hipMemcpyAsync( sy.p , L.y.p , L.N*sizeof(double) , hipMemcpyDeviceToDevice , stream ) );
				auto restore = [&]( void )
				{
					HIPCHECK( hipMemcpyAsync( L.x.p , sx.p , L.N*sizeof(double) , hipMemcpyDeviceToDevice , stream ) );
					HIPCHECK( hipMemcpyAsync( L.y.p , sy.p , L.N*sizeof(double) , hipMemcpyDeviceToDevice , stream ) );
				};
				std::vector< std::pair< int , int > > cand;
				if( L.N<=TuneOneGroupMax ) for( int k : { (int)GS_WAVE , (int)GS_BLOCK , (int)GS_TINY , (int)GS_SMALL_LDS } ) if( planKind( l , k ).kind>=0 ) cand.push_back( { k , 0 } );
				for( int th : TuneThreads ) cand.push_back( { (int)GS_PHASES , th } );
				double best = 1e300;
				std::pair< int , int > bc = cand.back();
				std::vector< std::string > lines;
				for( const auto &c : cand )
				{
					TuneTable::Get().set( Dim , L.res , false , false , TunePick{ c.first , c.second , 0 } );
					gs( l , its ) , restore();   // warm-up (and the graph's capture)
					const double us = 1000.*TimeOnStream( stream , [&]{ gs( l , its ); } )/its;
					restore();
					snprintf( line , sizeof(line) , "%d %d %lld fp64 %s %d %.3f" , l , L.res , L.N , GsKindName( c.first ) , c.first==GS_PHASES ? c.second : planKind( l , c.first ).threads , us );
					lines.push_back( line );
					if( us<best ) best = us , bc = c;
				}
				TuneTable::Get().set( Dim , L.res , false , false , TunePick{ bc.first , bc.second , best } );
				for( size_t k=0 ; k<cand.size() ; k++ ) rep += lines[k] + ( cand[k]==bc ? " *\n" : "\n" );
				if( ( flags & 1 ) && bc.first==GS_PHASES && !L.fm.p && L.N>=Fp32MinNodes )
				{
					narrowOperators( l );
					double bestF = 1e300;
					int bt = 256;
					std::vector< std::pair< std::string , int > > fl;
					for( int th : TuneThreads )
					{
						TuneTable::Get().set( Dim , L.res , false , true , TunePick{ -1 , th , 0 } );
						gsFloat( l , its ) , restore();
						const double us = 1000.*TimeOnStream( stream , [&]{ gsFloat( l , its ); } )/its;
						restore();
						snprintf( line , sizeof(line) , "%d %d %lld fp32 phases %d %.3f" , l , L.res , L.N , th , us );
						fl.push_back( { line , th } );
						if( us<bestF ) bestF = us , bt = th;
					}
					TuneTable::Get().set( Dim , L.res , false , true , TunePick{ -1 , bt , bestF } );
					for( const auto &f : fl ) rep += f.first + ( f.second==bt ? " *\n" : "\n" );
				}
				HIPCHECK( hipStreamSynchronize( stream ) );
			}
		}
		catch( ... ){ prm.fp32Sweeps = fp32 ; throw; }
		prm.fp32Sweeps = fp32;
		return rep;
	}

	// the cascadic schedule (ExteriorPoisson.h:316-409)
	void solve( double out[2] ){ solveFrom( out , nullptr ); }
	// iters (optional): the iteration count of every level d in [start, maxSolveDepth]
	// at iters[d] instead of the reference's ceil( iMult^(D-d) gsIters )
	void solveFrom( double out[2] , unet::Net *net , const int *sched=nullptr )
	{
		needBuilt();
		auto t0 = std::chrono::steady_clock::now();
		int lo = minD();
		const int hi = maxD();
		if( net )
		{
			lo = unet::Level( net );
			if( lo<minD() || lo>hi ) throw std::string( "the network's level is outside [minSolveDepth, maxSolveDepth]" );
			if( prm.singleLevel ) throw std::string( "--singleLevel has no coarse start to replace" );
		}
		{
			std::vector< std::pair< void * , size_t > > spans;
			for( int l=0 ; l<=D ; l++ ) spans.push_back( { (void *)levels[l].x.p , (size_t)levels[l].N*sizeof(double) } ) , spans.push_back( { (void *)levels[l].y.p , (size_t)levels[l].N*sizeof(double) } );
			zeroSpans( spans );
		}
		if( net ) learnedInit( lo , net , nullptr );
		else init( lo , nullptr , nullptr );
		auto Iters = [&]( int d ){ return sched ? sched[d] : (int)std::ceil( std::pow( prm.iMult , D-d ) * prm.gsIters ); };
		if( sched ) for( int d=lo ; d<=hi ; d++ ) if( sched[d]<0 ) throw std::string( "negative iteration count" );
		if( prm.singleLevel )
		{
			for( int d=lo ; d<hi ; d++ ) prolong2( d , levels[d].x.p , levels[d].y.p , levels[d+1].x.p , levels[d+1].y.p , 0 );
			for( int d=lo ; d<hi ; d++ ) levels[d].x.zero( stream ) , levels[d].y.zero( stream );
			gs( hi , sched ? sched[hi] : prm.gsIters );
		}
		else for( int d=lo ; d<=hi ; d++ )
		{
			gs( d , Iters( d ) );
			if( prm.verbose )
			{
				double e[2];
				energies( d , e );
				printf( "Error[%2d]: %g = %g + %g\n" , 1<<d , e[0]+e[1] , e[0] , e[1] );
			}
			if( d<hi )
			{
				prolong2( d , levels[d].x.p , levels[d].y.p , levels[d+1].x.p , levels[d+1].y.p , 1 );
			}
		}
		double e[2];
		energies( hi , e );
		if( out ) out[0] = e[0] , out[1] = e[1];
		if( prm.verbose ) printf( "Error: %g -> %g = %g + %g => %g\n" , c_ , e[0]+e[1] , e[0] , e[1] , std::sqrt( ( e[0]+e[1] )/c_ ) );
		for( int d=hi ; d<D ; d++ ) prolong2( d , levels[d].x.p , levels[d].y.p , levels[d+1].x.p , levels[d+1].y.p , 0 );
		for( int d=D ; d>0 ; d-- ) prolong2( d-1 , levels[d-1].x.p , levels[d-1].y.p , levels[d].x.p , levels[d].y.p , 2 );
		HIPCHECK( hipGetLastError() );
		HIPCHECK( hipStreamSynchronize( stream ) );
		solved = true;
		seconds[1] = std::chrono::duration< double >( std::chrono::steady_clock::now()-t0 ).count();
	}

	// ---------------- learned initializer ----------------
	DevBuf< double > dFeatE , dFeatN , dFeatS;
	DevBuf< unet::bf16 > dFeatRows;
	DevBuf< float > dFeatF , dNetOut;

	// the network input of level L: the finest target field averaged over each
	// level-L element, then over each node's incident elements, divided by its
	// root mean square (bf16 rows in dFeatRows; floats in out when given)
	void features( int L , float *out )
	{
		needBuilt() ; checkLevel( L );
		long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= 1LL<<L;
		const long long N = levels[L].N;
		dFeatE.ensure( (size_t)ne*CD ) , dFeatN.ensure( (size_t)N*CD ) , dFeatS.ensure( 1 );
		dFeatRows.ensure( (size_t)N*unet::CT );
		if( out ) dFeatF.ensure( (size_t)N*CD );
		KCoarseFieldMean< Dim ><<< Blocks( ne , 256 ) , 256 , 0 , stream >>>( dS.p , D , L , dFeatE.p );
		KNodeFieldMean< Dim ><<< Blocks( N , 256 ) , 256 , 0 , stream >>>( dFeatE.p , L , dFeatN.p );
		KSumSquares<<< 1 , 256 , 0 , stream >>>( dFeatN.p , N*CD , dFeatS.p );
		KFeatureRows<<< Blocks( N*unet::CT , 256 ) , 256 , 0 , stream >>>( dFeatN.p , N , CD , dFeatS.p , dFeatRows.p , out ? dFeatF.p : nullptr );
		HIPCHECK( hipGetLastError() );
		if( out ){ dFeatF.download( out , (size_t)N*CD , stream ) ; HIPCHECK( hipStreamSynchronize( stream ) ); }
	}
	void learnedFeatures( int L , float *out ){ features( L , out ); }

	// The start of level L from the network: its (x, y), reflected (y -> -y)
	// when that raises y^T B x (the energy's only odd term), scaled by the
	// factor a that minimises the fitting energy E_fit(a x, a y) =
	// a^4 W - 2 a^2 y^T B x + c, a^2 = y^T B x / W.  (The screening and
	// Dirichlet terms are left out of the scale: a prediction that is close to
	// the solution in L2 but not zero exactly at the samples pays a screening
	// energy far above the solution's, which would shrink the start towards
	// zero; the cascade's first sweeps remove that residual.)
	// Levels above L are zeroed (the cascade adds into them).  info: a, the
	// reflection (0/1), the energy after scaling.
	void learnedInit( int L , unet::Net *net , double info[3] )
	{
		needBuilt() ; checkLevel( L );
		if( unet::Dim( net )!=Dim || unet::Level( net )!=L || unet::InChannels( net )!=CD ) throw std::string( "the network's dimension, level or input channels do not match this reconstruction" );
		features( L , nullptr );
		Level &Lv = levels[L];
		dNetOut.ensure( (size_t)Lv.N*unet::CT );
		unet::ForwardDevice( net , dFeatRows.p , dNetOut.p , stream );
		KCoeffsFromNet<<< Blocks( Lv.N , 256 ) , 256 , 0 , stream >>>( dNetOut.p , Lv.N , Lv.fm.p , Lv.x.p , Lv.y.p );
		for( int l=L+1 ; l<=D ; l++ ) levels[l].x.zero( stream ) , levels[l].y.zero( stream );
		double t[4];
		energyTerms( L , t );
		bool reflect = false;
		if( t[1]<0 ){ reflect = true ; t[1] = -t[1]; }
		const double a2 = t[1] / t[0];
		if( !( a2>0 ) || !std::isfinite( a2 ) ) throw std::string( "the network's prediction has no fitting-energy-lowering scale; use the random start" );
		const double a = std::sqrt( a2 );
		dTmp2.ensure( 1 );
		dTmp2.upload( &a , 1 , stream );
		KScale2<<< Blocks( Lv.N , 256 ) , 256 , 0 , stream >>>( Lv.x.p , Lv.y.p , Lv.N , dTmp2.p );
		if( reflect ) KScaleArray<<< Blocks( Lv.N , 256 ) , 256 , 0 , stream >>>( Lv.y.p , Lv.N , -1. );
		HIPCHECK( hipGetLastError() );
		double e[2];
		energies( L , e );
		if( info ) info[0] = a , info[1] = reflect ? 1. : 0. , info[2] = e[0] + e[1];
	}

	// the cascade of solve() started at the network's level from its prediction
	void solveLearned( unet::Net *net , double out[2] ){ solveFrom( out , net ); }
	void solveSchedule( unet::Net *net , const int *iters , double out[2] ){ solveFrom( out , net , iters ); }

	void getCoeffs( int l , double *x , double *y )
	{
		needBuilt() ; checkLevel( l );
		if( x ) levels[l].x.download( x , levels[l].N , stream );
		if( y ) levels[l].y.download( y , levels[l].N , stream );
		HIPCHECK( hipStreamSynchronize( stream ) );
	}
	void setCoeffs( int l , const double *x , const double *y )
	{
		needBuilt() ; checkLevel( l );
		if( x ) levels[l].x.upload( x , levels[l].N , stream );
		if( y ) levels[l].y.upload( y , levels[l].N , stream );
		HIPCHECK( hipStreamSynchronize( stream ) );
	}
	void qfits( int l , double *q )
	{
		needBuilt() ; checkLevel( l );
		DevBuf< double > dq;
		dq.alloc( (size_t)levels[l].N*5 );
		KQFit< Dim ><<< Blocks( levels[l].N*C , 256 ) , 256 , 0 , stream >>>( view( l ) , dq.p );
		HIPCHECK( hipGetLastError() );
		dq.download( q , (size_t)levels[l].N*5 , stream );
		HIPCHECK( hipStreamSynchronize( stream ) );
	}
	void apply( int l , char which , const double *x , double *y )
	{
		needBuilt() ; checkLevel( l );
		if( which=='P' )
		{
			if( l>=D ) throw std::string( "no finer level" );
			// through the paired launch the solve uses, x in both slots
			DevBuf< double > a , b , b2;
			a.upload( x , levels[l].N , stream );
			b.alloc( levels[l+1].N ) , b2.alloc( levels[l+1].N );
			prolong2( l , a.p , a.p , b.p , b2.p , 0 );
			b.download( y , levels[l+1].N , stream );
			HIPCHECK( hipStreamSynchronize( stream ) );
			return;
		}
		if( which!='R' && which!='B' ) throw std::string( "operator must be R, B or P" );
		const Level &L = levels[l];
		DevBuf< double > a , b;
		a.upload( x , L.N , stream );
		b.alloc( L.N );
		KApplyCols< Dim ><<< Blocks( L.N , 256 ) , 256 , 0 , stream >>>( which=='R' ? L.R.p : L.B.p , a.p , b.p , view( l ) );
		b.download( y , L.N , stream );
		HIPCHECK( hipStreamSynchronize( stream ) );
	}
	size_t getArray( const std::string &name , double *out , size_t cap )
	{
		std::vector< double > v;
		auto dl = [&]( const DevBuf< double > &b ){ v.resize( b.n ) ; b.download( v.data() , b.n , stream ) ; HIPCHECK( hipStreamSynchronize( stream ) ); };
		if( name=="unit_pos" ) v = pos;
		else if( name=="skew_packed" ) v = skewPacked;
		else if( name=="cells" ) v.assign( cellIdx.begin() , cellIdx.end() );
		else if( name=="cell_counts" ){ for( int c=0 ; c<ncell ; c++ ) v.push_back( cellStart[c+1]-cellStart[c] ); }
		else if( name=="unit_to_world" ){ for( int c=0 ; c<=Dim ; c++ ) for( int r=0 ; r<=Dim ; r++ ) v.push_back( unitToWorld[c][r] ); }
		else if( name=="measure" ){ needBuilt() ; dl( dMeasure ); }
		else if( name=="noise" ){ needBuilt() ; dl( dNoise ); }
		else if( name=="sample_depth" ){ needBuilt() ; dl( dDepth ); }
		else if( name=="S_elem" ){ needBuilt() ; dl( dS ); }
		else if( name=="c" ){ needBuilt() ; v.push_back( c_ ); }
		else if( name.rfind( "adapted_" , 0 )==0 )
		{
			// adapted_{fmask,emask,forder,eorder}_<depth> (dense indices, in the hierarchy's numbering for the orders)
			needBuilt();
			if( !adapted() ) throw std::string( "not an adapted reconstruction" );
			const size_t u = name.rfind( '_' );
			const int l = atoi( name.c_str()+u+1 );
			checkLevel( l );
			const std::string what = name.substr( 8 , u-8 );
			const Level &L = levels[l];
			long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= L.res;
			auto put = [&]( const auto &h ){ v.assign( h.begin() , h.end() ); };
			if( what=="fmask" ) put( download( L.fm , (size_t)L.N ) );
			else if( what=="emask" ) put( download( L.em , (size_t)ne ) );
			else if( what=="forder" ) put( download( L.forder , L.nF ) );
			else if( what=="eorder" ) put( download( L.eorder , L.nE ) );
			else if( what=="act" ) put( download( L.act , L.nActive ) );
			else if( what=="nactive" ) v.push_back( (double)L.nActive );
			else throw std::string( "unknown array: " + name );
		}
		else throw std::string( "unknown array: " + name );
		if( out ) std::copy( v.begin() , v.begin()+std::min( cap , v.size() ) , out );
		return v.size();
	}

	void voxelToWorld( double v2w[Dim+1][Dim+1] ) const
	{
		double g2c[Dim+1][Dim+1];
		for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ) g2c[i][j] = ( i==j ? 1. : 0. );
		for( int d=0 ; d<Dim ; d++ ) g2c[d][d] = 1./(1<<D);
		for( int i=0 ; i<=Dim ; i++ ) for( int j=0 ; j<=Dim ; j++ ){ v2w[i][j] = 0 ; for( int k=0 ; k<=Dim ; k++ ) v2w[i][j] += g2c[i][k]*unitToWorld[k][j]; }
	}

	// .tree (ExteriorPoissonRecon.cpp:259-270; Misha/Streams.h, Misha/EigenStreams.h)
	void writeTree( const char *path , unsigned long long hash )
	{
		if( !solved ) throw std::string( "epr_solve has not run" );
		FILE *fp = fopen( path , "wb" );
		if( !fp ) throw std::string( "cannot open " ) + path;
		unsigned int dim = Dim;
		fwrite( &dim , sizeof(dim) , 1 , fp );
		fwrite( &hash , sizeof(hash) , 1 , fp );
		// XForm< double , Dim+1 > written raw: the vtable pointer of its Algebra
		// base (Misha/Algebra.h:168; meaningless in a file, written as 0), then the
		// (Dim+1)^2 coefficients
		const unsigned long long vptr = 0;
		fwrite( &vptr , sizeof(vptr) , 1 , fp );
		double v2w[Dim+1][Dim+1];
		voxelToWorld( v2w );
		fwrite( v2w , sizeof(double) , (Dim+1)*(Dim+1) , fp );
		for( int pass=0 ; pass<2 ; pass++ )
		{
			unsigned long long nl = D+1;
			fwrite( &nl , sizeof(nl) , 1 , fp );
			for( int d=0 ; d<=D ; d++ )
			{
				std::vector< double > h( levels[d].N );
				( pass==0 ? levels[d].x : levels[d].y ).download( h.data() , h.size() , stream );
				HIPCHECK( hipStreamSynchronize( stream ) );
				if( adapted() )   // the hierarchy's numbering (active functions, then padding)
				{
					const std::vector< long long > &fo = hostHierarchy().levels[d].fOrder;
					std::vector< double > a( fo.size() );
					for( size_t i=0 ; i<a.size() ; i++ ) a[i] = h[ fo[i] ];
					h.swap( a );
				}
				long long sz = (long long)h.size();
				fwrite( &sz , sizeof(sz) , 1 , fp );
				fwrite( h.data() , sizeof(double) , h.size() , fp );
			}
		}
		try
		{
			if( adapted() ) hostHierarchy().write( fp );   // HierarchicalAdaptedIndexer::write (Hat.adaptive.inl:408-414)
			else
			{
				unsigned int maxDepth = D;   // HierarchicalRegularIndexer::write
				fwrite( &maxDepth , sizeof(maxDepth) , 1 , fp );
			}
		}
		catch( ... ){ fclose( fp ) ; throw; }
		fclose( fp );
	}

	// .grid with 2 DOUBLE values per node (Misha/RegularGrid.inl:440-471), finest regular coefficients
	void writeGrid( const char *path )
	{
		if( !solved ) throw std::string( "epr_solve has not run" );
		// regular coefficients: sum of prolonged hierarchical levels
		DevBuf< double > ax , ay , bx , by;
		ax.alloc( levels[D].N ) , ay.alloc( levels[D].N ) , bx.alloc( levels[D].N ) , by.alloc( levels[D].N );
		HIPCHECK( hipMemcpyAsync( ax.p , levels[0].x.p , levels[0].N*sizeof(double) , hipMemcpyDeviceToDevice , stream ) );
		HIPCHECK( hipMemcpyAsync( ay.p , levels[0].y.p , levels[0].N*sizeof(double) , hipMemcpyDeviceToDevice , stream ) );
		for( int d=1 ; d<=D ; d++ )
		{
			prolong2( d-1 , ax.p , ay.p , bx.p , by.p , 0 , true );
			const long long N = levels[d].N;
			KAddInto<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( levels[d].x.p , bx.p , ax.p , N );
			KAddInto<<< Blocks( N , 256 ) , 256 , 0 , stream >>>( levels[d].y.p , by.p , ay.p , N );
		}
		HIPCHECK( hipGetLastError() );
		long long N = levels[D].N;
		std::vector< double > hx( N ) , hy( N );
		ax.download( hx.data() , N , stream ) , ay.download( hy.data() , N , stream );
		HIPCHECK( hipStreamSynchronize( stream ) );
		FILE *fp = fopen( path , "wb" );
		if( !fp ) throw std::string( "cannot open " ) + path;
		fprintf( fp , "G%d\n" , Dim );
		fprintf( fp , "2 DOUBLE\n" );
		for( int d=0 ; d<Dim ; d++ ) fprintf( fp , d ? " %d" : "%d" , ( 1<<D )+1 );
		fprintf( fp , "\n" );
		double v2w[Dim+1][Dim+1];
		voxelToWorld( v2w );
		for( int c=0 ; c<=Dim ; c++ ){ for( int r=0 ; r<=Dim ; r++ ) fprintf( fp , r ? " %f" : "%f" , v2w[c][r] ) ; fprintf( fp , "\n" ); }
		for( long long i=0 ; i<N ; i++ ){ fwrite( &hx[i] , sizeof(double) , 1 , fp ) ; fwrite( &hy[i] , sizeof(double) , 1 , fp ); }
		fclose( fp );
	}

	// <out>.density.tree (ExteriorPoissonRecon.cpp:272-277): the voxel-to-world
	// XForm, then TreeEstimator::write( stream , false ) (GridSamples.inl:734-738):
	// the tree depth first (Misha/RegularTree.inl:614-640: the root, then every
	// brood followed by its children's subtrees) as raw
	// RegularTreeNode< Dim , DensityAndNoiseInfo< Dim > , unsigned short > records
	// (u16 depth and offsets padded to 8 bytes, parent and children pointers --
	// written 0, and 1 for "has children" -- then weightSum, positionSum, the
	// covariance matrix's vtable slot (0) and its Dim^2 entries), then
	// _targetSamplesPerCell = 2^(Dim-CoDim).
	void writeDensity( const char *path )
	{
		if( !built ) throw std::string( "epr_build has not run" );
		// the nodes in file order: a node exists when its parent has children
		std::vector< std::vector< unsigned char > > kids( D+1 );
		for( int t=1 ; t<=D ; t++ )
		{
			kids[t].resize( estKids[t]->n );
			estKids[t]->download( kids[t].data() , kids[t].size() , stream );
		}
		HIPCHECK( hipStreamSynchronize( stream ) );
		struct Rec { unsigned char t ; unsigned short o[Dim] ; bool has; };
		std::vector< Rec > seq;
		auto Has = [&]( int t , const unsigned short *o )
		{
			if( t==0 ) return true;
			if( t>D ) return false;
			long long i = 0;
			for( int d=0 ; d<Dim ; d++ ) i = i*( 1LL<<t ) + o[d];
			return kids[t][i]!=0;
		};
		Rec root{};
		root.has = true;
		seq.push_back( root );
		std::function< void ( const Rec & ) > Children = [&]( const Rec &r )
		{
			if( !r.has ) return;
			const size_t first = seq.size();
			for( int c=0 ; c<( 1<<Dim ) ; c++ )
			{
				Rec k;
				k.t = (unsigned char)( r.t+1 );
				for( int d=0 ; d<Dim ; d++ ) k.o[d] = (unsigned short)( 2*r.o[d] + ( ( c>>d ) & 1 ) );
				k.has = Has( k.t , k.o );
				seq.push_back( k );
			}
			for( int c=0 ; c<( 1<<Dim ) ; c++ ){ const Rec k = seq[ first+c ] ; Children( k ); }
		};
		Children( root );
		// gather the statistics of every node from the dense corner grids, depth by depth
		std::vector< std::vector< long long > > idx( D+2 );
		for( const Rec &r : seq )
		{
			if( r.t<1 ) continue;
			const long long rr = 1LL<<( r.t-1 );
			long long i = 0;
			for( int d=0 ; d<Dim && i>=0 ; d++ ) i = r.o[d]<=rr ? i*( rr+1 ) + r.o[d] : -1;
			idx[ r.t ].push_back( i );
		}
		std::vector< std::vector< double > > rows( D+2 );
		DevBuf< long long > dIdx;
		DevBuf< double > dRows;
		for( int t=1 ; t<=D+1 ; t++ )
		{
			const size_t m = idx[t].size();
			if( !m ) continue;
			dIdx.upload( idx[t].data() , m , stream );
			dRows.ensure( m*NV );
			KGatherRows<<< Blocks( (long long)m , 256 ) , 256 , 0 , stream >>>( estData[t]->p , dIdx.p , (long long)m , NV , dRows.p );
			HIPCHECK( hipGetLastError() );
			rows[t].resize( m*NV );
			dRows.download( rows[t].data() , m*NV , stream );
			HIPCHECK( hipStreamSynchronize( stream ) );
		}
		FILE *fp = fopen( path , "wb" );
		if( !fp ) throw std::string( "cannot open " ) + path;
		const unsigned long long zero = 0;
		fwrite( &zero , sizeof(zero) , 1 , fp );
		double v2w[Dim+1][Dim+1];
		voxelToWorld( v2w );
		fwrite( v2w , sizeof(double) , (Dim+1)*(Dim+1) , fp );
		constexpr size_t head = ( ( 2*( Dim+1 ) + 7 )/8 )*8 , rec = head + 16 + 8*( 2 + Dim + Dim*Dim );
		std::vector< size_t > cur( D+2 , 0 );
		std::vector< unsigned char > buf( rec );
		for( const Rec &r : seq )
		{
			memset( buf.data() , 0 , rec );
			unsigned short h[Dim+1];
			h[0] = r.t;
			for( int d=0 ; d<Dim ; d++ ) h[1+d] = r.o[d];
			memcpy( buf.data() , h , sizeof(h) );
			const unsigned long long ch = r.has ? 1 : 0;
			memcpy( buf.data() + head + 8 , &ch , 8 );
			if( r.t>=1 )
			{
				const double *v = &rows[ r.t ][ ( cur[ r.t ]++ )*NV ];
				double *o = (double *)( buf.data() + head + 16 );
				o[0] = v[0];
				for( int d=0 ; d<Dim ; d++ ) o[1+d] = v[1+d];
				for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) o[ 2+Dim+i*Dim+j ] = v[ 1+Dim+( i>j ? j*Dim+i : i*Dim+j ) ];   // lower covariance mirrored (EstLower)
			}
			fwrite( buf.data() , 1 , rec , fp );
		}
		const double target = std::pow( 2. , Dim-CoDim );
		fwrite( &target , sizeof(double) , 1 , fp );
		fclose( fp );
	}
};

// ---------------- many reconstructions ----------------
// Builds every context (Ctx::build) from up to `threads` host threads, each
// context's kernels on its own stream; the first failure is rethrown after all
// threads end, prefixed with the failing context's position.
inline void BuildBatch( const std::vector< CtxBase * > &cs , int threads )
{
	if( cs.empty() ) return;
	int dev = 0;
	HIPCHECK( hipGetDevice( &dev ) );
	// default 4: the runtime's hardware queues per process (GPU_MAX_HW_QUEUES, 4 by
	// default); on MI355X more threads measured slower (256 builds: 0.11 s with
	// 4 threads, 0.14 s with 16)
	const size_t nt = std::min< size_t >( threads>0 ? (size_t)threads : 4 , cs.size() );
	std::atomic< size_t > next( 0 );
	std::atomic< bool > failed( false );
	std::mutex m;
	std::exception_ptr first;
	auto fail = [&]( std::exception_ptr e ){ std::lock_guard< std::mutex > lock( m ) ; if( !first ) first = e ; failed = true; };
	auto work = [&]( void )
	{
		if( hipSetDevice( dev )!=hipSuccess ){ fail( std::make_exception_ptr( HipError{ "hipSetDevice failed in a build thread" } ) ) ; return; }
		for( size_t i ; !failed && ( i = next++ )<cs.size() ; )
		{
			const std::string where = "context " + std::to_string( i ) + ": ";
			try { cs[i]->build(); }
			catch( const HipError &e ){ fail( std::make_exception_ptr( HipError{ where + e.msg } ) ); }
			catch( const std::string &s ){ fail( std::make_exception_ptr( where + s ) ); }
			catch( const std::exception &e ){ fail( std::make_exception_ptr( where + e.what() ) ); }
		}
	};
	std::vector< std::thread > ts;
	for( size_t t=1 ; t<nt ; t++ ) ts.emplace_back( work );
	work();
	for( auto &t : ts ) t.join();
	if( first ) std::rethrow_exception( first );
}

// ---------------- the batched solve ----------------
// The cascadic schedule of ExteriorPoisson.h:316-409 for many built contexts
// of one dimension.  Contexts of equal depth form a group (their cascades
// have the same number of stages); level by level, the sweeps of every
// context of a group that solves at that level share launches: on the levels
// swept by one-workgroup kernels one workgroup per context runs all of that
// context's iterations (the latency-bound coarse levels of different contexts
// run side by side on different CUs); on the others each colour phase is one
// grid over that colour's nodes in every context still iterating.  Fills,
// the random start (its draws made on the host, scattered and rescaled on the
// device), prolongations and the hierarchical split are batched the same way.
// With --fp32 a context's colour-launch levels run all iterations but the
// last on the float copies, as its own solve does.  Every context goes
// through the kernels and per-node arithmetic of its own Ctx::solve, in the
// same order, so its coefficients and energies are the same bit for bit.  The
// --verbose per-level energies are not printed.
//
// The device work of all groups -- each group on its own stream, the groups
// side by side, so that the throughput-bound fine levels of the shallower
// groups fill the GPU while the deepest group's latency-bound coarse levels
// run -- is enqueued without a host synchronisation and captured once into a
// HIP graph; later calls on the same contexts replay it with one
// hipGraphLaunch, as long as nothing it depends on (the buffers, counts,
// iteration numbers, parameters and plans of every level of every context,
// summarised in a signature) has changed.  Only the target-field constants c
// are uploaded per call.

// host-side job tables of a batched solve, referenced by offset
struct BatchHost
{
	std::vector< char > table;
	std::vector< double > draws;
	size_t partials = 0;
	size_t put( const void *p , size_t bytes ){ const size_t o = ( table.size() + 15 ) & ~(size_t)15 ; table.resize( o + bytes ) ; if( bytes ) memcpy( table.data()+o , p , bytes ) ; return o; }
};
// the device copies and the per-call arrays
struct BatchDevPtrs
{
	const char *table = nullptr;
	const double *draws = nullptr , *c = nullptr;
	double *partials = nullptr , *terms = nullptr;
};

template< int Dim >
struct BatchGroup
{
	using Cx = Ctx< Dim >;
	using Lv = typename Cx::Level;
	static constexpr int C = DimC< Dim >::C;
	struct Jobs { size_t off = 0 , count = 0; };
	struct Stage
	{
		typename Cx::GsPlan plan = { -1 , 256 , 0 , 0 };
		int floatThreads = 256;
		Jobs all;                              // one-workgroup kernels: every context sweeping here
		Jobs dense , dense32 , adapted , lists;  // colour launches: by decreasing iteration count
		Jobs narrow;                           // --fp32: x, y of the contexts with float iterations
		std::vector< int > denseIters , dense32Iters , adaptedIters;
		std::vector< std::vector< long long > > listCount;   // [context of `adapted`][colour]
		std::vector< size_t > smallRs;         // GS_SMALL: the contexts, launched one by one
		Jobs prolongAdd , prolongSet , zeroAfter;
	};
	std::vector< Cx * > cs;
	std::vector< int > ids;                    // positions in the whole batch (c, terms)
	std::vector< int > lo , hi;
	int maxD = 0;
	long long Nd[ MAXT ];
	std::vector< Jobs > zeroAll , splitSet , splitSub;
	Jobs start , scale;
	long long maxStart = 1;
	std::vector< size_t > partOff;             // each context's partial sums in the shared buffer
	std::vector< Stage > stages;

	int iters( size_t r , int d ) const
	{
		const Cx &c = *cs[r];
		if( d<lo[r] || d>hi[r] ) return 0;
		if( c.prm.singleLevel ) return d==hi[r] ? c.prm.gsIters : 0;
		return (int)std::ceil( std::pow( c.prm.iMult , c.D-d ) * c.prm.gsIters );
	}
	// the context's float iterations at a colour-launch level (all but the last, as Ctx::gs)
	int floatIters( size_t r , int d ) const
	{
		const int it = iters( r , d );
		return ( cs[r]->floatLevel( d ) && it>1 ) ? it-1 : 0;
	}
	Lv &lev( size_t r , int d ) const { return cs[r]->levels[d]; }

	void plan( const std::vector< Cx * > &_cs , const std::vector< int > &_ids , BatchHost &H )
	{
		cs = _cs , ids = _ids;
		const size_t n = cs.size();
		lo.resize( n ) , hi.resize( n );
		for( size_t r=0 ; r<n ; r++ ) lo[r] = cs[r]->minD() , hi[r] = cs[r]->maxD() , maxD = std::max( maxD , cs[r]->D );
		for( int d=0 ; d<=maxD ; d++ ){ Nd[d] = 1 ; for( int k=0 ; k<Dim ; k++ ) Nd[d] *= ( 1LL<<d ) + 1; }
		auto zeroJobs = [&]( int d , const std::vector< size_t > &rs )
		{
			std::vector< BatchZero > z;
			for( size_t r : rs ){ Lv &L = lev( r , d ) ; z.push_back( { L.x.p , L.N } ) , z.push_back( { L.y.p , L.N } ); }
			return Jobs{ H.put( z.data() , z.size()*sizeof(BatchZero) ) , z.size() };
		};
		auto prolongJobs = [&]( int d , const std::vector< size_t > &rs , int mode )
		{
			std::vector< BatchProlong > j;
			for( size_t r : rs )
			{
				Lv &Lc = lev( r , d ) , &Lf = lev( r , d+1 );
				j.push_back( { Lc.x.p , Lf.x.p , Lf.fm.p , mode } ) , j.push_back( { Lc.y.p , Lf.y.p , Lf.fm.p , mode } );
			}
			return Jobs{ H.put( j.data() , j.size()*sizeof(BatchProlong) ) , j.size() };
		};
		auto levelJobs = [&]( int d , const std::vector< size_t > &rs , const std::vector< int > &its )
		{
			std::vector< BatchLevel > b;
			for( size_t q=0 ; q<rs.size() ; q++ )
			{
				const size_t r = rs[q];
				Lv &L = lev( r , d );
				BatchLevel e;
				e.x = L.x.p , e.y = L.y.p , e.R = L.R.p , e.B = L.B.p , e.K = L.K.p , e.fmask = L.fm.p;
				e.dW = cs[r]->prm.dWeight , e.separate = cs[r]->prm.separate , e.iters = its[q];
				b.push_back( e );
			}
			return Jobs{ H.put( b.data() , b.size()*sizeof(BatchLevel) ) , b.size() };
		};
		auto floatJobs = [&]( int d , const std::vector< size_t > &rs , const std::vector< int > &its )
		{
			std::vector< BatchLevelT< float > > b;
			for( size_t q=0 ; q<rs.size() ; q++ )
			{
				const size_t r = rs[q];
				Lv &L = lev( r , d );
				BatchLevelT< float > e;
				e.x = L.fx.p , e.y = L.fy.p , e.R = L.fR.p , e.B = L.fB.p , e.K = L.fK.p , e.fmask = nullptr;
				e.dW = (float)cs[r]->prm.dWeight , e.separate = cs[r]->prm.separate , e.iters = its[q];
				b.push_back( e );
			}
			return Jobs{ H.put( b.data() , b.size()*sizeof(BatchLevelT< float >) ) , b.size() };
		};

		// fills of every level
		zeroAll.resize( maxD+1 );
		for( int d=0 ; d<=maxD ; d++ )
		{
			std::vector< size_t > rs;
			for( size_t r=0 ; r<n ; r++ ) if( d<=cs[r]->D ) rs.push_back( r );
			zeroAll[d] = zeroJobs( d , rs );
		}
		// the random start of Ctx::init (ExteriorPoisson.h:333-341): the draws of
		// each context's private rand() sequence, in the order they are assigned
		// (the hierarchy's numbering of an adapted level, function order on a dense one)
		{
			std::vector< BatchStart > st;
			std::vector< BatchScale > sc;
			for( size_t r=0 ; r<n ; r++ )
			{
				Cx &c = *cs[r];
				Lv &L = lev( r , lo[r] );
				const long long cnt = c.adapted() ? (long long)L.nF : L.N;
				GlibcRand g( (unsigned int)c.prm.seed );
				const long long off = (long long)H.draws.size();
				for( long long k=0 ; k<2*cnt ; k++ ) H.draws.push_back( g.random01() );
				st.push_back( { off , L.x.p , L.y.p , c.adapted() ? L.forder.p : nullptr , cnt } );
				sc.push_back( { L.x.p , L.y.p , L.N , ids[r] } );
				maxStart = std::max( maxStart , L.N );
			}
			start = Jobs{ H.put( st.data() , st.size()*sizeof(BatchStart) ) , st.size() };
			scale = Jobs{ H.put( sc.data() , sc.size()*sizeof(BatchScale) ) , sc.size() };
		}
		// partial sums of the energy terms (the start's and the final level's)
		partOff.resize( n );
		for( size_t r=0 ; r<n ; r++ )
		{
			const long long nb = std::max( EnergyRows< Dim >( lev( r , lo[r] ).N , lev( r , lo[r] ).res ) , EnergyRows< Dim >( lev( r , hi[r] ).N , lev( r , hi[r] ).res ) );
			partOff[r] = H.partials , H.partials += (size_t)nb*4;
		}

		// per level: the sweeps, then the prolongation to the next level
		stages.resize( maxD+1 );
		for( int d=0 ; d<=maxD ; d++ )
		{
			Stage &S = stages[d];
			std::vector< size_t > rs;
			for( size_t r=0 ; r<n ; r++ ) if( d<=cs[r]->D && iters( r , d )>0 ) rs.push_back( r );
			if( !rs.empty() )
			{
				S.plan = cs[ rs[0] ]->gsPlan( d , true );
				S.floatThreads = cs[ rs[0] ]->floatThreads( d , true );
				if( S.plan.kind==Cx::GS_PHASES )
				{
					// FP64 iterations: all of a context's, or the last one after its float iterations
					auto dIters = [&]( size_t r ){ return floatIters( r , d ) ? 1 : iters( r , d ); };
					std::vector< size_t > rd , ra , rf;
					for( size_t r : rs ) ( cs[r]->adapted() ? ra : rd ).push_back( r );
					std::stable_sort( rd.begin() , rd.end() , [&]( size_t a , size_t b ){ return dIters( a )>dIters( b ); } );
					std::stable_sort( ra.begin() , ra.end() , [&]( size_t a , size_t b ){ return iters( a , d )>iters( b , d ); } );
					for( size_t r : rd ) S.denseIters.push_back( dIters( r ) );
					for( size_t r : ra ) S.adaptedIters.push_back( iters( r , d ) );
					S.dense = levelJobs( d , rd , S.denseIters ) , S.adapted = levelJobs( d , ra , S.adaptedIters );
					for( size_t r : rd ) if( floatIters( r , d ) ) rf.push_back( r );
					std::stable_sort( rf.begin() , rf.end() , [&]( size_t a , size_t b ){ return floatIters( a , d )>floatIters( b , d ); } );
					for( size_t r : rf ) S.dense32Iters.push_back( floatIters( r , d ) );
					S.dense32 = floatJobs( d , rf , S.dense32Iters );
					{
						std::vector< BatchCast > nc;
						for( size_t r : rf ){ Lv &L = lev( r , d ) ; nc.push_back( { L.x.p , L.fx.p , L.N } ) , nc.push_back( { L.y.p , L.fy.p , L.N } ); }
						S.narrow = Jobs{ H.put( nc.data() , nc.size()*sizeof(BatchCast) ) , nc.size() };
					}
					std::vector< BatchList > ls;
					for( size_t r : ra )
					{
						Lv &L = lev( r , d );
						S.listCount.emplace_back( C );
						for( int c=0 ; c<C ; c++ ){ const long long cnt = L.actStart[c+1]-L.actStart[c] ; ls.push_back( { L.act.p + L.actStart[c] , cnt } ) ; S.listCount.back()[c] = cnt; }
					}
					S.lists = Jobs{ H.put( ls.data() , ls.size()*sizeof(BatchList) ) , ls.size() };
				}
				else if( S.plan.kind==Cx::GS_SMALL ) S.smallRs = rs;
				else
				{
					std::vector< int > its;
					for( size_t r : rs ) its.push_back( iters( r , d ) );
					S.all = levelJobs( d , rs , its );
				}
			}
			if( d<maxD )
			{
				std::vector< size_t > add , set;
				for( size_t r=0 ; r<n ; r++ ) if( d>=lo[r] && d<hi[r] ) ( cs[r]->prm.singleLevel ? set : add ).push_back( r );
				S.prolongAdd = prolongJobs( d , add , 1 ) , S.prolongSet = prolongJobs( d , set , 0 );
				S.zeroAfter = zeroJobs( d , set );   // --singleLevel: the coarser levels end at zero
			}
		}
		// the hierarchical split after the solve: levels above the last solved one
		// are prolongations (=), then every level keeps only its own detail (-=)
		splitSet.resize( maxD+1 ) , splitSub.resize( maxD+1 );
		for( int d=0 ; d<maxD ; d++ )
		{
			std::vector< size_t > rs;
			for( size_t r=0 ; r<n ; r++ ) if( d>=hi[r] && d<cs[r]->D ) rs.push_back( r );
			splitSet[d] = prolongJobs( d , rs , 0 );
		}
		for( int d=maxD ; d>0 ; d-- )
		{
			std::vector< size_t > rs;
			for( size_t r=0 ; r<n ; r++ ) if( d<=cs[r]->D ) rs.push_back( r );
			splitSub[d] = prolongJobs( d-1 , rs , 2 );
		}
	}

	// the sweeps of level d of every context that solves there
	void enqueueSweeps( hipStream_t st , const BatchDevPtrs &P , int d ) const
	{
		auto at = [&]( const Jobs &j ){ return (const void *)( P.table + j.off ); };
		{
			const Stage &S = stages[d];
			// any context with this level gives the shared constants (resolution, strides, layout, wedge scale)
			size_t rAny = 0;
			while( cs[rAny]->D<d ) rAny++;
			const LevelView< Dim > L0 = cs[rAny]->view( d );
			if( S.all.count )
			{
				const unsigned g = (unsigned)S.all.count;
				const BatchLevel *J = (const BatchLevel *)at( S.all );
				if( S.plan.kind==Cx::GS_WAVE )
				{
					if constexpr( Dim==3 ){ if( L0.res==1 ) KGsWaveBatch< 1 ><<< g , 64 , 0 , st >>>( L0 , J ) ; else KGsWaveBatch< 2 ><<< g , 64 , 0 , st >>>( L0 , J ); }
				}
				else if( S.plan.kind==Cx::GS_BLOCK ){ if constexpr( Dim==3 ) KGsBlockBatch<<< g , 512 , S.plan.lds , st >>>( L0 , J ); }
				else if( S.plan.kind==Cx::GS_TINY )
				{
					if( S.plan.rb ) KGsTinyBatch< Dim , true ><<< g , S.plan.threads , S.plan.lds , st >>>( L0 , J );
					else KGsTinyBatch< Dim , false ><<< g , S.plan.threads , S.plan.lds , st >>>( L0 , J );
				}
				else KGsSmallLdsBatch< Dim ><<< g , S.plan.threads , S.plan.lds , st >>>( L0 , J , S.plan.rb );
			}
			for( size_t r : S.smallRs ) KGsSmall< Dim ><<< 1 , 512 , 0 , st >>>( cs[r]->view( d ) , iters( r , d ) );
			const int th = S.plan.threads;
			const bool idx32 = L0.N*DimC< Dim >::W < ( 1LL<<31 ) , b32 = L0.N*(long long)sizeof(double) < ( 1LL<<32 );
			if( S.dense32.count )
			{
				// --fp32: x, y rounded, the float iterations, x, y widened (Ctx::gsFloat)
				const unsigned g = (unsigned)std::min< long long >( Blocks( Nd[d] , 256 ) , 1024 );
				KNarrowBatch<<< dim3( g , (unsigned)S.narrow.count ) , 256 , 0 , st >>>( (const BatchCast *)at( S.narrow ) );
				LevelView< Dim , float > F0;
				F0.res = L0.res , F0.N = L0.N , F0.x = F0.y = nullptr , F0.R = F0.B = F0.K = nullptr;
				F0.sW = (float)L0.sW , F0.dW = 0 , F0.lay = L0.lay , F0.fmask = nullptr , F0.separate = 0;
				for( int k=0 ; k<Dim ; k++ ) F0.stride[k] = L0.stride[k];
				const BatchLevelT< float > *J = (const BatchLevelT< float > *)at( S.dense32 );
				size_t k = S.dense32.count;
				for( int it=0 ; it<S.dense32Iters[0] ; it++ )
				{
					while( k && S.dense32Iters[k-1]<=it ) k--;
					for( int c=0 ; c<C ; c++ )
					{
						ColourSet< Dim > cset ; cset.set( 1<<d , c );
						const dim3 grid( Blocks( cset.total*C , S.floatThreads ) , (unsigned)k );
						if( b32 ){ ColourDiv< Dim > cdiv ; cdiv.set( cset ) ; KGsPhaseBatch32< Dim , float ><<< grid , S.floatThreads , 0 , st >>>( F0 , J , cdiv , L0.lay.cstart[c] ); }
						else if( idx32 ) KGsPhaseBatch< Dim , float , int ><<< grid , S.floatThreads , 0 , st >>>( F0 , J , cset , L0.lay.cstart[c] );
						else KGsPhaseBatch< Dim , float , long long ><<< grid , S.floatThreads , 0 , st >>>( F0 , J , cset , L0.lay.cstart[c] );
					}
				}
				KWidenBatch<<< dim3( g , (unsigned)S.narrow.count ) , 256 , 0 , st >>>( (const BatchCast *)at( S.narrow ) );
			}
			if( S.dense.count )
			{
				const BatchLevel *J = (const BatchLevel *)at( S.dense );
				size_t k = S.dense.count;
				for( int it=0 ; it<S.denseIters[0] ; it++ )
				{
					while( k && S.denseIters[k-1]<=it ) k--;
					for( int c=0 ; c<C ; c++ )
					{
						ColourSet< Dim > cset ; cset.set( 1<<d , c );
						const dim3 grid( Blocks( cset.total*C , th ) , (unsigned)k );
						if( b32 ){ ColourDiv< Dim > cdiv ; cdiv.set( cset ) ; KGsPhaseBatch32< Dim , double ><<< grid , th , 0 , st >>>( L0 , J , cdiv , L0.lay.cstart[c] ); }
						else if( idx32 ) KGsPhaseBatch< Dim , double , int ><<< grid , th , 0 , st >>>( L0 , J , cset , L0.lay.cstart[c] );
						else KGsPhaseBatch< Dim , double , long long ><<< grid , th , 0 , st >>>( L0 , J , cset , L0.lay.cstart[c] );
					}
				}
			}
			if( S.adapted.count )
			{
				size_t k = S.adapted.count;
				for( int it=0 ; it<S.adaptedIters[0] ; it++ )
				{
					while( k && S.adaptedIters[k-1]<=it ) k--;
					for( int c=0 ; c<C ; c++ )
					{
						long long mx = 0;
						for( size_t q=0 ; q<k ; q++ ) mx = std::max( mx , S.listCount[q][c] );
						if( mx ) KGsPhaseListBatch< Dim ><<< dim3( Blocks( mx*C , th ) , (unsigned)k ) , th , 0 , st >>>( L0 , (const BatchLevel *)at( S.adapted ) , (const BatchList *)at( S.lists ) , c );
					}
				}
			}
		}
		HIPCHECK( hipGetLastError() );
	}

	// Every device step of the group's solve, on stream st, with no host
	// synchronisation (so that it can be captured into a graph).
	void enqueue( hipStream_t st , const BatchDevPtrs &P ) const
	{
		const size_t n = cs.size();
		auto at = [&]( const Jobs &j ){ return (const void *)( P.table + j.off ); };
		auto launchZero = [&]( const Jobs &j , int d )
		{
			if( j.count ) KZeroBatch<<< dim3( (unsigned)std::min< long long >( Blocks( Nd[d] , 256 ) , 1024 ) , (unsigned)j.count ) , 256 , 0 , st >>>( (const BatchZero *)at( j ) );
		};
		auto launchProlong = [&]( const Jobs &j , int d )
		{
			if( j.count ) KProlongBatch< Dim ><<< dim3( Blocks( Nd[d+1] , 256 ) , (unsigned)j.count ) , 256 , 0 , st >>>( (const BatchProlong *)at( j ) , 1<<d );
		};
		// (W, yBx, R-form, K-form) of every context at its level into terms[4 id] (Ctx::energyTerms)
		auto terms = [&]( const std::vector< int > &level )
		{
			for( size_t r=0 ; r<n ; r++ )
			{
				LaunchEnergy< Dim >( cs[r]->view( level[r] ) , P.partials + partOff[r] , P.terms + 4*ids[r] , st );
			}
		};
		// ---- zero, random start, rescale s = ( c / W(x,y) )^(1/4)
		for( int d=0 ; d<=maxD ; d++ ) launchZero( zeroAll[d] , d );
		KStartBatch<<< dim3( (unsigned)std::min< long long >( Blocks( maxStart , 256 ) , 1024 ) , (unsigned)start.count ) , 256 , 0 , st >>>( (const BatchStart *)at( start ) , P.draws );
		terms( lo );
		KRescaleBatch<<< dim3( (unsigned)std::min< long long >( Blocks( maxStart , 256 ) , 1024 ) , (unsigned)scale.count ) , 256 , 0 , st >>>( (const BatchScale *)at( scale ) , P.c , P.terms );
		// ---- the levels
		for( int d=0 ; d<=maxD ; d++ )
		{
			const Stage &S = stages[d];
			enqueueSweeps( st , P , d );
			if( d<maxD )
			{
				launchProlong( S.prolongAdd , d );
				launchProlong( S.prolongSet , d );
				launchZero( S.zeroAfter , d );
			}
		}
		// ---- final energy terms (Ctx::energies), then the hierarchical split
		terms( hi );
		for( int d=0 ; d<maxD ; d++ ) launchProlong( splitSet[d] , d );
		for( int d=maxD ; d>0 ; d-- ) launchProlong( splitSub[d] , d-1 );
		HIPCHECK( hipGetLastError() );
	}
};

// Everything a batched solve's recorded device work depends on: the contexts,
// their parameters, and for every level the buffers, counts and plans.
template< int Dim >
std::vector< long long > BatchSignature( const std::vector< Ctx< Dim > * > &cs )
{
	constexpr int C = DimC< Dim >::C;
	std::vector< long long > s;
	auto bits = []( double v ){ long long b ; memcpy( &b , &v , sizeof(b) ) ; return b; };
	auto ptr = []( const void *p ){ return (long long)(uintptr_t)p; };
	s.push_back( (long long)cs.size() );
	for( const Ctx< Dim > *c : cs )
	{
		const epr_params &p = c->prm;
		for( long long v : { ptr( c ) , (long long)c->D , (long long)c->minD() , (long long)c->maxD() , (long long)p.seed , (long long)p.gsIters , bits( p.iMult ) , (long long)p.singleLevel ,
		                     (long long)p.separate , bits( p.dWeight ) , (long long)p.fp32Sweeps , (long long)p.aRadius } ) s.push_back( v );
		for( int d=0 ; d<=c->D ; d++ )
		{
			const auto &L = c->levels[d];
			for( const void *q : { (const void *)L.x.p , (const void *)L.y.p , (const void *)L.R.p , (const void *)L.B.p , (const void *)L.K.p , (const void *)L.fm.p , (const void *)L.act.p , (const void *)L.forder.p ,
			                       (const void *)L.fx.p , (const void *)L.fy.p , (const void *)L.fR.p , (const void *)L.fB.p , (const void *)L.fK.p } ) s.push_back( ptr( q ) );
			s.push_back( L.N ) , s.push_back( (long long)L.nF );
			for( int k=0 ; k<=C ; k++ ) s.push_back( L.actStart[k] );
			const auto pl = c->gsPlan( d , true );
			s.push_back( pl.kind ) , s.push_back( pl.threads ) , s.push_back( pl.rb ) , s.push_back( (long long)pl.lds ) , s.push_back( c->floatThreads( d , true ) );
		}
	}
	return s;
}

// A recorded batched solve: the plans, their device tables, the streams and the graph.
template< int Dim >
struct BatchExec
{
	std::vector< long long > sig;
	std::vector< BatchGroup< Dim > > groups;
	DevBuf< char > table;
	DevBuf< double > draws , partials , terms , c;
	double *hC = nullptr , *hT = nullptr;      // pinned: the per-call c values and the energy terms
	size_t n = 0;
	hipStream_t origin = nullptr;
	std::vector< hipStream_t > streams;
	std::vector< hipEvent_t > events;          // [0]: fork, [1..]: one join per group
	hipGraphExec_t graph = nullptr;
	unsigned long long uses = 0;
	std::mutex busy;
	~BatchExec()
	{
		if( origin ) (void)hipStreamSynchronize( origin );   // the last replay has ended
		if( graph ) (void)hipGraphExecDestroy( graph );
		for( hipEvent_t e : events ) (void)hipEventDestroy( e );
		for( hipStream_t s : streams ) (void)hipStreamDestroy( s );
		if( origin ) (void)hipStreamDestroy( origin );
		if( hC ) (void)hipHostFree( hC );
		if( hT ) (void)hipHostFree( hT );
	}
	// all groups: forked from the origin stream, each on its own, joined back
	void enqueueAll( void )
	{
		BatchDevPtrs P;
		P.table = table.p , P.draws = draws.p , P.c = c.p , P.partials = partials.p , P.terms = terms.p;
		HIPCHECK( hipEventRecord( events[0] , origin ) );
		for( size_t k=0 ; k<groups.size() ; k++ )
		{
			HIPCHECK( hipStreamWaitEvent( streams[k] , events[0] , 0 ) );
			groups[k].enqueue( streams[k] , P );
			HIPCHECK( hipEventRecord( events[k+1] , streams[k] ) );
			HIPCHECK( hipStreamWaitEvent( origin , events[k+1] , 0 ) );
		}
	}
	void setup( const std::vector< Ctx< Dim > * > &cs , std::vector< long long > &&s , bool useGraph )
	{
		sig = std::move( s );
		n = cs.size();
		std::vector< std::vector< Ctx< Dim > * > > gc( MAXT );
		std::vector< std::vector< int > > gi( MAXT );
		for( size_t r=0 ; r<n ; r++ ) gc[ cs[r]->D ].push_back( cs[r] ) , gi[ cs[r]->D ].push_back( (int)r );
		BatchHost H;
		for( int d=0 ; d<MAXT ; d++ ) if( !gc[d].empty() ){ groups.emplace_back() ; groups.back().plan( gc[d] , gi[d] , H ); }
		HIPCHECK( hipStreamCreate( &origin ) );
		table.alloc( std::max< size_t >( H.table.size() , 16 ) ) , draws.alloc( std::max< size_t >( H.draws.size() , 1 ) );
		partials.alloc( std::max< size_t >( H.partials , 1 ) ) , terms.alloc( 4*n ) , c.alloc( n );
		HIPCHECK( hipMemcpyAsync( table.p , H.table.data() , H.table.size() , hipMemcpyHostToDevice , origin ) );
		if( !H.draws.empty() ) HIPCHECK( hipMemcpyAsync( draws.p , H.draws.data() , H.draws.size()*sizeof(double) , hipMemcpyHostToDevice , origin ) );
		HIPCHECK( hipStreamSynchronize( origin ) );
		HIPCHECK( hipHostMalloc( (void **)&hC , n*sizeof(double) , hipHostMallocDefault ) );
		HIPCHECK( hipHostMalloc( (void **)&hT , 4*n*sizeof(double) , hipHostMallocDefault ) );
		streams.resize( groups.size() , nullptr ) , events.resize( groups.size()+1 , nullptr );
		for( auto &s : streams ) HIPCHECK( hipStreamCreate( &s ) );
		for( auto &e : events ) HIPCHECK( hipEventCreateWithFlags( &e , hipEventDisableTiming ) );
		if( useGraph )
		{
			hipGraph_t g;
			HIPCHECK( hipStreamBeginCapture( origin , hipStreamCaptureModeThreadLocal ) );
			try { enqueueAll(); }
			catch( ... ){ hipGraph_t dropped ; (void)hipStreamEndCapture( origin , &dropped ) ; if( dropped ) (void)hipGraphDestroy( dropped ) ; throw; }
			HIPCHECK( hipStreamEndCapture( origin , &g ) );
			const hipError_t e = hipGraphInstantiate( &graph , g , nullptr , nullptr , 0 );
			(void)hipGraphDestroy( g );
			HIPCHECK( e );
		}
	}
	// one solve: the contexts' c, the recorded work, the energy terms back
	void run( const std::vector< Ctx< Dim > * > &cs , double *energies )
	{
		for( size_t r=0 ; r<n ; r++ ) hC[r] = cs[r]->c_;
		HIPCHECK( hipMemcpyAsync( c.p , hC , n*sizeof(double) , hipMemcpyHostToDevice , origin ) );
		if( graph ) HIPCHECK( hipGraphLaunch( graph , origin ) );
		else enqueueAll();
		HIPCHECK( hipMemcpyAsync( hT , terms.p , 4*n*sizeof(double) , hipMemcpyDeviceToHost , origin ) );
		HIPCHECK( hipStreamSynchronize( origin ) );
		for( size_t r=0 ; r<n ; r++ )
		{
			const double *t = hT + 4*r;
			double e0 = t[0] - 2.*t[1] + cs[r]->c_ , e1 = t[2];
			if( cs[r]->prm.dWeight ) e1 += t[3]*cs[r]->prm.dWeight;
			if( energies ) energies[2*r] = e0 , energies[2*r+1] = e1;
		}
		uses++;
	}
};

// The recorded solves of the process (per dimension), the most recently used
// first; a few are kept (each holds its tables, streams and graph).
template< int Dim >
struct BatchExecCache
{
	static constexpr size_t Keep = 4;
	std::mutex m;
	std::vector< std::shared_ptr< BatchExec< Dim > > > list;
	static BatchExecCache &Get( void ){ static BatchExecCache *c = new BatchExecCache ; return *c; }
	std::shared_ptr< BatchExec< Dim > > find( const std::vector< long long > &sig , bool useGraph )
	{
		std::lock_guard< std::mutex > lock( m );
		for( size_t i=0 ; i<list.size() ; i++ ) if( list[i]->sig==sig && ( list[i]->graph!=nullptr )==useGraph )
		{
			std::shared_ptr< BatchExec< Dim > > e = list[i];
			list.erase( list.begin()+i ) , list.insert( list.begin() , e );
			return e;
		}
		return nullptr;
	}
	void add( const std::shared_ptr< BatchExec< Dim > > &e )
	{
		std::lock_guard< std::mutex > lock( m );
		list.insert( list.begin() , e );
		while( list.size()>Keep ) list.pop_back();
	}
	void clear( void ){ std::lock_guard< std::mutex > lock( m ) ; list.clear(); }
};

// The batched solve (see above).  EPR_BATCH_GRAPH=0 enqueues the same work
// directly on every call instead of recording and replaying it (for tools
// that do not follow graph launches).
template< int Dim >
void SolveBatch( const std::vector< Ctx< Dim > * > &cs , double *energies )
{
	using Cx = Ctx< Dim >;
	const size_t n = cs.size();
	if( !n ) return;
	// grid.y of the fill, copy and prolongation launches is twice the contexts of a level (x and y), at most 65535
	if( n>32767 ) throw std::string( "at most 32767 contexts per batch" );
	{
		std::set< const Cx * > seen;
		for( const Cx *c : cs ) if( !seen.insert( c ).second ) throw std::string( "a context appears twice in the batch" );
	}
	for( Cx *c : cs ) c->needBuilt();
	auto t0 = std::chrono::steady_clock::now();
	// the float copies of the operators of every level that sweeps in FP32 (made by the build; again after a re-plan)
	for( Cx *c : cs ) if( c->prm.fp32Sweeps && !c->adapted() )
	{
		bool any = false;
		for( int d=0 ; d<=c->D ; d++ ) if( c->floatLevel( d ) && c->gsPlan( d , true ).kind==Cx::GS_PHASES && !c->levels[d].narrowed ) c->narrowOperators( d ) , any = true;
		if( any ) HIPCHECK( hipStreamSynchronize( c->stream ) );
	}
	static const bool useGraph = !( getenv( "EPR_BATCH_GRAPH" ) && atoi( getenv( "EPR_BATCH_GRAPH" ) )==0 );
	std::vector< long long > sig = BatchSignature< Dim >( cs );
	std::shared_ptr< BatchExec< Dim > > e = BatchExecCache< Dim >::Get().find( sig , useGraph );
	if( !e )
	{
		e = std::make_shared< BatchExec< Dim > >();
		e->setup( cs , std::move( sig ) , useGraph );
		BatchExecCache< Dim >::Get().add( e );
	}
	{
		std::lock_guard< std::mutex > lock( e->busy );
		e->run( cs , energies );
	}
	const double sec = std::chrono::duration< double >( std::chrono::steady_clock::now()-t0 ).count();
	for( Cx *c : cs ) c->solved = true , c->seconds[1] = sec;
}

// Tunes the batched solve's plans (the batched entries of the tuning table)
// on these contexts: for every level, each sweep kernel that can run it --
// the one-workgroup-per-context kernels on levels of at most TuneOneGroupMax
// nodes, colour grids of 64, 128 and 256 threads -- times the level's sweeps
// of every context with that level (their schedule's iterations, the depth
// groups one after another on one stream, x and y restored after every
// trial); the fastest is kept.  flags bit 0 also times the workgroup size of
// the FP32 colour grids (contexts with --fp32).  Returns one line per trial:
// "level res nodes contexts precision kind threads us", the kept ones marked '*'.
template< int Dim >
std::string AutotuneBatch( const std::vector< Ctx< Dim > * > &cs , int flags )
{
	using Cx = Ctx< Dim >;
	const size_t n = cs.size();
	if( !n ) return "";
	{
		std::set< const Cx * > seen;
		for( const Cx *c : cs ) if( !seen.insert( c ).second ) throw std::string( "a context appears twice in the batch" );
	}
	for( Cx *c : cs ) c->needBuilt();
	std::vector< std::vector< Cx * > > gc( MAXT );
	std::vector< std::vector< int > > gi( MAXT );
	int maxD = 0;
	for( size_t r=0 ; r<n ; r++ ) gc[ cs[r]->D ].push_back( cs[r] ) , gi[ cs[r]->D ].push_back( (int)r ) , maxD = std::max( maxD , cs[r]->D );
	hipStream_t st = nullptr;
	HIPCHECK( hipStreamCreate( &st ) );
	std::string rep;
	char line[200];
	DevBuf< char > table;
	DevBuf< double > save;
	try
	{
		for( int d=0 ; d<=maxD ; d++ )
		{
			// the contexts with this level, and room for their x, y
			std::vector< Cx * > with;
			for( Cx *c : cs ) if( c->D>=d && c->minD()<=d && c->maxD()>=d ) with.push_back( c );
			if( with.empty() ) continue;
			const long long N = with[0]->levels[d].N;
			const int res = with[0]->levels[d].res;
			save.ensure( (size_t)2*N*with.size() );
			for( size_t q=0 ; q<with.size() ; q++ )
			{
				HIPCHECK( hipMemcpyAsync( save.p + 2*q*N , with[q]->levels[d].x.p , N*sizeof(double) , hipMemcpyDeviceToDevice , st ) );
				HIPCHECK( hipMemcpyAsync( save.p + ( 2*q+1 )*N , with[q]->levels[d].y.p , N*sizeof(double) , hipMemcpyDeviceToDevice , st ) );
			}
			auto restore = [&]( void )
			{
				for( size_t q=0 ; q<with.size() ; q++ )
				{
					HIPCHECK( hipMemcpyAsync( with[q]->levels[d].x.p , save.p + 2*q*N , N*sizeof(double) , hipMemcpyDeviceToDevice , st ) );
					HIPCHECK( hipMemcpyAsync( with[q]->levels[d].y.p , save.p + ( 2*q+1 )*N , N*sizeof(double) , hipMemcpyDeviceToDevice , st ) );
				}
			};
			// the level's sweeps of every group under the current table
			auto trial = [&]( void )
			{
				for( Cx *c : with ) if( c->floatLevel( d ) && c->gsPlan( d , true ).kind==Cx::GS_PHASES && !c->levels[d].narrowed )
				{
					c->narrowOperators( d );
					HIPCHECK( hipStreamSynchronize( c->stream ) );
				}
				BatchHost H;
				std::vector< BatchGroup< Dim > > groups;
				for( int g=0 ; g<MAXT ; g++ ) if( !gc[g].empty() && g>=d ){ groups.emplace_back() ; groups.back().plan( gc[g] , gi[g] , H ); }
				table.ensure( std::max< size_t >( H.table.size() , 16 ) );
				HIPCHECK( hipMemcpyAsync( table.p , H.table.data() , H.table.size() , hipMemcpyHostToDevice , st ) );
				BatchDevPtrs P;
				P.table = table.p;
				auto sweeps = [&]{ for( const auto &g : groups ) g.enqueueSweeps( st , P , d ); };
				sweeps() , restore();   // warm-up
				const double us = 1000.*TimeOnStream( st , sweeps );
				restore();
				HIPCHECK( hipStreamSynchronize( st ) );   // the host table outlives no launch
				return us;
			};
			std::vector< std::pair< int , int > > cand;
			if( N<=TuneOneGroupMax ) for( int k : { (int)Cx::GS_WAVE , (int)Cx::GS_BLOCK , (int)Cx::GS_TINY , (int)Cx::GS_SMALL_LDS } ) if( with[0]->planKind( d , k ).kind>=0 ) cand.push_back( { k , 0 } );
			for( int th : TuneThreads ) cand.push_back( { (int)Cx::GS_PHASES , th } );
			double best = 1e300;
			std::pair< int , int > bc = cand.back();
			std::vector< std::string > lines;
			for( const auto &c : cand )
			{
				TuneTable::Get().set( Dim , res , true , false , TunePick{ c.first , c.second , 0 } );
				const double us = trial();
				snprintf( line , sizeof(line) , "%d %d %lld %zu fp64 %s %d %.3f" , d , res , N , with.size() , GsKindName( c.first ) , c.first==Cx::GS_PHASES ? c.second : with[0]->planKind( d , c.first ).threads , us );
				lines.push_back( line );
				if( us<best ) best = us , bc = c;
			}
			TuneTable::Get().set( Dim , res , true , false , TunePick{ bc.first , bc.second , best } );
			for( size_t k=0 ; k<cand.size() ; k++ ) rep += lines[k] + ( cand[k]==bc ? " *\n" : "\n" );
			bool anyF = false;
			for( Cx *c : with ) anyF |= c->floatLevel( d );
			if( ( flags & 1 ) && bc.first==Cx::GS_PHASES && anyF )
			{
				double bestF = 1e300;
				int bt = 256;
				std::vector< std::pair< std::string , int > > fl;
				for( int th : TuneThreads )
				{
					TuneTable::Get().set( Dim , res , true , true , TunePick{ -1 , th , 0 } );
					const double us = trial();
					snprintf( line , sizeof(line) , "%d %d %lld %zu fp32 phases %d %.3f" , d , res , N , with.size() , th , us );
					fl.push_back( { line , th } );
					if( us<bestF ) bestF = us , bt = th;
				}
				TuneTable::Get().set( Dim , res , true , true , TunePick{ -1 , bt , bestF } );
				for( const auto &f : fl ) rep += f.first + ( f.second==bt ? " *\n" : "\n" );
			}
		}
	}
	catch( ... ){ (void)hipStreamSynchronize( st ) ; (void)hipStreamDestroy( st ) ; throw; }
	(void)hipStreamDestroy( st );
	return rep;
}

// Measurement of the batched solve's colour grids: iters iterations of level
// d's colour launches (KGsPhaseBatch, the batched plan's workgroup size) over
// every dense context of the batch with that level, timed with HIP events on
// one stream.  fp32: on the float copies (x and y rounded first; the FP64
// iterate is left as it is); FP64: x and y are restored afterwards.  Returns
// ms for all launches, the launch count and the node updates per iteration
// (summed over the contexts).
template< int Dim >
void TimeBatchPhases( const std::vector< Ctx< Dim > * > &all , int d , bool fp32 , int iters , double *ms , int *launches , long long *nodes )
{
	using Cx = Ctx< Dim >;
	constexpr int C = DimC< Dim >::C , W = DimC< Dim >::W;
	if( iters<=0 ) throw std::string( "iters must be positive" );
	std::vector< Cx * > cs;
	for( Cx *c : all ){ c->needBuilt() ; if( c->D>=d && !c->adapted() && d>=0 ) cs.push_back( c ); }
	if( cs.empty() ) throw std::string( "no dense context of the batch has this level" );
	const long long N = cs[0]->levels[d].N;
	hipStream_t st = nullptr;
	HIPCHECK( hipStreamCreate( &st ) );
	DevBuf< char > table;
	DevBuf< double > save;
	try
	{
		std::vector< char > host;
		if( fp32 )
		{
			std::vector< BatchLevelT< float > > b;
			for( Cx *c : cs )
			{
				auto &L = c->levels[d];
				c->narrowOperators( d );
				KNarrow<<< Blocks( N , 256 ) , 256 , 0 , c->stream >>>( L.x.p , L.fx.p , N );
				KNarrow<<< Blocks( N , 256 ) , 256 , 0 , c->stream >>>( L.y.p , L.fy.p , N );
				HIPCHECK( hipStreamSynchronize( c->stream ) );
				b.push_back( { L.fx.p , L.fy.p , L.fR.p , L.fB.p , L.fK.p , nullptr , (float)c->prm.dWeight , c->prm.separate , iters } );
			}
			host.assign( (const char *)b.data() , (const char *)( b.data()+b.size() ) );
		}
		else
		{
			std::vector< BatchLevel > b;
			save.ensure( (size_t)2*N*cs.size() );
			for( size_t q=0 ; q<cs.size() ; q++ )
			{
				auto &L = cs[q]->levels[d];
				HIPCHECK( hipMemcpyAsync( save.p + 2*q*N , L.x.p , N*sizeof(double) , hipMemcpyDeviceToDevice , st ) );
				HIPCHECK( hipMemcpyAsync( save.p + ( 2*q+1 )*N , L.y.p , N*sizeof(double) , hipMemcpyDeviceToDevice , st ) );
				b.push_back( { L.x.p , L.y.p , L.R.p , L.B.p , L.K.p , nullptr , cs[q]->prm.dWeight , cs[q]->prm.separate , iters } );
			}
			host.assign( (const char *)b.data() , (const char *)( b.data()+b.size() ) );
		}
		table.alloc( host.size() );
		HIPCHECK( hipMemcpyAsync( table.p , host.data() , host.size() , hipMemcpyHostToDevice , st ) );
		HIPCHECK( hipStreamSynchronize( st ) );
		const LevelView< Dim > L0 = cs[0]->view( d );
		LevelView< Dim , float > F0;
		F0.res = L0.res , F0.N = L0.N , F0.x = F0.y = nullptr , F0.R = F0.B = F0.K = nullptr;
		F0.sW = (float)L0.sW , F0.dW = 0 , F0.lay = L0.lay , F0.fmask = nullptr , F0.separate = 0;
		for( int k=0 ; k<Dim ; k++ ) F0.stride[k] = L0.stride[k];
		const int th = fp32 ? cs[0]->floatThreads( d , true ) : cs[0]->gsPlan( d , true ).threads;
		const bool idx32 = N*W < ( 1LL<<31 ) , b32 = N*(long long)sizeof(double) < ( 1LL<<32 );
		const unsigned k = (unsigned)cs.size();
		auto run = [&]
		{
			for( int it=0 ; it<iters ; it++ ) for( int c=0 ; c<C ; c++ )
			{
				ColourSet< Dim > cset ; cset.set( 1<<d , c );
				const dim3 grid( Blocks( cset.total*C , th ) , k );
				if( fp32 )
				{
					const BatchLevelT< float > *J = (const BatchLevelT< float > *)table.p;
					if( b32 ){ ColourDiv< Dim > cdiv ; cdiv.set( cset ) ; KGsPhaseBatch32< Dim , float ><<< grid , th , 0 , st >>>( F0 , J , cdiv , L0.lay.cstart[c] ); }
						else if( idx32 ) KGsPhaseBatch< Dim , float , int ><<< grid , th , 0 , st >>>( F0 , J , cset , L0.lay.cstart[c] );
					else KGsPhaseBatch< Dim , float , long long ><<< grid , th , 0 , st >>>( F0 , J , cset , L0.lay.cstart[c] );
				}
				else
				{
					const BatchLevel *J = (const BatchLevel *)table.p;
					if( b32 ){ ColourDiv< Dim > cdiv ; cdiv.set( cset ) ; KGsPhaseBatch32< Dim , double ><<< grid , th , 0 , st >>>( L0 , J , cdiv , L0.lay.cstart[c] ); }
						else if( idx32 ) KGsPhaseBatch< Dim , double , int ><<< grid , th , 0 , st >>>( L0 , J , cset , L0.lay.cstart[c] );
					else KGsPhaseBatch< Dim , double , long long ><<< grid , th , 0 , st >>>( L0 , J , cset , L0.lay.cstart[c] );
				}
			}
		};
		const double t = TimeOnStream( st , run );
		HIPCHECK( hipGetLastError() );
		if( !fp32 ) for( size_t q=0 ; q<cs.size() ; q++ )
		{
			auto &L = cs[q]->levels[d];
			HIPCHECK( hipMemcpyAsync( L.x.p , save.p + 2*q*N , N*sizeof(double) , hipMemcpyDeviceToDevice , st ) );
			HIPCHECK( hipMemcpyAsync( L.y.p , save.p + ( 2*q+1 )*N , N*sizeof(double) , hipMemcpyDeviceToDevice , st ) );
		}
		HIPCHECK( hipStreamSynchronize( st ) );
		if( ms ) *ms = t;
		if( launches ) *launches = iters*C;
		if( nodes ) *nodes = N*(long long)cs.size();
	}
	catch( ... ){ (void)hipStreamSynchronize( st ) ; (void)hipStreamDestroy( st ) ; throw; }
	(void)hipStreamDestroy( st );
}
} // namespace epr

struct epr_ctx
{
	int dim;
	epr_params prm;
	std::unique_ptr< epr::CtxBase > impl;
	std::string err;
};

const char *EprMeshLastError( void );
epr::unet::Net *EprUnetNet( epr_unet *u );   // epr_unet.hip   // epr_levelset.hip: the calling thread's last context-free error

template< typename F >
static int Guard( epr_ctx *ctx , F f )
{
	if( !ctx ) return EPR_ERR_ARG;
	try { f() ; ctx->err.clear() ; return EPR_OK; }
	catch( const epr::HipError &e ){ ctx->err = e.msg ; return EPR_ERR_HIP; }
	catch( const std::string &s ){ ctx->err = s ; return EPR_ERR_STATE; }
	catch( const std::exception &e ){ ctx->err = e.what() ; return EPR_ERR_STATE; }
}

extern "C" {

void epr_default_params( epr_params *p )
{
	p->depth = 5 , p->seed = 0 , p->tRadius = 1 , p->dRadius = 1 , p->minSolveDepth = 0 , p->maxSolveDepth = -1 , p->gsIters = 10;
	p->singleLevel = 0 , p->verbose = 0 , p->separate = 0 , p->jitter = 0 , p->fp32Sweeps = 0;
	p->nSigma = 0.25 , p->scale = 1.1 , p->sWeight = 100. , p->dWeight = 0.00625 , p->iMult = 2.;
	p->aRadius = 3;
}

int epr_adapted_indexer_write( int dim , int depth , int aRadius , double scale , size_t n , const double *pos , const char *path )
{
	return epr::ContextFree( [&]()
	{
		if( ( dim!=3 && dim!=4 ) || !pos || !path || depth<0 || depth>12 ) throw std::string( "bad arguments" );
		auto run = [&]( auto tag )
		{
			constexpr int Dim = decltype( tag )::value;
			epr::SampleCells< Dim > sc;
			sc.bin( n , pos , scale , depth );
			epr::AdaptedHierarchy< Dim > h;
			h.build( sc.cellIdx.data() , (size_t)sc.ncell , depth , aRadius );
			FILE *fp = fopen( path , "wb" );
			if( !fp ) throw std::string( "cannot open " ) + path;
			try { h.write( fp ); } catch( ... ){ fclose( fp ) ; throw; }
			fclose( fp );
		};
		if( dim==3 ) run( std::integral_constant< int , 3 >() );
		else         run( std::integral_constant< int , 4 >() );
	} );
}

int epr_create( int dim , const epr_params *p , epr_ctx **out )
{
	if( !out || !p ) return EPR_ERR_ARG;
	*out = nullptr;
	if( dim!=3 && dim!=4 ) return EPR_ERR_ARG;
	epr_ctx *ctx = new epr_ctx;
	ctx->dim = dim , ctx->prm = *p;
	int rc = Guard( ctx , [&]()
	{
		if( dim==3 ) ctx->impl.reset( new epr::Ctx< 3 >( *p ) );
		else         ctx->impl.reset( new epr::Ctx< 4 >( *p ) );
	} );
	*out = ctx;
	return rc;
}

void epr_destroy( epr_ctx *ctx ){ delete ctx; }
const char *epr_last_error( const epr_ctx *ctx ){ return ctx ? ctx->err.c_str() : EprMeshLastError(); }

#define IMPL( body ) return Guard( ctx , [&](){ if( !ctx->impl ) throw std::string( "context not initialised" ) ; body; } )

int epr_set_samples( epr_ctx *ctx , size_t n , const double *pos , const double *skew )
{
	if( ctx && ( !pos || !skew ) ){ ctx->err = "null sample buffer" ; return EPR_ERR_ARG; }
	IMPL( ctx->impl->setSamples( n , pos , skew ) );
}
int epr_build( epr_ctx *ctx ){ IMPL( ctx->impl->build() ); }
int epr_init( epr_ctx *ctx , int level , const double *x0 , const double *y0 ){ IMPL( ctx->impl->init( level , x0 , y0 ) ); }
int epr_gs_sweeps( epr_ctx *ctx , int level , int iters ){ IMPL( ctx->impl->gs( level , iters ) ; HIPCHECK( hipDeviceSynchronize() ) ); }
int epr_gs_colour( epr_ctx *ctx , int level , int colour ){ IMPL( ctx->impl->gsColour( level , colour ) ); }
int epr_time_gs_sweeps_fp32( epr_ctx *ctx , int level , int iters , double *ms , int *launches ){ IMPL( ctx->impl->timeGsF32( level , iters , ms , launches ) ); }
int epr_time_gs_sweeps( epr_ctx *ctx , int level , int iters , double *ms , int *launches ){ IMPL( ctx->impl->timeGs( level , iters , ms , launches ) ); }
int epr_prolong_add( epr_ctx *ctx , int level ){ IMPL( ctx->impl->prolongAdd( level ) ); }
int epr_energies( epr_ctx *ctx , int level , double out[2] ){ IMPL( ctx->impl->energies( level , out ) ); }
int epr_energies_noalloc( epr_ctx *ctx , int level , double out[2] ){ if( !out ) return EPR_ERR_ARG ; IMPL( ctx->impl->energiesNoAlloc( level , out ) ); }
int epr_solve( epr_ctx *ctx , double out_E[2] ){ IMPL( ctx->impl->solve( out_E ) ); }
int epr_learned_features( epr_ctx *ctx , int level , float *out ){ if( !out ) return EPR_ERR_ARG ; IMPL( ctx->impl->learnedFeatures( level , out ) ); }
int epr_init_learned( epr_ctx *ctx , epr_unet *net , int level , double info[3] ){ if( !EprUnetNet( net ) ) return EPR_ERR_ARG ; IMPL( ctx->impl->learnedInit( level , EprUnetNet( net ) , info ) ); }
int epr_solve_schedule( epr_ctx *ctx , epr_unet *net , const int *iters , double out_E[2] )
{
	if( net && !EprUnetNet( net ) ) return EPR_ERR_ARG;
	IMPL( ctx->impl->solveSchedule( net ? EprUnetNet( net ) : nullptr , iters , out_E ) );
}
int epr_solve_learned( epr_ctx *ctx , epr_unet *net , double out_E[2] ){ if( !EprUnetNet( net ) ) return EPR_ERR_ARG ; IMPL( ctx->impl->solveLearned( EprUnetNet( net ) , out_E ) ); }
size_t epr_num_functions( const epr_ctx *ctx , int level ){ return ( ctx && ctx->impl && level>=0 && level<=ctx->prm.depth ) ? ctx->impl->numFunctions( level ) : 0; }
int epr_get_coeffs( epr_ctx *ctx , int level , double *x , double *y ){ IMPL( ctx->impl->getCoeffs( level , x , y ) ); }
int epr_set_coeffs( epr_ctx *ctx , int level , const double *x , const double *y ){ IMPL( ctx->impl->setCoeffs( level , x , y ) ); }
int epr_quadratic_fits( epr_ctx *ctx , int level , double *q ){ IMPL( ctx->impl->qfits( level , q ) ); }
int epr_apply( epr_ctx *ctx , int level , char which , const double *x , double *y ){ IMPL( ctx->impl->apply( level , which , x , y ) ); }
int epr_get_array( epr_ctx *ctx , const char *name , double *out , size_t cap , size_t *count )
{
	IMPL( size_t c = ctx->impl->getArray( name ? name : "" , out , cap ) ; if( count ) *count = c );
}
int epr_write_tree( epr_ctx *ctx , const char *path , unsigned long long hash ){ IMPL( ctx->impl->writeTree( path , hash ) ); }
int epr_write_grid( epr_ctx *ctx , const char *path ){ IMPL( ctx->impl->writeGrid( path ) ); }
int epr_write_density( epr_ctx *ctx , const char *path ){ IMPL( ctx->impl->writeDensity( path ) ); }
int epr_time_copy( size_t bytes , int reps , double *ms )
{
	if( !ms || reps<=0 || bytes<16 ) return EPR_ERR_ARG;
	const long long n = (long long)( bytes/16 );
	double2 *a = nullptr , *b = nullptr;
	hipEvent_t e0 = nullptr , e1 = nullptr;
	int rc = EPR_OK;
	auto ok = [&]( hipError_t e ){ if( e!=hipSuccess && rc==EPR_OK ) rc = EPR_ERR_HIP ; return rc==EPR_OK; };
	if( ok( hipMalloc( &a , n*16 ) ) && ok( hipMalloc( &b , n*16 ) ) && ok( hipMemset( a , 0 , n*16 ) ) && ok( hipEventCreate( &e0 ) ) && ok( hipEventCreate( &e1 ) ) )
	{
		const unsigned int blocks = (unsigned int)std::min< long long >( ( n + 255 )/256 , 256LL*64 );
		epr::KCopy<<< blocks , 256 >>>( a , b , n );
		ok( hipEventRecord( e0 , nullptr ) );
		for( int r=0 ; r<reps ; r++ ) epr::KCopy<<< blocks , 256 >>>( a , b , n );
		ok( hipEventRecord( e1 , nullptr ) );
		float t = 0;
		if( ok( hipEventSynchronize( e1 ) ) && ok( hipEventElapsedTime( &t , e0 , e1 ) ) ) *ms = t / reps;
	}
	if( e0 ) (void)hipEventDestroy( e0 );
	if( e1 ) (void)hipEventDestroy( e1 );
	if( a ) (void)hipFree( a );
	if( b ) (void)hipFree( b );
	return rc;
}
static std::vector< epr::CtxBase * > BatchImpls( size_t n , epr_ctx *const *ctxs , bool oneDim , int &dim )
{
	if( n && !ctxs ) throw std::string( "null context array" );
	std::vector< epr::CtxBase * > v;
	std::set< const epr_ctx * > seen;
	dim = 0;
	for( size_t i=0 ; i<n ; i++ )
	{
		if( !ctxs[i] || !ctxs[i]->impl ) throw std::string( "context " + std::to_string( i ) + " is null or not initialised" );
		if( !seen.insert( ctxs[i] ).second ) throw std::string( "a context appears twice in the batch" );
		if( oneDim && dim && ctxs[i]->dim!=dim ) throw std::string( "contexts of different dimensions in one batch" );
		dim = ctxs[i]->dim;
		v.push_back( ctxs[i]->impl.get() );
	}
	return v;
}
int epr_build_batch( size_t n , epr_ctx *const *ctxs , int threads )
{
	return epr::ContextFree( [&](){ int dim ; epr::BuildBatch( BatchImpls( n , ctxs , false , dim ) , threads ); } );
}
int epr_solve_batch( size_t n , epr_ctx *const *ctxs , double *energies )
{
	return epr::ContextFree( [&]()
	{
		int dim;
		std::vector< epr::CtxBase * > v = BatchImpls( n , ctxs , true , dim );
		auto run = [&]( auto tag )
		{
			constexpr int Dim = decltype( tag )::value;
			std::vector< epr::Ctx< Dim > * > c;
			for( epr::CtxBase *b : v ) c.push_back( static_cast< epr::Ctx< Dim > * >( b ) );
			epr::SolveBatch< Dim >( c , energies );
		};
		if( dim==3 ) run( std::integral_constant< int , 3 >() );
		else if( dim==4 ) run( std::integral_constant< int , 4 >() );
	} );
}
static void CopyReport( const std::string &r , char *out , size_t cap )
{
	if( !out || !cap ) return;
	const size_t k = std::min( r.size() , cap-1 );
	memcpy( out , r.data() , k ) , out[k] = 0;
}
int epr_autotune( epr_ctx *ctx , int flags , char *report , size_t cap ){ IMPL( CopyReport( ctx->impl->autotune( flags ) , report , cap ) ); }
int epr_autotune_batch( size_t n , epr_ctx *const *ctxs , int flags , char *report , size_t cap )
{
	return epr::ContextFree( [&]()
	{
		int dim;
		std::vector< epr::CtxBase * > v = BatchImpls( n , ctxs , true , dim );
		auto run = [&]( auto tag )
		{
			constexpr int Dim = decltype( tag )::value;
			std::vector< epr::Ctx< Dim > * > c;
			for( epr::CtxBase *b : v ) c.push_back( static_cast< epr::Ctx< Dim > * >( b ) );
			CopyReport( epr::AutotuneBatch< Dim >( c , flags ) , report , cap );
		};
		if( dim==3 ) run( std::integral_constant< int , 3 >() );
		else if( dim==4 ) run( std::integral_constant< int , 4 >() );
	} );
}
int epr_tune_get( char *buf , size_t cap , size_t *len )
{
	return epr::ContextFree( [&]()
	{
		const std::string t = epr::TuneTable::Get().text();
		if( len ) *len = t.size();
		CopyReport( t , buf , cap );
	} );
}
int epr_time_op( epr_ctx *ctx , const char *op , int level , int reps , double *ms , double *bytes ){ IMPL( ctx->impl->timeOp( op ? op : "" , level , reps , ms , bytes ) ); }
int epr_time_batch_sweeps( size_t n , epr_ctx *const *ctxs , int level , int fp32 , int iters , double *ms , int *launches , long long *nodes )
{
	return epr::ContextFree( [&]()
	{
		int dim;
		std::vector< epr::CtxBase * > v = BatchImpls( n , ctxs , true , dim );
		auto run = [&]( auto tag )
		{
			constexpr int Dim = decltype( tag )::value;
			std::vector< epr::Ctx< Dim > * > c;
			for( epr::CtxBase *b : v ) c.push_back( static_cast< epr::Ctx< Dim > * >( b ) );
			epr::TimeBatchPhases< Dim >( c , level , fp32!=0 , iters , ms , launches , nodes );
		};
		if( dim==3 ) run( std::integral_constant< int , 3 >() );
		else if( dim==4 ) run( std::integral_constant< int , 4 >() );
		else throw std::string( "empty batch" );
	} );
}
int epr_tune_set( const char *text ){ return epr::ContextFree( [&](){ epr::TuneTable::Get().parse( text ); } ); }
double epr_last_seconds( const epr_ctx *ctx , int which ){ return ( ctx && ctx->impl && ( which==0 || which==1 ) ) ? ctx->impl->seconds[which] : -1.; }

}
