// Kernels of the batched solve: the same level of many independent
// reconstructions in one launch.  All reconstructions of a launch share the
// level's resolution (hence its LevelView constants); what differs per
// reconstruction -- its buffers, Dirichlet weight, step kind and iteration
// count -- comes from a BatchLevel table in device memory.
#pragma once
#include "epr_kernels.h"

namespace epr
{
template< typename Real = double >
struct BatchLevelT
{
	Real *x , *y;
	const Real *R , *B , *K;
	const unsigned char *fmask;
	Real dW;
	int separate , iters;
};
using BatchLevel = BatchLevelT< double >;
struct BatchList { const unsigned int *list ; long long count; };   // the active nodes of one colour (adapted levels)
struct BatchProlong { const double *xc ; double *xf ; const unsigned char *fmaskF ; int mode; };
struct BatchZero { double *p ; long long n; };

template< int Dim , typename Real = double >
__device__ __forceinline__ LevelView< Dim , Real > BatchView( const LevelView< Dim , Real > &L0 , const BatchLevelT< Real > &b )
{
	LevelView< Dim , Real > L = L0;
	L.x = b.x , L.y = b.y , L.R = b.R , L.B = b.B , L.K = b.K , L.fmask = b.fmask , L.dW = b.dW , L.separate = b.separate;
	return L;
}

// one workgroup per reconstruction, all of its iterations (KGsTiny)
template< int Dim , bool RB >
__global__ void __launch_bounds__( TinyMaxThreads< Dim >() ) KGsTinyBatch( LevelView< Dim > L0 , const BatchLevel *P )
{
	const BatchLevel &b = P[ blockIdx.x ];
	GsTiny< Dim , RB >( BatchView( L0 , b ) , b.iters );
}
// one wavefront per reconstruction, all of its iterations (KGsWave: 3D levels of resolution 1 or 2)
template< int RES >
__global__ void __launch_bounds__( 64 ) KGsWaveBatch( LevelView< 3 > L0 , const BatchLevel *P )
{
	const BatchLevel &b = P[ blockIdx.x ];
	GsWave< RES >( BatchView( L0 , b ) , b.iters );
}
// one workgroup of 512 threads per reconstruction (KGsBlock: the 3D resolution-4 level)
__global__ void __launch_bounds__( 512 ) KGsBlockBatch( LevelView< 3 > L0 , const BatchLevel *P )
{
	const BatchLevel &b = P[ blockIdx.x ];
	GsBlock( BatchView( L0 , b ) , b.iters );
}
// one workgroup per reconstruction, all of its iterations (KGsSmallLds)
template< int Dim >
__global__ void __launch_bounds__( Dim==3 ? 1024 : 512 ) KGsSmallLdsBatch( LevelView< Dim > L0 , const BatchLevel *P , int rbInLds )
{
	const BatchLevel &b = P[ blockIdx.x ];
	GsSmallLds< Dim >( BatchView( L0 , b ) , b.iters , rbInLds );
}
// one colour phase of the first gridDim.y reconstructions; the workgroups of
// one XCD take a contiguous range of (reconstruction, block) pairs
// (the FP32 form: Real = float on the float copies of x, y, R, B, K)
template< int Dim , typename Real = double , typename Idx = long long >
__global__ void __launch_bounds__( 256 ) KGsPhaseBatch( LevelView< Dim , Real > L0 , const BatchLevelT< Real > *P , ColourSet< Dim > cs , long long colBase )
{
	const unsigned lin = XcdBlock( blockIdx.y*gridDim.x + blockIdx.x , gridDim.x*gridDim.y );
	const unsigned r = lin / gridDim.x , b = lin - r*gridDim.x;
	GsPhase< Dim , Real , Idx >( BatchView( L0 , P[r] ) , cs , colBase , b );
}
// the same with GsPhase32 (dense levels whose coefficients sit below 4 GiB)
template< int Dim , typename Real >
__global__ void __launch_bounds__( 256 ) KGsPhaseBatch32( LevelView< Dim , Real > L0 , const BatchLevelT< Real > *P , ColourDiv< Dim > cs , long long colBase )
{
	const unsigned lin = XcdBlock( blockIdx.y*gridDim.x + blockIdx.x , gridDim.x*gridDim.y );
	const unsigned r = lin / gridDim.x , b = lin - r*gridDim.x;
	GsPhase32< Dim , Real >( BatchView( L0 , P[r] ) , cs , colBase , b );
}
// the adapted counterpart: colour c's active-node list of reconstruction r at lists[ r*2^Dim + c ]
template< int Dim >
__global__ void __launch_bounds__( 256 ) KGsPhaseListBatch( LevelView< Dim > L0 , const BatchLevel *P , const BatchList *lists , int c )
{
	constexpr int C = DimC< Dim >::C;
	const unsigned lin = XcdBlock( blockIdx.y*gridDim.x + blockIdx.x , gridDim.x*gridDim.y );
	const unsigned r = lin / gridDim.x , b = lin - r*gridDim.x;
	const BatchList &l = lists[ (size_t)r*C + c ];
	GsPhaseList< Dim >( BatchView( L0 , P[r] ) , l.list , l.count , b );
}
// prolongation job blockIdx.y (KProlong)
template< int Dim >
__global__ void __launch_bounds__( 256 ) KProlongBatch( const BatchProlong *J , int cres )
{
	const BatchProlong &j = J[ blockIdx.y ];
	Prolong< Dim >( j.xc , j.xf , cres , j.mode , j.fmaskF , (long long)blockIdx.x*blockDim.x + threadIdx.x );
}
// copy of job blockIdx.y
struct BatchCopy { const double *src ; double *dst ; long long n; };
__global__ void __launch_bounds__( 256 ) KCopyBatch( const BatchCopy *J )
{
	const BatchCopy &j = J[ blockIdx.y ];
	for( long long i=(long long)blockIdx.x*blockDim.x + threadIdx.x ; i<j.n ; i+=(long long)gridDim.x*blockDim.x ) j.dst[i] = j.src[i];
}
// rounding (KNarrow) or widening (KWiden) copy of job blockIdx.y
struct BatchCast { double *d ; float *f ; long long n; };
__global__ void __launch_bounds__( 256 ) KNarrowBatch( const BatchCast *J )
{
	const BatchCast &j = J[ blockIdx.y ];
	for( long long i=(long long)blockIdx.x*blockDim.x + threadIdx.x ; i<j.n ; i+=(long long)gridDim.x*blockDim.x ) j.f[i] = (float)j.d[i];
}
__global__ void __launch_bounds__( 256 ) KWidenBatch( const BatchCast *J )
{
	const BatchCast &j = J[ blockIdx.y ];
	for( long long i=(long long)blockIdx.x*blockDim.x + threadIdx.x ; i<j.n ; i+=(long long)gridDim.x*blockDim.x ) j.d[i] = (double)j.f[i];
}
// The random start of job blockIdx.y (ExteriorPoisson.h:333-341): the k-th
// pair of the context's draws goes to function order[k] (the adapted
// hierarchy's numbering) or k (dense); the other functions stay zero.
struct BatchStart { long long drawOff ; double *x , *y ; const unsigned int *order ; long long count; };
__global__ void __launch_bounds__( 256 ) KStartBatch( const BatchStart *J , const double *draws )
{
	const BatchStart &j = J[ blockIdx.y ];
	const double *dr = draws + j.drawOff;
	for( long long k=(long long)blockIdx.x*blockDim.x + threadIdx.x ; k<j.count ; k+=(long long)gridDim.x*blockDim.x )
	{
		const long long f = j.order ? (long long)j.order[k] : k;
		j.x[f] = dr[2*k] , j.y[f] = dr[2*k+1];
	}
}
// The start's rescale s = ( c / W(x,y) )^(1/4) of job blockIdx.y
// (ExteriorPoisson.h:338-341): c of its context at c[ctx], W at terms[4 ctx]
// (written by KEnergy + KSumRows earlier on the stream).  sqrt and the
// division are correctly rounded, as the host's are.
struct BatchScale { double *x , *y ; long long n ; int ctx; };
__global__ void __launch_bounds__( 256 ) KRescaleBatch( const BatchScale *J , const double *c , const double *terms )
{
	const BatchScale &j = J[ blockIdx.y ];
	const double s = sqrt( sqrt( c[ j.ctx ] / terms[ 4*j.ctx ] ) );
	for( long long i=(long long)blockIdx.x*blockDim.x + threadIdx.x ; i<j.n ; i+=(long long)gridDim.x*blockDim.x ) j.x[i] *= s , j.y[i] *= s;
}
// zero-fill of job blockIdx.y
__global__ void __launch_bounds__( 256 ) KZeroBatch( const BatchZero *J )
{
	const BatchZero &j = J[ blockIdx.y ];
	for( long long i=(long long)blockIdx.x*blockDim.x + threadIdx.x ; i<j.n ; i+=(long long)gridDim.x*blockDim.x ) j.p[i] = 0;
}
} // namespace epr
