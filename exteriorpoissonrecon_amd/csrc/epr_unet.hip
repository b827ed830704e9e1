// The learned initializer's network: a 4-level convolutional U-Net on the
// hat-function node grids (2^l+1)^Dim, BF16 storage, FP32 accumulation, with
// hand-written gfx950 MFMA kernels for every convolution of the forward and
// backward passes and an Adam optimizer on the device.
//
// No reference counterpart: it predicts the coefficients (x, y) of a coarse
// level, which replace the random start of ExteriorPoisson.h:333-341 and are
// then refined by the reference's own cascadic solve.
//
// Architecture (channels c_l = 16 * 2^l on level l = 0..3, level 0 the finest
// grid, side s_l = 2^(L-l) + 1 for the network level L):
//   encoder  h_l = relu(conv(in_l)),  s_l = relu(conv(h_l)),  in_0 = the
//            input field, in_l = down(s_{l-1})
//   decoder  g_l = relu(conv(up(d_{l+1}))), a_l = g_l + s_l, d_l = relu(conv(a_l)),
//            d_3 = s_3
//   output   y = conv(d_0) (2 channels, no activation)
// conv: 3^Dim taps, zero outside the grid, bias; tap t has the offsets
// (digit_d(t) - 1), base-3 digits with the first coordinate most significant.
// down: hat-function restriction P^T normalised by its weight sum (an average
// over the fine nodes 2c + {-1,0,1}^Dim with weights prod 1 or 1/2); up: the
// hat-function prolongation P (multilinear interpolation).
//
// Loss of a sample with prediction p and target t over N nodes (the solution
// is defined up to a rotation of the (x, y) plane, an exact symmetry of the
// energy, so the loss is the distance after the best rotation):
//   L = ( |p|^2 + |t|^2 - 2 sqrt( A^2 + B^2 ) ) / N,
//   A = sum px tx + py ty,  B = sum px ty - py tx,
// averaged over the samples of a batch.
//
// Storage: activations and activation gradients bf16, node-major with the
// channels of a node contiguous and padded to 16 (zero); weights fp32 master
// (Adam) with a bf16 copy [tap][cout][cin] (padded) and a bf16 transposed,
// tap-mirrored copy [tap][cin][cout] for the input gradients; biases fp32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/epr.h"
#include "epr_dev.h"
#include "epr_unet.h"

namespace epr
{
namespace unet
{
typedef short v4s __attribute__( ( ext_vector_type( 4 ) ) );
typedef float v4f __attribute__( ( ext_vector_type( 4 ) ) );

__host__ __device__ constexpr int Pow3( int d ){ return d==0 ? 1 : 3*Pow3( d-1 ); }

template< int Dim >
__device__ __forceinline__ void Coords( long long n , int side , int *F )
{
#pragma unroll
	for( int d=Dim-1 ; d>=0 ; d-- ){ F[d] = (int)( n % side ) ; n /= side; }
}

// ---------------------------------------------------------------------------
// Convolution, forward (and input gradient with the transposed weights):
//   out[b][n][co] = act( bias[co] + sum_{tap,ci} w[tap][co][ci] in[b][n+off(tap)][ci] )
// act = relu or identity; with skip, out2[b][n][co] = out + skip[b][n][co].
// One wavefront computes 16 nodes x 16*NT output channels with
// v_mfma_f32_16x16x16_bf16: A = 16 nodes x 16 input channels (lane l: node
// l%16, channels 4(l/16)..+3), B = 16 input x 16 output channels (lane l:
// output channel l%16, input channels 4(l/16)..+3), D = 16 nodes x 16 output
// channels (lane l: output channel l%16, nodes 4(l/16)+j, j = 0..3).
// grid: x = ceil(N/64), y = batch, z = coutP/(16 NT); 256 threads.
template< int Dim , int NT >
__global__ void __launch_bounds__( 256 ) KConv( const bf16 *__restrict__ in , int cinP , const bf16 *__restrict__ w , const float *__restrict__ bias , int coutP ,
                                                int side , long long N , int relu , const bf16 *__restrict__ skip , bf16 *__restrict__ out , bf16 *__restrict__ out2 , float *__restrict__ outF )
{
	constexpr int W = Pow3( Dim );
	const int lane = threadIdx.x & 63 , wave = threadIdx.x >> 6;
	const int row = lane & 15 , kg = lane >> 4;
	const long long b = blockIdx.y;
	const long long n0 = (long long)blockIdx.x*64 + wave*16;
	if( n0>=N ) return;
	const long long n = n0 + row;
	int F[Dim];
	Coords< Dim >( n<N ? n : 0 , side , F );
	long long stride[Dim];
	{
		long long s = 1;
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ stride[d] = s ; s *= side; }
	}
	const int coBase = blockIdx.z*16*NT;
	v4f acc[NT];
#pragma unroll
	for( int t=0 ; t<NT ; t++ ) acc[t] = v4f{ 0.f , 0.f , 0.f , 0.f };
	const bf16 *inb = in + (size_t)b*N*cinP + 4*kg;
	const bf16 *wl = w + (size_t)( coBase + row )*cinP + 4*kg;
	for( int tap=0 ; tap<W ; tap++ )
	{
		int t = tap;
		bool valid = n<N;
		long long nb = n;
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- )
		{
			const int o = t%3 - 1 ; t /= 3;
			const int g = F[d] + o;
			valid = valid && g>=0 && g<side;
			nb += o*stride[d];
		}
		const bf16 *ap = inb + (size_t)( valid ? nb : 0 )*cinP;
		const bf16 *wp = wl + (size_t)tap*coutP*cinP;
		for( int ci0=0 ; ci0<cinP ; ci0+=16 )
		{
			v4s a = *(const v4s *)( ap + ci0 );
			if( !valid ) a = v4s{ 0 , 0 , 0 , 0 };
#pragma unroll
			for( int tt=0 ; tt<NT ; tt++ )
			{
				const v4s bb = *(const v4s *)( wp + (size_t)tt*16*cinP + ci0 );
				acc[tt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k( a , bb , acc[tt] , 0 , 0 , 0 );
			}
		}
	}
#pragma unroll
	for( int tt=0 ; tt<NT ; tt++ )
	{
		const int co = coBase + tt*16 + row;
		const float bv = bias ? bias[co] : 0.f;
#pragma unroll
		for( int j=0 ; j<4 ; j++ )
		{
			const long long node = n0 + 4*kg + j;
			if( node>=N ) continue;
			float v = acc[tt][j] + bv;
			if( relu ) v = fmaxf( v , 0.f );
			const size_t o = ( (size_t)b*N + node )*coutP + co;
			if( outF ) outF[o] = v;
			else out[o] = ToBf16( v );
			if( skip ) out2[o] = ToBf16( FromBf16( ToBf16( v ) ) + FromBf16( skip[o] ) );
		}
	}
}

// Weight gradient: partial[chunk][tap][co][ci] = sum over the chunk's nodes g
// (all samples, g = b N + n) of dout[g][co] in[g+off(tap)][ci].  MFMA with M =
// output channels, K = nodes, N = input channels (D: lane l holds co = co0 +
// 4(l/16)+j, ci = ci0 + l%16).  grid: x = tap, y = chunk, z = co tile x ci
// tile; 256 threads, the four waves take interleaved 16-node steps.
template< int Dim >
__global__ void __launch_bounds__( 256 ) KConvWGrad( const bf16 *__restrict__ in , int cinP , const bf16 *__restrict__ dout , int coutP ,
                                                    int side , long long N , long long total , long long chunk , float *__restrict__ partial )
{
	constexpr int W = Pow3( Dim );
	const int lane = threadIdx.x & 63 , wave = threadIdx.x >> 6;
	const int row = lane & 15 , kg = lane >> 4;
	const int tap = blockIdx.x;
	const int nCi = cinP/16;
	const int co0 = ( blockIdx.z / nCi )*16 , ci0 = ( blockIdx.z % nCi )*16;
	int o[Dim];
	long long off = 0;
	{
		int t = tap;
		long long s = 1;
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ o[d] = t%3 - 1 ; t /= 3 ; off += o[d]*s ; s *= side; }
	}
	const long long g0 = (long long)blockIdx.y*chunk , g1 = g0+chunk<total ? g0+chunk : total;
	v4f acc = v4f{ 0.f , 0.f , 0.f , 0.f };
	for( long long g=g0+wave*16 ; g<g1 ; g+=64 )
	{
		v4s a , bv;
#pragma unroll
		for( int j=0 ; j<4 ; j++ )
		{
			const long long gk = g + 4*kg + j;
			const bool ok = gk<g1;
			const long long gs = ok ? gk : g0;
			const long long nn = gs % N;
			int F[Dim];
			Coords< Dim >( nn , side , F );
			bool v = ok;
#pragma unroll
			for( int d=0 ; d<Dim ; d++ ){ const int q = F[d] + o[d] ; v = v && q>=0 && q<side; }
			const bf16 av = dout[ (size_t)gs*coutP + co0 + row ];
			const bf16 iv = in[ (size_t)( v ? gs+off : gs )*cinP + ci0 + row ];
			a[j] = ok ? (short)av : (short)0;
			bv[j] = v ? (short)iv : (short)0;
		}
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k( a , bv , acc , 0 , 0 , 0 );
	}
	__shared__ float red[4][64*4];
#pragma unroll
	for( int j=0 ; j<4 ; j++ ) red[wave][ lane*4+j ] = acc[j];
	__syncthreads();
	if( wave==0 )
	{
#pragma unroll
		for( int j=0 ; j<4 ; j++ )
		{
			const float s = ( ( red[0][lane*4+j] + red[1][lane*4+j] ) + red[2][lane*4+j] ) + red[3][lane*4+j];
			const int co = co0 + 4*kg + j , ci = ci0 + row;
			partial[ ( (size_t)blockIdx.y*W + tap )*coutP*cinP + (size_t)co*cinP + ci ] = s;
		}
	}
}

// out[k] = sum_c in[c*count + k], c in increasing order
__global__ void __launch_bounds__( 256 ) KSumChunks( const float *in , long long nChunks , long long count , float *out )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( k>=count ) return;
	float s = 0.f;
	for( long long c=0 ; c<nChunks ; c++ ) s += in[ c*count + k ];
	out[k] = s;
}

// out[c] = sum_g d[g][c] (fixed order: strided per thread, then a tree); one workgroup per channel
__global__ void __launch_bounds__( 256 ) KBiasGrad( const bf16 *d , int cP , long long total , float *out )
{
	__shared__ float sh[256];
	const int c = blockIdx.x;
	float s = 0.f;
	for( long long g=threadIdx.x ; g<total ; g+=256 ) s += FromBf16( d[ (size_t)g*cP + c ] );
	sh[threadIdx.x] = s;
	__syncthreads();
	for( int h=128 ; h>0 ; h>>=1 ){ if( (int)threadIdx.x<h ) sh[threadIdx.x] += sh[threadIdx.x+h] ; __syncthreads(); }
	if( threadIdx.x==0 ) out[c] = sh[0];
}

// dAcc = h>0 ? dOut : 0
__global__ void __launch_bounds__( 256 ) KReluMask( const bf16 *dOut , const bf16 *h , long long count , bf16 *dAcc )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( k<count ) dAcc[k] = FromBf16( h[k] )>0.f ? dOut[k] : (bf16)0;
}

// wt[tap'][ci][co] = w[W-1-tap'][co][ci] (the input gradient's weights: transposed, taps mirrored)
__global__ void __launch_bounds__( 256 ) KTransposeMirror( const bf16 *w , int W , int coutP , int cinP , bf16 *wt )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	const long long total = (long long)W*coutP*cinP;
	if( k>=total ) return;
	const int co = (int)( k % coutP ) , ci = (int)( ( k/coutP ) % cinP ) , t = (int)( k/( (long long)coutP*cinP ) );
	wt[k] = w[ ( (size_t)( W-1-t )*coutP + co )*cinP + ci ];
}

// Hat prolongation P on the node grids, coarse side sc -> fine side 2 sc - 1:
//   fine[f][c] = sum over parents p (f_d even: p_d = f_d/2, weight 1; odd:
//   (f_d -+ 1)/2, weight 1/2 each) of w coarse[p][c] (/ wsum(p) when norm).
// acc: fine += (in fp32, stored bf16).  One thread per (fine node, channel).
template< int Dim >
__device__ __forceinline__ float DownWeightSum( const int *C , int sf )
{
	float s = 1.f;
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) s *= 1.f + ( 2*C[d]-1>=0 ? 0.5f : 0.f ) + ( 2*C[d]+1<=sf-1 ? 0.5f : 0.f );
	return s;
}
template< int Dim >
__global__ void __launch_bounds__( 256 ) KUp( const bf16 *coarse , int sc , long long Nc , bf16 *fine , long long Nf , int cP , long long count , int norm , int acc )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( k>=count ) return;
	const int c = (int)( k % cP );
	const long long g = k / cP , b = g / Nf , f = g % Nf;
	const int sf = 2*sc - 1;
	int F[Dim];
	Coords< Dim >( f , sf , F );
	float s = 0.f;
	for( int m=0 ; m<(1<<Dim) ; m++ )
	{
		int P[Dim];
		float wgt = 1.f;
		bool ok = true;
#pragma unroll
		for( int d=0 ; d<Dim ; d++ )
		{
			const int bit = ( m>>d ) & 1;
			if( F[d] & 1 ){ P[d] = ( F[d]-1 )/2 + bit ; wgt *= 0.5f; }
			else { P[d] = F[d]/2 ; if( bit ) ok = false; }
		}
		if( !ok ) continue;
		long long p = 0;
#pragma unroll
		for( int d=0 ; d<Dim ; d++ ) p = p*sc + P[d];
		float v = FromBf16( coarse[ ( (size_t)b*Nc + p )*cP + c ] );
		if( norm ) v /= DownWeightSum< Dim >( P , sf );
		s += wgt*v;
	}
	if( acc ) s += FromBf16( fine[k] );
	fine[k] = ToBf16( s );
}
// P^T: coarse[p][c] = sum over fine nodes f in 2p + {-1,0,1}^Dim of w fine[f][c]
// (/ wsum(p) when norm: the down-sampling average).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KDown( const bf16 *fine , int sc , long long Nc , bf16 *coarse , long long Nf , int cP , long long count , int norm )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( k>=count ) return;
	const int c = (int)( k % cP );
	const long long g = k / cP , b = g / Nc , p = g % Nc;
	const int sf = 2*sc - 1;
	int P[Dim];
	Coords< Dim >( p , sc , P );
	float s = 0.f;
	for( int m=0 ; m<Pow3( Dim ) ; m++ )
	{
		int t = m;
		long long f = 0;
		float wgt = 1.f;
		bool ok = true;
		int Fd[Dim];
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ const int o = t%3 - 1 ; t /= 3 ; Fd[d] = 2*P[d] + o ; ok = ok && Fd[d]>=0 && Fd[d]<sf ; if( o ) wgt *= 0.5f; }
		if( !ok ) continue;
#pragma unroll
		for( int d=0 ; d<Dim ; d++ ) f = f*sf + Fd[d];
		s += wgt*FromBf16( fine[ ( (size_t)b*Nf + f )*cP + c ] );
	}
	if( norm ) s /= DownWeightSum< Dim >( P , sf );
	coarse[k] = ToBf16( s );
}

// out = bf16( a + b )
__global__ void __launch_bounds__( 256 ) KAddBf16( const bf16 *a , const bf16 *b , long long count , bf16 *out )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( k<count ) out[k] = ToBf16( FromBf16( a[k] ) + FromBf16( b[k] ) );
}

// Loss statistics per sample: part[(b*nb + block)*4 + {pp, tt, A, B}]
__global__ void __launch_bounds__( 256 ) KLossPartials( const float *y , const float *t , long long N , float *part )
{
	__shared__ float sh[4][256];
	const long long b = blockIdx.y;
	float s[4] = { 0.f , 0.f , 0.f , 0.f };
	for( long long n=(long long)blockIdx.x*256+threadIdx.x ; n<N ; n+=(long long)gridDim.x*256 )
	{
		const float px = y[ ( b*N + n )*CT ] , py = y[ ( b*N + n )*CT + 1 ] , tx = t[ ( b*N + n )*2 ] , ty = t[ ( b*N + n )*2 + 1 ];
		s[0] += px*px + py*py , s[1] += tx*tx + ty*ty , s[2] += px*tx + py*ty , s[3] += px*ty - py*tx;
	}
	for( int k=0 ; k<4 ; k++ ) sh[k][threadIdx.x] = s[k];
	__syncthreads();
	for( int h=128 ; h>0 ; h>>=1 ){ if( (int)threadIdx.x<h ) for( int k=0 ; k<4 ; k++ ) sh[k][threadIdx.x] += sh[k][threadIdx.x+h] ; __syncthreads(); }
	if( threadIdx.x==0 ) for( int k=0 ; k<4 ; k++ ) part[ ( b*gridDim.x + blockIdx.x )*4 + k ] = sh[k][0];
}
// stats[b] = sum of the sample's partials (in block order)
__global__ void KLossStats( const float *part , int nb , int batch , double *stats )
{
	const int b = threadIdx.x;
	if( b>=batch ) return;
	double s[4] = { 0 , 0 , 0 , 0 };
	for( int i=0 ; i<nb ; i++ ) for( int k=0 ; k<4 ; k++ ) s[k] += part[ ( (size_t)b*nb + i )*4 + k ];
	for( int k=0 ; k<4 ; k++ ) stats[ b*4+k ] = s[k];
}
// dy = dL/dy of the batch-mean loss (channels 0, 1; the others zero)
__global__ void __launch_bounds__( 256 ) KLossGrad( const float *y , const float *t , const double *stats , long long N , int batch , bf16 *dy )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( k>=(long long)batch*N*CT ) return;
	const int c = (int)( k % CT );
	const long long g = k / CT , b = g / N;
	if( c>=2 ){ dy[k] = 0 ; return; }
	const double A = stats[ b*4+2 ] , B = stats[ b*4+3 ] , rho = sqrt( A*A + B*B );
	const double cs = rho>0 ? A/rho : 1. , sn = rho>0 ? B/rho : 0.;
	const double px = y[ g*CT ] , py = y[ g*CT+1 ] , tx = t[ g*2 ] , ty = t[ g*2+1 ];
	const double f = 2. / ( (double)N*batch );
	const double v = c==0 ? f*( px - cs*tx - sn*ty ) : f*( py - cs*ty + sn*tx );
	dy[k] = ToBf16( (float)v );
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the padded fp32 parameters; the
// bf16 copy follows.  Padding entries have zero gradients and stay zero.
__global__ void __launch_bounds__( 256 ) KAdam( float *p , const float *g , float *m , float *v , long long count , float lr , float c1 , float c2 , bf16 *pb )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( k>=count ) return;
	const float gk = g[k];
	const float mk = 0.9f*m[k] + 0.1f*gk , vk = 0.999f*v[k] + 0.001f*gk*gk;
	m[k] = mk , v[k] = vk;
	const float pk = p[k] - lr*( mk/c1 )/( sqrtf( vk/c2 ) + 1e-8f );
	p[k] = pk;
	pb[k] = ToBf16( pk );
}
__global__ void __launch_bounds__( 256 ) KToBf16( const float *p , long long count , bf16 *pb )
{
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( k<count ) pb[k] = ToBf16( p[k] );
}
} // namespace unet
} // namespace epr

// ---------------------------------------------------------------------------
// Host side
namespace epr
{
namespace unet
{
struct Layer { int level , cin , cout , relu ; size_t offW , offB ; int cinP , coutP; };

struct Net
{
	int dim , level , inC;
	int side[4];
	long long N[4];
	std::vector< Layer > layers;
	size_t nPad = 0 , nParams = 0;
	hipStream_t stream = nullptr;
	std::string err;
	DevBuf< float > master , grad , m , v , part , stats32;
	DevBuf< double > stats;
	DevBuf< bf16 > wb , wt;
	int adamStep = 0;
	// activations and their gradients for a batch capacity
	int cap = 0;
	enum { X0 , H0 , S0 , IN1 , H1 , S1 , IN2 , H2 , S2 , IN3 , H3 , S3 , U2 , G2 , A2 , D2 , U1 , G1 , A1 , D1 , U0 , G0 , A0 , D0 , NACT };
	DevBuf< bf16 > act[NACT] , gact[NACT] , dAcc , dy;
	DevBuf< float > y , target;
	int actLevel[NACT] , actC[NACT];
	// the training set on the device
	DevBuf< bf16 > dsIn;
	DevBuf< float > dsT;
	size_t dsCount = 0;

	Net( int d , int l , int ic ) : dim( d ) , level( l ) , inC( ic )
	{
		if( d!=3 && d!=4 ) throw std::string( "dimension must be 3 or 4" );
		if( l<3 || l>8 ) throw std::string( "network level must be in [3,8] (four grids of 2^l+1, 2^(l-1)+1, ... nodes per side)" );
		if( ic<1 || ic>CT ) throw std::string( "input channels must be in [1,16]" );
		for( int k=0 ; k<4 ; k++ ){ side[k] = ( 1<<( l-k ) ) + 1 ; N[k] = 1 ; for( int a=0 ; a<d ; a++ ) N[k] *= side[k]; }
		const int c[4] = { 16 , 32 , 64 , 128 };
		auto add = [&]( int lev , int ci , int co , int relu ){ Layer L ; L.level = lev , L.cin = ci , L.cout = co , L.relu = relu , L.cinP = PadC( ci ) , L.coutP = PadC( co ) ; layers.push_back( L ); };
		add( 0 , ic , c[0] , 1 ) , add( 0 , c[0] , c[0] , 1 );
		add( 1 , c[0] , c[1] , 1 ) , add( 1 , c[1] , c[1] , 1 );
		add( 2 , c[1] , c[2] , 1 ) , add( 2 , c[2] , c[2] , 1 );
		add( 3 , c[2] , c[3] , 1 ) , add( 3 , c[3] , c[3] , 1 );
		add( 2 , c[3] , c[2] , 1 ) , add( 2 , c[2] , c[2] , 1 );
		add( 1 , c[2] , c[1] , 1 ) , add( 1 , c[1] , c[1] , 1 );
		add( 0 , c[1] , c[0] , 1 ) , add( 0 , c[0] , c[0] , 1 );
		add( 0 , c[0] , 2 , 0 );
		const int W = Pow3( d );
		for( Layer &L : layers )
		{
			L.offW = nPad ; nPad += (size_t)W*L.coutP*L.cinP;
			L.offB = nPad ; nPad += L.coutP;
			nParams += (size_t)W*L.cout*L.cin + L.cout;
		}
		const int al[NACT] = { 0,0,0 , 1,1,1 , 2,2,2 , 3,3,3 , 2,2,2,2 , 1,1,1,1 , 0,0,0,0 };
		const int ac[NACT] = { PadC( ic ) , 16 , 16 , 16 , 32 , 32 , 32 , 64 , 64 , 64 , 128 , 128 , 128 , 64 , 64 , 64 , 64 , 32 , 32 , 32 , 32 , 16 , 16 , 16 };
		for( int k=0 ; k<NACT ; k++ ) actLevel[k] = al[k] , actC[k] = ac[k];
		HIPCHECK( hipStreamCreate( &stream ) );
		master.alloc( nPad ) , grad.alloc( nPad ) , m.alloc( nPad ) , v.alloc( nPad ) , wb.alloc( nPad ) , wt.alloc( nPad );
		HIPCHECK( hipMemsetAsync( master.p , 0 , nPad*sizeof(float) , stream ) );
		HIPCHECK( hipMemsetAsync( m.p , 0 , nPad*sizeof(float) , stream ) );
		HIPCHECK( hipMemsetAsync( v.p , 0 , nPad*sizeof(float) , stream ) );
		HIPCHECK( hipMemsetAsync( wb.p , 0 , nPad*sizeof(bf16) , stream ) );
		HIPCHECK( hipMemsetAsync( wt.p , 0 , nPad*sizeof(bf16) , stream ) );
		HIPCHECK( hipStreamSynchronize( stream ) );
	}
	~Net(){ if( stream ) (void)hipStreamDestroy( stream ); }

	long long count( int a , int batch ) const { return (long long)batch*N[ actLevel[a] ]*actC[a]; }
	void ensureBatch( int batch )
	{
		if( batch<=cap ) return;
		for( int a=0 ; a<NACT ; a++ ){ act[a].alloc( (size_t)count( a , batch ) ) , gact[a].alloc( (size_t)count( a , batch ) ); }
		dAcc.alloc( (size_t)batch*N[0]*128 );
		dy.alloc( (size_t)batch*N[0]*CT );
		y.alloc( (size_t)batch*N[0]*CT );
		target.alloc( (size_t)batch*N[0]*2 );
		stats.alloc( (size_t)batch*4 );
		cap = batch;
	}

	// canonical parameter order: per layer W[tap][co][ci] then b[co] (unpadded)
	void setParams( const float *p )
	{
		const int W = Pow3( dim );
		std::vector< float > h( nPad , 0.f );
		size_t k = 0;
		for( const Layer &L : layers )
		{
			for( int t=0 ; t<W ; t++ ) for( int co=0 ; co<L.cout ; co++ ) for( int ci=0 ; ci<L.cin ; ci++ ) h[ L.offW + ( (size_t)t*L.coutP + co )*L.cinP + ci ] = p[k++];
			for( int co=0 ; co<L.cout ; co++ ) h[ L.offB + co ] = p[k++];
		}
		master.upload( h.data() , nPad , stream );
		KToBf16<<< Blocks( (long long)nPad , 256 ) , 256 , 0 , stream >>>( master.p , (long long)nPad , wb.p );
		HIPCHECK( hipMemsetAsync( m.p , 0 , nPad*sizeof(float) , stream ) );
		HIPCHECK( hipMemsetAsync( v.p , 0 , nPad*sizeof(float) , stream ) );
		adamStep = 0;
		refreshTransposed();
		HIPCHECK( hipStreamSynchronize( stream ) );
	}
	void getPadded( const DevBuf< float > &buf , float *p )
	{
		const int W = Pow3( dim );
		std::vector< float > h( nPad );
		buf.download( h.data() , nPad , stream );
		HIPCHECK( hipStreamSynchronize( stream ) );
		size_t k = 0;
		for( const Layer &L : layers )
		{
			for( int t=0 ; t<W ; t++ ) for( int co=0 ; co<L.cout ; co++ ) for( int ci=0 ; ci<L.cin ; ci++ ) p[k++] = h[ L.offW + ( (size_t)t*L.coutP + co )*L.cinP + ci ];
			for( int co=0 ; co<L.cout ; co++ ) p[k++] = h[ L.offB + co ];
		}
	}
	void refreshTransposed( void )
	{
		const int W = Pow3( dim );
		for( const Layer &L : layers )
		{
			const long long tot = (long long)W*L.coutP*L.cinP;
			KTransposeMirror<<< Blocks( tot , 256 ) , 256 , 0 , stream >>>( wb.p + L.offW , W , L.coutP , L.cinP , wt.p + L.offW );
		}
		HIPCHECK( hipGetLastError() );
	}

	template< int Dim > void conv( const bf16 *in , int cinP , const bf16 *w , const float *bias , int coutP , int lev , int batch , int relu , const bf16 *skip , bf16 *out , bf16 *out2 , float *outF )
	{
		const dim3 grid( Blocks( N[lev] , 64 ) , batch , 1 );
		if( coutP%64==0 ){ dim3 g = grid ; g.z = coutP/64 ; KConv< Dim , 4 ><<< g , 256 , 0 , stream >>>( in , cinP , w , bias , coutP , side[lev] , N[lev] , relu , skip , out , out2 , outF ); }
		else if( coutP%32==0 ){ dim3 g = grid ; g.z = coutP/32 ; KConv< Dim , 2 ><<< g , 256 , 0 , stream >>>( in , cinP , w , bias , coutP , side[lev] , N[lev] , relu , skip , out , out2 , outF ); }
		else { dim3 g = grid ; g.z = coutP/16 ; KConv< Dim , 1 ><<< g , 256 , 0 , stream >>>( in , cinP , w , bias , coutP , side[lev] , N[lev] , relu , skip , out , out2 , outF ); }
	}
	template< int Dim > void down( const bf16 *fine , int lf , bf16 *coarse , int cP , int batch , int norm )
	{
		const long long cnt = (long long)batch*N[lf+1]*cP;
		KDown< Dim ><<< Blocks( cnt , 256 ) , 256 , 0 , stream >>>( fine , side[lf+1] , N[lf+1] , coarse , N[lf] , cP , cnt , norm );
	}
	template< int Dim > void up( const bf16 *coarse , int lc , bf16 *fine , int cP , int batch , int norm , int acc )
	{
		const long long cnt = (long long)batch*N[lc-1]*cP;
		KUp< Dim ><<< Blocks( cnt , 256 ) , 256 , 0 , stream >>>( coarse , side[lc] , N[lc] , fine , N[lc-1] , cP , cnt , norm , acc );
	}

	template< int Dim > void forward( int batch )
	{
		auto C = [&]( int k , int in , int out , const bf16 *skip=nullptr , int out2=-1 )
		{
			const Layer &L = layers[k];
			conv< Dim >( act[in].p , L.cinP , wb.p + L.offW , master.p + L.offB , L.coutP , L.level , batch , L.relu , skip , act[out].p , out2>=0 ? act[out2].p : nullptr , nullptr );
		};
		C( 0 , X0 , H0 ) , C( 1 , H0 , S0 );
		down< Dim >( act[S0].p , 0 , act[IN1].p , 16 , batch , 1 );
		C( 2 , IN1 , H1 ) , C( 3 , H1 , S1 );
		down< Dim >( act[S1].p , 1 , act[IN2].p , 32 , batch , 1 );
		C( 4 , IN2 , H2 ) , C( 5 , H2 , S2 );
		down< Dim >( act[S2].p , 2 , act[IN3].p , 64 , batch , 1 );
		C( 6 , IN3 , H3 ) , C( 7 , H3 , S3 );
		up< Dim >( act[S3].p , 3 , act[U2].p , 128 , batch , 0 , 0 );
		C( 8 , U2 , G2 , act[S2].p , A2 ) , C( 9 , A2 , D2 );
		up< Dim >( act[D2].p , 2 , act[U1].p , 64 , batch , 0 , 0 );
		C( 10 , U1 , G1 , act[S1].p , A1 ) , C( 11 , A1 , D1 );
		up< Dim >( act[D1].p , 1 , act[U0].p , 32 , batch , 0 , 0 );
		C( 12 , U0 , G0 , act[S0].p , A0 ) , C( 13 , A0 , D0 );
		const Layer &L = layers[14];
		conv< Dim >( act[D0].p , L.cinP , wb.p + L.offW , master.p + L.offB , L.coutP , 0 , batch , 0 , nullptr , nullptr , nullptr , y.p );
		HIPCHECK( hipGetLastError() );
	}

	// gradients of layer k from the gradient of its (post-activation) output
	// dOut: weight and bias gradients into grad, and (when dIn) the gradient
	// of its input
	template< int Dim > void layerBack( int k , const bf16 *in , const bf16 *outAct , const bf16 *dOut , bf16 *dIn , int batch )
	{
		const int W = Pow3( Dim );
		const Layer &L = layers[k];
		const long long total = (long long)batch*N[L.level];
		const bf16 *d = dOut;
		if( L.relu )
		{
			const long long cnt = total*L.coutP;
			KReluMask<<< Blocks( cnt , 256 ) , 256 , 0 , stream >>>( dOut , outAct , cnt , dAcc.p );
			d = dAcc.p;
		}
		const long long chunk = 1024;
		const long long nChunks = ( total + chunk - 1 )/chunk;
		const long long cnt = (long long)W*L.coutP*L.cinP;
		part.ensure( (size_t)( nChunks*cnt ) );
		KConvWGrad< Dim ><<< dim3( W , (unsigned)nChunks , ( L.coutP/16 )*( L.cinP/16 ) ) , 256 , 0 , stream >>>( in , L.cinP , d , L.coutP , side[L.level] , N[L.level] , total , chunk , part.p );
		KSumChunks<<< Blocks( cnt , 256 ) , 256 , 0 , stream >>>( part.p , nChunks , cnt , grad.p + L.offW );
		KBiasGrad<<< L.coutP , 256 , 0 , stream >>>( d , L.coutP , total , grad.p + L.offB );
		if( dIn ) conv< Dim >( d , L.coutP , wt.p + L.offW , nullptr , L.cinP , L.level , batch , 0 , nullptr , dIn , nullptr , nullptr );
		HIPCHECK( hipGetLastError() );
	}

	template< int Dim > void backward( int batch )
	{
		auto cntOf = [&]( int a ){ return count( a , batch ); };
		// output layer (its input D0)
		layerBack< Dim >( 14 , act[D0].p , nullptr , dy.p , gact[D0].p , batch );
		layerBack< Dim >( 13 , act[A0].p , act[D0].p , gact[D0].p , gact[A0].p , batch );
		// a0 = g0 + s0
		HIPCHECK( hipMemcpyAsync( gact[S0].p , gact[A0].p , cntOf( A0 )*sizeof(bf16) , hipMemcpyDeviceToDevice , stream ) );
		layerBack< Dim >( 12 , act[U0].p , act[G0].p , gact[A0].p , gact[U0].p , batch );
		down< Dim >( gact[U0].p , 0 , gact[D1].p , 32 , batch , 0 );                 // up^T
		layerBack< Dim >( 11 , act[A1].p , act[D1].p , gact[D1].p , gact[A1].p , batch );
		HIPCHECK( hipMemcpyAsync( gact[S1].p , gact[A1].p , cntOf( A1 )*sizeof(bf16) , hipMemcpyDeviceToDevice , stream ) );
		layerBack< Dim >( 10 , act[U1].p , act[G1].p , gact[A1].p , gact[U1].p , batch );
		down< Dim >( gact[U1].p , 1 , gact[D2].p , 64 , batch , 0 );
		layerBack< Dim >( 9 , act[A2].p , act[D2].p , gact[D2].p , gact[A2].p , batch );
		HIPCHECK( hipMemcpyAsync( gact[S2].p , gact[A2].p , cntOf( A2 )*sizeof(bf16) , hipMemcpyDeviceToDevice , stream ) );
		layerBack< Dim >( 8 , act[U2].p , act[G2].p , gact[A2].p , gact[U2].p , batch );
		down< Dim >( gact[U2].p , 2 , gact[S3].p , 128 , batch , 0 );
		layerBack< Dim >( 7 , act[H3].p , act[S3].p , gact[S3].p , gact[H3].p , batch );
		layerBack< Dim >( 6 , act[IN3].p , act[H3].p , gact[H3].p , gact[IN3].p , batch );
		up< Dim >( gact[IN3].p , 3 , gact[S2].p , 64 , batch , 1 , 1 );            // down^T, added to the skip's gradient
		layerBack< Dim >( 5 , act[H2].p , act[S2].p , gact[S2].p , gact[H2].p , batch );
		layerBack< Dim >( 4 , act[IN2].p , act[H2].p , gact[H2].p , gact[IN2].p , batch );
		up< Dim >( gact[IN2].p , 2 , gact[S1].p , 32 , batch , 1 , 1 );
		layerBack< Dim >( 3 , act[H1].p , act[S1].p , gact[S1].p , gact[H1].p , batch );
		layerBack< Dim >( 2 , act[IN1].p , act[H1].p , gact[H1].p , gact[IN1].p , batch );
		up< Dim >( gact[IN1].p , 1 , gact[S0].p , 16 , batch , 1 , 1 );
		layerBack< Dim >( 1 , act[H0].p , act[S0].p , gact[S0].p , gact[H0].p , batch );
		layerBack< Dim >( 0 , act[X0].p , act[H0].p , gact[H0].p , nullptr , batch );
	}

	// loss of the batch in y against target (both on the device), its
	// gradient into dy; returns the batch-mean loss
	double loss( int batch )
	{
		if( batch>256 ) throw std::string( "at most 256 samples per batch" );
		const int nb = (int)std::min< long long >( 64 , Blocks( N[0] , 256 ) );
		stats32.ensure( (size_t)batch*nb*4 );
		KLossPartials<<< dim3( nb , batch ) , 256 , 0 , stream >>>( y.p , target.p , N[0] , stats32.p );
		KLossStats<<< 1 , 256 , 0 , stream >>>( stats32.p , nb , batch , stats.p );
		KLossGrad<<< Blocks( (long long)batch*N[0]*CT , 256 ) , 256 , 0 , stream >>>( y.p , target.p , stats.p , N[0] , batch , dy.p );
		HIPCHECK( hipGetLastError() );
		std::vector< double > s( (size_t)batch*4 );
		stats.download( s.data() , s.size() , stream );
		HIPCHECK( hipStreamSynchronize( stream ) );
		double L = 0;
		for( int b=0 ; b<batch ; b++ ) L += ( s[b*4] + s[b*4+1] - 2.*std::sqrt( s[b*4+2]*s[b*4+2] + s[b*4+3]*s[b*4+3] ) ) / (double)N[0];
		return L / batch;
	}

	void adam( double lr )
	{
		adamStep++;
		const float c1 = 1.f - std::pow( 0.9f , (float)adamStep ) , c2 = 1.f - std::pow( 0.999f , (float)adamStep );
		KAdam<<< Blocks( (long long)nPad , 256 ) , 256 , 0 , stream >>>( master.p , grad.p , m.p , v.p , (long long)nPad , (float)lr , c1 , c2 , wb.p );
		refreshTransposed();
	}

	// host float samples (batch x N x inC) -> bf16 padded input rows
	void uploadInput( int batch , const float *in )
	{
		const int cP = PadC( inC );
		std::vector< bf16 > h( (size_t)batch*N[0]*cP , 0 );
		for( long long g=0 ; g<(long long)batch*N[0] ; g++ ) for( int c=0 ; c<inC ; c++ ) h[ g*cP + c ] = ToBf16( in[ g*inC + c ] );
		act[X0].upload( h.data() , h.size() , stream );
	}

	void runForward( int batch ){ if( dim==3 ) forward< 3 >( batch ) ; else forward< 4 >( batch ); }
	void runBackward( int batch ){ if( dim==3 ) backward< 3 >( batch ) ; else backward< 4 >( batch ); }
};

int Dim( const Net *n ){ return n->dim; }
int Level( const Net *n ){ return n->level; }
int InChannels( const Net *n ){ return n->inC; }
void ForwardDevice( Net *n , const bf16 *in , float *out , hipStream_t s )
{
	n->ensureBatch( 1 );
	HIPCHECK( hipStreamSynchronize( s ) );
	HIPCHECK( hipMemcpyAsync( n->act[Net::X0].p , in , (size_t)n->N[0]*PadC( n->inC )*sizeof(bf16) , hipMemcpyDeviceToDevice , n->stream ) );
	n->runForward( 1 );
	HIPCHECK( hipMemcpyAsync( out , n->y.p , (size_t)n->N[0]*CT*sizeof(float) , hipMemcpyDeviceToDevice , n->stream ) );
	HIPCHECK( hipStreamSynchronize( n->stream ) );
}
} // namespace unet
} // namespace epr

struct epr_unet { std::unique_ptr< epr::unet::Net > net; std::string err; };

template< typename F >
static int UGuard( epr_unet *u , F f )
{
	if( !u || !u->net ) return EPR_ERR_ARG;
	try { f() ; u->err.clear() ; return EPR_OK; }
	catch( const epr::HipError &e ){ u->err = e.msg ; return EPR_ERR_HIP; }
	catch( const std::string &s ){ u->err = s ; return EPR_ERR_STATE; }
	catch( const std::exception &e ){ u->err = e.what() ; return EPR_ERR_STATE; }
}

epr::unet::Net *EprUnetNet( epr_unet *u ){ return u ? u->net.get() : nullptr; }

extern "C" {

int epr_unet_create( int dim , int level , int inChannels , epr_unet **out )
{
	if( !out ) return EPR_ERR_ARG;
	*out = nullptr;
	return epr::ContextFree( [&]{ epr_unet *u = new epr_unet ; try { u->net.reset( new epr::unet::Net( dim , level , inChannels ) ); } catch( ... ){ delete u ; throw; } *out = u; } );
}
void epr_unet_destroy( epr_unet *u ){ delete u; }
const char *epr_unet_last_error( const epr_unet *u ){ return u ? u->err.c_str() : epr::ThreadError().c_str(); }
size_t epr_unet_num_params( const epr_unet *u ){ return u && u->net ? u->net->nParams : 0; }
size_t epr_unet_num_nodes( const epr_unet *u ){ return u && u->net ? (size_t)u->net->N[0] : 0; }

int epr_unet_set_params( epr_unet *u , const float *p ){ if( !p ) return EPR_ERR_ARG ; return UGuard( u , [&]{ u->net->setParams( p ); } ); }
int epr_unet_get_params( epr_unet *u , float *p ){ if( !p ) return EPR_ERR_ARG ; return UGuard( u , [&]{ u->net->getPadded( u->net->master , p ); } ); }

int epr_unet_forward( epr_unet *u , int batch , const float *in , float *out )
{
	if( !in || !out || batch<1 ) return EPR_ERR_ARG;
	return UGuard( u , [&]
	{
		epr::unet::Net &n = *u->net;
		n.ensureBatch( batch );
		n.uploadInput( batch , in );
		n.runForward( batch );
		std::vector< float > y( (size_t)batch*n.N[0]*epr::unet::CT );
		n.y.download( y.data() , y.size() , n.stream );
		HIPCHECK( hipStreamSynchronize( n.stream ) );
		for( long long g=0 ; g<(long long)batch*n.N[0] ; g++ ) out[2*g] = y[ g*epr::unet::CT ] , out[2*g+1] = y[ g*epr::unet::CT+1 ];
	} );
}

int epr_unet_loss_gradients( epr_unet *u , int batch , const float *in , const float *target , double *loss , float *grad )
{
	if( !in || !target || batch<1 ) return EPR_ERR_ARG;
	return UGuard( u , [&]
	{
		epr::unet::Net &n = *u->net;
		n.ensureBatch( batch );
		n.uploadInput( batch , in );
		n.target.upload( target , (size_t)batch*n.N[0]*2 , n.stream );
		n.runForward( batch );
		const double L = n.loss( batch );
		n.runBackward( batch );
		if( loss ) *loss = L;
		if( grad ) n.getPadded( n.grad , grad );
		HIPCHECK( hipStreamSynchronize( n.stream ) );
	} );
}

int epr_unet_set_dataset( epr_unet *u , size_t count , const float *in , const float *target )
{
	if( !count || !in || !target ) return EPR_ERR_ARG;
	return UGuard( u , [&]
	{
		epr::unet::Net &n = *u->net;
		const int cP = epr::unet::PadC( n.inC );
		const size_t rows = count*(size_t)n.N[0];
		std::vector< epr::unet::bf16 > h( rows*cP , 0 );
		for( size_t g=0 ; g<rows ; g++ ) for( int c=0 ; c<n.inC ; c++ ) h[ g*cP + c ] = epr::unet::ToBf16( in[ g*n.inC + c ] );
		n.dsIn.upload( h.data() , h.size() , n.stream );
		n.dsT.upload( target , rows*2 , n.stream );
		HIPCHECK( hipStreamSynchronize( n.stream ) );
		n.dsCount = count;
	} );
}

int epr_unet_train( epr_unet *u , int steps , int batch , double lr , unsigned seed , double *losses )
{
	if( steps<0 || batch<1 ) return EPR_ERR_ARG;
	return UGuard( u , [&]
	{
		epr::unet::Net &n = *u->net;
		if( !n.dsCount ) throw std::string( "no training set (epr_unet_set_dataset)" );
		if( (size_t)batch>n.dsCount ) throw std::string( "batch larger than the training set" );
		n.ensureBatch( batch );
		const int cP = epr::unet::PadC( n.inC );
		const size_t rowIn = (size_t)n.N[0]*cP , rowT = (size_t)n.N[0]*2;
		std::mt19937 rng( seed );
		std::vector< size_t > perm( n.dsCount );
		std::iota( perm.begin() , perm.end() , (size_t)0 );
		size_t pos = perm.size();
		for( int s=0 ; s<steps ; s++ )
		{
			for( int b=0 ; b<batch ; b++ )
			{
				if( pos==perm.size() ){ std::shuffle( perm.begin() , perm.end() , rng ) ; pos = 0; }
				const size_t k = perm[pos++];
				HIPCHECK( hipMemcpyAsync( n.act[epr::unet::Net::X0].p + b*rowIn , n.dsIn.p + k*rowIn , rowIn*sizeof(epr::unet::bf16) , hipMemcpyDeviceToDevice , n.stream ) );
				HIPCHECK( hipMemcpyAsync( n.target.p + b*rowT , n.dsT.p + k*rowT , rowT*sizeof(float) , hipMemcpyDeviceToDevice , n.stream ) );
			}
			n.runForward( batch );
			const double L = n.loss( batch );
			n.runBackward( batch );
			n.adam( lr );
			if( losses ) losses[s] = L;
		}
		HIPCHECK( hipStreamSynchronize( n.stream ) );
	} );
}

} // extern "C"
