// The learned initializer's network as the rest of the library sees it: BF16
// storage helpers and a forward pass over device buffers (csrc/epr_unet.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>

namespace epr
{
namespace unet
{
typedef unsigned short bf16;   // bfloat16 bits

// round to nearest even (NaN stays NaN)
__host__ __device__ inline bf16 ToBf16( float f )
{
	uint32_t u;
#ifdef __HIP_DEVICE_COMPILE__
	u = __float_as_uint( f );
#else
	memcpy( &u , &f , 4 );
#endif
	if( ( u & 0x7f800000u )==0x7f800000u ) return (bf16)( ( u>>16 ) | ( ( u & 0xffffu ) ? 0x40u : 0u ) );
	u += 0x7fffu + ( ( u>>16 ) & 1u );
	return (bf16)( u>>16 );
}
__host__ __device__ inline float FromBf16( bf16 h )
{
	const uint32_t u = (uint32_t)h<<16;
#ifdef __HIP_DEVICE_COMPILE__
	return __uint_as_float( u );
#else
	float f ; memcpy( &f , &u , 4 ) ; return f;
#endif
}

// channel storage: every activation row is padded to a multiple of 16 channels
// (one MFMA K/N tile), the padding held at zero
constexpr int CT = 16;
__host__ __device__ constexpr int PadC( int c ){ return ( c + CT - 1 )/CT*CT; }

struct Net;
int Dim( const Net *n );
int Level( const Net *n );          // the network's finest grid: (2^level+1)^dim nodes
int InChannels( const Net *n );
// One sample on the device: in is N x PadC(InChannels) bf16 (node index with
// the last coordinate fastest, the hat functions' order), out receives
// N x CT floats, channels 0 and 1 the predicted (x, y).
void ForwardDevice( Net *n , const bf16 *in , float *out , hipStream_t s );
} // namespace unet
} // namespace epr
