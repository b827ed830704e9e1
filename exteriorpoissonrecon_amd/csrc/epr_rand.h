// A private copy of glibc's rand() generator, so that every reconstruction
// draws the reference's sequence without touching the process-wide rand()
// state (which other contexts, threads and the host program share).
//
// The reference draws its random initial guess with srand(seed) followed by
// Random<double>() (Misha/Geometry.inl:429-438: two rand() calls per value),
// ExteriorPoisson.h:333-341 and ExteriorPoissonRecon.cpp:220; --jitter draws
// from the unseeded sequence (ExteriorPoissonRecon.cpp:174-180).  glibc's
// rand() is random() with its default state: the additive lagged-Fibonacci
// generator r[i] = r[i-3] + r[i-31] (mod 2^32) over 31 words, seeded by the
// Park-Miller minimal standard generator (16807 x mod 2^31-1, computed with
// Schrage's decomposition), with the first 310 outputs discarded and each
// output the word shifted right by one bit.
#pragma once
#include <cstdint>

namespace epr
{
struct GlibcRand
{
	static constexpr int Deg = 31 , Sep = 3;
	static constexpr int Max = 2147483647;   // RAND_MAX of glibc
	uint32_t r[Deg];
	int f = Sep , b = 0;   // the two taps, Sep apart

	explicit GlibcRand( unsigned int seed=1 ){ srand( seed ); }

	void srand( unsigned int seed )
	{
		if( seed==0 ) seed = 1;
		int32_t w = (int32_t)seed;
		r[0] = (uint32_t)w;
		for( int i=1 ; i<Deg ; i++ )
		{
			// w = 16807 w mod (2^31-1), without overflow
			const int32_t hi = w / 127773 , lo = w % 127773;
			w = 16807*lo - 2836*hi;
			if( w<0 ) w += Max;
			r[i] = (uint32_t)w;
		}
		f = Sep , b = 0;
		for( int i=0 ; i<10*Deg ; i++ ) (void)rand();
	}

	int rand( void )
	{
		r[f] += r[b];
		const int out = (int)( r[f]>>1 );
		if( ++f==Deg ) f = 0;
		if( ++b==Deg ) b = 0;
		return out;
	}

	// Random<double>() of Misha/Geometry.inl:429-438 over this sequence
	double random01( void )
	{
		const unsigned long long M = (unsigned long long)Max + 1;
		const double M2 = double( M*M - 1 );
		const unsigned long long r1 = (unsigned long long)rand() , r2 = (unsigned long long)rand();
		const long long v = (long long)( r1*M + r2 );
		return double( v ) / M2;
	}
};
} // namespace epr
