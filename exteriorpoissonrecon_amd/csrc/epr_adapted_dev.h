// The adapted hierarchy built on the device (epr_adapted.hip).
#pragma once
#include <vector>

#include "epr_dev.h"

namespace epr
{
// One depth of an adapted hierarchy: the same sets and numbering as the host
// AdaptedHierarchy< Dim >::buildLevels (epr_adapted.h), in device memory.
struct AdaptedLevelDev
{
	DevBuf< unsigned char > fm , em;          // function flags over (2^d+1)^Dim (1 active, 2 padding), element flags over (2^d)^Dim
	DevBuf< unsigned int > forder , eorder;   // dense index of every indexed function / element in the hierarchy's numbering
	DevBuf< unsigned int > act;               // the active functions grouped by colour, increasing within a colour
	size_t nF = 0 , nActive = 0 , nE = 0;
	long long actStart[17] = {};              // colour c occupies act[ actStart[c] , actStart[c+1] )
};

// Device temporaries of the construction, kept between builds (a build of
// the same size allocates nothing).
struct AdaptedScratch
{
	DevBuf< unsigned char > tmp , ga , gb;
	DevBuf< unsigned int > sel , vals , ckeys , ckeys2;
	DevBuf< unsigned long long > keys , keys2 , pkey;
	DevBuf< unsigned int > dn;
	DevBuf< long long > dstarts;
	// pinned host words for the counts and colour starts read back at every depth
	unsigned int *hCount = nullptr;
	long long *hStarts = nullptr;
	AdaptedScratch( void ){}
	AdaptedScratch( const AdaptedScratch & ) = delete;
	AdaptedScratch &operator=( const AdaptedScratch & ) = delete;
	~AdaptedScratch(){ if( hCount ) (void)hipHostFree( hCount ) ; if( hStarts ) (void)hipHostFree( hStarts ); }
};

// dCells: ncell x Dim cell indices at depth D (device); out[d] (d = 0..D)
// receives depth d, its buffers grown as needed and reused otherwise
template< int Dim >
void BuildAdaptedDevice( const int *dCells , int ncell , int D , int K , hipStream_t st , const std::vector< AdaptedLevelDev * > &out , AdaptedScratch &scratch );
} // namespace epr
