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

// dCells: ncell x Dim cell indices at depth D (device); out receives depths 0..D
template< int Dim >
void BuildAdaptedDevice( const int *dCells , int ncell , int D , int K , hipStream_t st , std::vector< AdaptedLevelDev > &out );
} // namespace epr
