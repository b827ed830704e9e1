// Golden-vector harness around the reference's MarchingSimplices::Orient
// (Misha/MarchingSimplices.h:179-246).
//
// Test infrastructure only: compiled against the reference's own headers where
// they lie (oracle/Makefile, target `ref`) into oracle/_ref/ref_orient and run
// here to write fixtures under tests/golden/.
//
// Input (binary, stdin path argv[1]): u32 k (vertices per simplex, 2 or 3),
// u64 n, n*k u32 indices.  Output (argv[2]): the oriented indices, n*k u32.

#include <stdio.h>
#include <stdlib.h>
#include <vector>

#include "Misha/Miscellany.h"
#include "Misha/Geometry.h"
#include "Misha/MarchingSimplices.h"

using namespace MishaK;

template< unsigned int K >
static void Run( const std::vector< unsigned int > &in , FILE *out )
{
	std::vector< SimplexIndex< K , unsigned int > > S( in.size()/(K+1) );
	for( size_t i=0 ; i<S.size() ; i++ ) for( unsigned int j=0 ; j<=K ; j++ ) S[i][j] = in[ i*(K+1)+j ];
	MarchingSimplices::Orient( S );
	for( size_t i=0 ; i<S.size() ; i++ ) for( unsigned int j=0 ; j<=K ; j++ ){ unsigned int v = S[i][j] ; fwrite( &v , sizeof(v) , 1 , out ); }
}

int main( int argc , char *argv[] )
{
	if( argc!=3 ){ fprintf( stderr , "usage: %s <in> <out>\n" , argv[0] ) ; return 1; }
	FILE *fp = fopen( argv[1] , "rb" );
	if( !fp ) return 1;
	unsigned int k ; unsigned long long n;
	if( fread( &k , sizeof(k) , 1 , fp )!=1 || fread( &n , sizeof(n) , 1 , fp )!=1 ) return 1;
	std::vector< unsigned int > in( n*k );
	if( n && fread( in.data() , sizeof(unsigned int) , in.size() , fp )!=in.size() ) return 1;
	fclose( fp );
	FILE *out = fopen( argv[2] , "wb" );
	if( !out ) return 1;
	if( k==2 ) Run< 1 >( in , out );
	else if( k==3 ) Run< 2 >( in , out );
	else return 1;
	fclose( out );
	return 0;
}
