// The reference's Sample command line (Sample/Sample.cpp:39-238) over
// libepr_hip.so's epr_sample_generate: the same flags and defaults, and the
// same PLY bytes (PLY::WritePoints with a position and a "skew" matrix
// factory, binary little endian: x y z [w] then skew_c_r, all double).
// Without --type it prints the usage and exits; errors print "[ERROR] ..."
// to stderr and exit with status 0 (MK_ERROR_OUT, Misha/Exceptions.h:101-108).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "epr.h"

static void Usage( const char *ex )
{
	printf( "Usage %s:\n" , ex );
	printf( "\t --type <sample type>\n" );
	for( const char *t : { "torus_knot:<p>:<q>" , "spiral:<r>" , "borromean_rings" , "spherical:1:<n>:<l>" , "clifford_torus" , "hopf_torus:<number of lobes>:<amplitude>" } ) printf( "\t\t%s\n" , t );
	printf( "\t[--res <sample resolution>=%d]\n" , 1024 );
	printf( "\t[--aNoise <angular noise sigma (in units of degrees)>=%f]\n" , 0. );
	printf( "\t[--pNoise <positional noise sigma (in normalized units)>=%f]\n" , 0. );
	printf( "\t[--sNoise <spot noise (fraction of input samples)>=%f]\n" , 0. );
	printf( "\t[--out <output file>]\n" );
	printf( "\t[--regular]\n" );
}

int main( int argc , char **argv )
{
	std::string type , out;
	unsigned int res = 1024;
	double aNoise = 0 , pNoise = 0 , sNoise = 0;
	bool regular = false;
	for( int i=1 ; i<argc ; i++ )
	{
		const std::string a = argv[i];
		auto val = [&]( void ) -> const char * { if( i+1>=argc ){ fprintf( stderr , "[ERROR] missing value for %s\n" , a.c_str() ) ; exit( 0 ); } return argv[++i]; };
		if( a=="--type" ) type = val();
		else if( a=="--res" ) res = (unsigned int)strtoul( val() , nullptr , 10 );
		else if( a=="--aNoise" ) aNoise = atof( val() );
		else if( a=="--pNoise" ) pNoise = atof( val() );
		else if( a=="--sNoise" ) sNoise = atof( val() );
		else if( a=="--out" ) out = val();
		else if( a=="--regular" ) regular = true;
		else fprintf( stderr , "[WARNING] unknown argument: %s\n" , a.c_str() );
	}
	if( type.empty() ){ Usage( argv[0] ) ; return EXIT_SUCCESS; }
	int dim = 0;
	size_t n = 0;
	if( epr_sample_generate( type.c_str() , res , regular ? 1 : 0 , aNoise , pNoise , sNoise , &dim , &n , nullptr , nullptr , 0 ) )
	{
		fprintf( stderr , "[ERROR] %s\n" , epr_last_error( nullptr ) );
		return 0;
	}
	std::vector< double > pos( n*dim ) , skew( n*dim*dim );
	if( epr_sample_generate( type.c_str() , res , regular ? 1 : 0 , aNoise , pNoise , sNoise , &dim , &n , pos.data() , skew.data() , n ) )
	{
		fprintf( stderr , "[ERROR] %s\n" , epr_last_error( nullptr ) );
		return 0;
	}
	if( out.empty() ) return EXIT_SUCCESS;
	FILE *fp = fopen( out.c_str() , "wb" );
	if( !fp ){ fprintf( stderr , "[ERROR] cannot open %s\n" , out.c_str() ) ; return 0; }
	const char *names = "xyzw";
	fprintf( fp , "ply\nformat binary_little_endian 1.0\nelement vertex %zu\n" , n );
	for( int d=0 ; d<dim ; d++ ) fprintf( fp , "property double %c\n" , names[d] );
	for( int c=0 ; c<dim ; c++ ) for( int r=0 ; r<dim ; r++ ) fprintf( fp , "property double skew_%d_%d\n" , c , r );
	fprintf( fp , "end_header\n" );
	for( size_t i=0 ; i<n ; i++ )
	{
		fwrite( &pos[i*dim] , sizeof(double) , dim , fp );
		fwrite( &skew[i*dim*dim] , sizeof(double) , dim*dim , fp );
	}
	fclose( fp );
	return EXIT_SUCCESS;
}
