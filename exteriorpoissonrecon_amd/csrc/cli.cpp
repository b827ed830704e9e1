// ExteriorPoissonRecon command line over libepr_hip.so (dense --aRadius -1 and adapted --aRadius 0..5 paths).
//
// Same flags, defaults and outputs as ExteriorPoissonRecon/ExteriorPoissonRecon.cpp:52-279:
// reads <in>.ply, writes <out>.tree and <out>.density.tree (reference byte layouts) and, with --grid,
// a .grid of the finest (phi1,phi2) values, with --densityGrid a .grid of the
// estimated samples per cell at the finest nodes.  Errors print "[ERROR] ..." to
// stderr and exit with status 0, as MK_ERROR_OUT does (Misha/Exceptions.h:101-108).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "epr.h"
#include "ply.h"

static void Error( const std::string &msg )
{
	fprintf( stderr , "[ERROR] %s\n" , msg.c_str() );
	exit( 0 );
}

static void Usage( const char *ex )
{
	printf( "Usage %s:\n" , ex );
	printf( "\t --in <input points and frames>\n" );
	printf( "\t[--out <output voxel grid header>]\n" );
	printf( "\t[--grid <output .grid of the finest (phi1,phi2)>]\n" );
	printf( "\t[--densityGrid <output .grid of the samples per cell at the finest nodes; needs --out>]\n" );
	printf( "\t[--depth <grid depth>=5]\n" );
	printf( "\t[--minSolveDepth <minimum hierarchy solver depth>=0]\n" );
	printf( "\t[--maxSolveDepth <maximum hierarchy solver depth>]\n" );
	printf( "\t[--nSigma <noise-to-weight sigma>=0.250000]\n" );
	printf( "\t[--dRadius <density kernel radius>=1]\n" );
	printf( "\t[--tRadius <tensor kernel radius>=1]\n" );
	printf( "\t[--aRadius <adaptive radius>=3]\n" );
	printf( "\t[--jitter <jitter magnitude>=%f]\n" , 0. );
	printf( "\t[--fp32] (Gauss-Seidel iterations in FP32 but each level's last, on the larger levels)\n" );
	printf( "\t[--gsIters <gs iterations>=10]\n" );
	printf( "\t[--seed <random seed>=0]\n" );
	printf( "\t[--sWeight <screening weight>=100]\n" );
	printf( "\t[--dWeight <dirichlet weight>=0.00625]\n" );
	printf( "\t[--scale <bounding box scale>=1.100000]\n" );
	printf( "\t[--iMult <iteration multiplier>=2.000000]\n" );
	printf( "\t[--separate]\n" );
	printf( "\t[--singleLevel]\n" );
	printf( "\t[--verbose]\n" );
}

int main( int argc , char *argv[] )
{
	epr_params p;
	epr_default_params( &p );
	std::string in , out , grid , densityGrid;
	long long aRadius = 3;
	bool project = false;
	for( int i=1 ; i<argc ; i++ )
	{
		std::string a = argv[i];
		if( a.rfind( "--" , 0 )!=0 ){ fprintf( stderr , "[WARNING] Expected parameter name: %s\n" , a.c_str() ) ; continue; }
		a = a.substr( 2 );
		auto val = [&]() -> std::string { if( i+1>=argc ){ fprintf( stderr , "[WARNING] Missing value for --%s\n" , a.c_str() ) ; return ""; } return argv[++i]; };
		if     ( a=="in" ) in = val();
		else if( a=="out" ) out = val();
		else if( a=="grid" ) grid = val();
		else if( a=="densityGrid" ) densityGrid = val();
		else if( a=="depth" ) p.depth = atoi( val().c_str() );
		else if( a=="seed" ) p.seed = atoi( val().c_str() );
		else if( a=="tRadius" ) p.tRadius = atoi( val().c_str() );
		else if( a=="dRadius" ) p.dRadius = atoi( val().c_str() );
		else if( a=="aRadius" ) aRadius = atoll( val().c_str() );
		else if( a=="minSolveDepth" ) p.minSolveDepth = atoi( val().c_str() );
		else if( a=="maxSolveDepth" ) p.maxSolveDepth = atoi( val().c_str() );
		else if( a=="gsIters" ) p.gsIters = atoi( val().c_str() );
		else if( a=="nSigma" ) p.nSigma = atof( val().c_str() );
		else if( a=="scale" ) p.scale = atof( val().c_str() );
		else if( a=="sWeight" ) p.sWeight = atof( val().c_str() );
		else if( a=="dWeight" ) p.dWeight = atof( val().c_str() );
		else if( a=="iMult" ) p.iMult = atof( val().c_str() );
		else if( a=="jitter" ) p.jitter = atof( val().c_str() );
		else if( a=="singleLevel" ) p.singleLevel = 1;
		else if( a=="separate" ) p.separate = 1;
		else if( a=="fp32" ) p.fp32Sweeps = 1;
		else if( a=="project" ) project = true;
		else if( a=="verbose" ) p.verbose = 1;
		else fprintf( stderr , "[WARNING] Invalid parameter: --%s\n" , a.c_str() );
	}
	if( in.empty() ){ Usage( argv[0] ) ; return EXIT_SUCCESS; }
	// "-1" reads as UINT_MAX in the reference (an unsigned parameter) and selects the dense grid
	if( aRadius==-1 || aRadius==(long long)0xFFFFFFFFLL ) p.aRadius = -1;
	else if( aRadius<0 || aRadius>5 ) Error( "Only kernel radius in the range [0,5] supported: " + std::to_string( (unsigned int)aRadius ) );
	else p.aRadius = (int)aRadius;
	if( project ) Error( "--project is not supported" );
	if( !densityGrid.empty() && out.empty() ) Error( "--densityGrid needs --out (it is computed from <out>.density.tree)" );
	if( p.maxSolveDepth>p.depth ) p.maxSolveDepth = p.depth;
	if( in.size()<4 || in.substr( in.size()-4 )!=".ply" ) Error( "Only .ply files supported" );
	std::string err;
	epr::PlySamples s = epr::ReadPlySamples( in , err );
	if( !s.dim ) Error( "Invalid sample file: " + err );
	auto t0 = std::chrono::steady_clock::now();
	epr_ctx *ctx = nullptr;
	auto Check = [&]( int rc ){ if( rc!=EPR_OK ) Error( ctx ? epr_last_error( ctx ) : "epr_create failed" ); };
	Check( epr_create( s.dim , &p , &ctx ) );
	Check( epr_set_samples( ctx , s.size() , s.pos.data() , s.skew.data() ) );
	Check( epr_build( ctx ) );
	if( p.verbose ) printf( "%32s: %.3f (s)\n" , "Reconstructor" , epr_last_seconds( ctx , 0 ) );
	double E[2];
	Check( epr_solve( ctx , E ) );
	printf( "Error: %.17g -> %.17g = %.17g + %.17g\n" , 0. , E[0]+E[1] , E[0] , E[1] );
	if( !out.empty() )
	{
		Check( epr_write_tree( ctx , ( out + ".tree" ).c_str() , epr_indexer_hash( s.dim , p.aRadius!=-1 ) ) );
		Check( epr_write_density( ctx , ( out + ".density.tree" ).c_str() ) );
	}
	if( !grid.empty() ) Check( epr_write_grid( ctx , grid.c_str() ) );
	if( !densityGrid.empty() && epr_density_grid( ( out + ".density.tree" ).c_str() , s.dim , p.depth , densityGrid.c_str() )!=EPR_OK ) Error( epr_last_error( nullptr ) );
	if( p.verbose ) printf( "%32s: %.3f (s)\n" , "Executed" , std::chrono::duration< double >( std::chrono::steady_clock::now()-t0 ).count() );
	epr_destroy( ctx );
	return EXIT_SUCCESS;
}
