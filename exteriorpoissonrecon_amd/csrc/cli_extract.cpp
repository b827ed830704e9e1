// Extract command line over libepr_hip.so: the reference's Extract
// (Extract/Extract.cpp:48-366) with the same flags and outputs.
//
//   --in <.tree>           a dense or adapted hierarchy written by ExteriorPoissonRecon
//   --out <.ply>           the joint zero set with per-vertex frames n_c_d (:263-282)
//   --levelSets <.ply>     both codimension-1 zero sets, coloured (:285-327)
//   --grid <header>        header.0.grid, header.1.grid of the finest values (:132-144)
//   --density <tree> --trimDensity <v>   component trimming by sample density (:217-238)
//   --verbose
// Errors print "[ERROR] ..." to stderr and exit with status 0, as MK_ERROR_OUT
// does (Misha/Exceptions.h:101-108).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "epr.h"

static void Error( const std::string &msg )
{
	fprintf( stderr , "[ERROR] %s\n" , msg.c_str() );
	exit( 0 );
}

static void Usage( const char *ex )
{
	printf( "Usage %s:\n" , ex );
	printf( "\t --in <input grid>\n" );
	printf( "\t[--density <density grid>]\n" );
	printf( "\t[--out <output mesh>]\n" );
	printf( "\t[--levelSets <output level sets>]\n" );
	printf( "\t[--grid <output grid header>]\n" );
	printf( "\t[--trimDensity <trim value>=%f]\n" , 0. );
	printf( "\t[--verbose]\n" );
}

struct Mesh
{
	size_t nV = 0 , nS = 0;
	int vd = 0 , k = 0;
	std::vector< double > v;
	std::vector< unsigned > s;
};

static Mesh Take( epr_mesh *m )
{
	Mesh o;
	epr_mesh_info( m , &o.nV , &o.vd , &o.nS , &o.k );
	o.v.resize( o.nV*o.vd ) , o.s.resize( o.nS*o.k );
	epr_mesh_copy( m , o.v.data() , o.s.data() );
	epr_mesh_destroy( m );
	return o;
}

// PLY::WritePolygons in PLY_BINARY_NATIVE (Misha/Ply.inl:673-737, PlyFile.inl:314-356):
// vertex properties as given, faces as "list int int vertex_indices"
static void WritePly( const std::string &path , const std::vector< std::string > &props , size_t nV , size_t rowBytes ,
                      const std::vector< unsigned char > &rows , size_t nS , int k , const unsigned *S )
{
	FILE *fp = fopen( path.c_str() , "wb" );
	if( !fp ) Error( "could not write ply file: " + path );
	fprintf( fp , "ply\nformat binary_little_endian 1.0\n" );
	fprintf( fp , "element vertex %llu\n" , (unsigned long long)nV );
	for( const std::string &p : props ) fprintf( fp , "property %s\n" , p.c_str() );
	fprintf( fp , "element face %llu\n" , (unsigned long long)nS );
	fprintf( fp , "property list int int vertex_indices\nend_header\n" );
	if( nV ) fwrite( rows.data() , rowBytes , nV , fp );
	std::vector< int > face( k+1 );
	face[0] = k;
	for( size_t i=0 ; i<nS ; i++ )
	{
		for( int j=0 ; j<k ; j++ ) face[1+j] = (int)S[ i*k+j ];
		fwrite( face.data() , sizeof(int) , k+1 , fp );
	}
	fclose( fp );
}

static const char *Axis[] = { "x" , "y" , "z" , "w" };

int main( int argc , char *argv[] )
{
	std::string in , out , levelSets , grid , densityPath;
	double trim = 0;
	bool verbose = false;
	for( int i=1 ; i<argc ; i++ )
	{
		std::string a = argv[i];
		if( a.rfind( "--" , 0 )!=0 ){ fprintf( stderr , "[WARNING] Unrecognized argument: %s\n" , argv[i] ) ; continue; }
		a = a.substr( 2 );
		auto Next = [&]( void ) -> std::string { if( i+1>=argc ) Error( "missing value for --" + a ) ; return argv[++i]; };
		if     ( a=="in" ) in = Next();
		else if( a=="out" ) out = Next();
		else if( a=="levelSets" ) levelSets = Next();
		else if( a=="grid" ) grid = Next();
		else if( a=="density" ) densityPath = Next();
		else if( a=="trimDensity" ) trim = atof( Next().c_str() );
		else if( a=="verbose" ) verbose = true;
		else fprintf( stderr , "[WARNING] Unrecognized argument: --%s\n" , a.c_str() );
	}
	if( in.empty() ){ Usage( argv[0] ) ; return EXIT_SUCCESS; }

	epr_iso *iso = nullptr;
	if( epr_iso_open( in.c_str() , &iso ) ) Error( epr_iso_last_error( nullptr ) );
	int dim = 0 , depth = 0 , adapted = 0;
	epr_iso_info( iso , &dim , &depth , &adapted , nullptr );
	if( verbose ) printf( "%25s: %.3f (s)\n" , "Iso-tree" , epr_iso_seconds( iso , 0 ) );
	if( !grid.empty() && epr_iso_write_grid( iso , grid.c_str() ) ) Error( epr_iso_last_error( iso ) );
	if( !densityPath.empty() && epr_iso_read_density( iso , densityPath.c_str() ) ) Error( epr_iso_last_error( iso ) );

	const bool all = levelSets.empty();
	if( !out.empty() )
	{
		epr_mesh *m = nullptr;
		if( epr_iso_extract_trimmed( iso , all ? 1 : 0 , densityPath.empty() ? 0. : trim , &m ) ) Error( epr_iso_last_error( iso ) );
		Mesh M = Take( m );
		if( verbose && !densityPath.empty() && trim>0 )
		{
			size_t n = 0;
			epr_iso_component_densities( iso , 0 , nullptr , &n );
			std::vector< double > mx( n );
			epr_iso_component_densities( iso , n , mx.data() , &n );
			// std::cout's default formatting (6 significant digits), as Extract.cpp:227 prints
			for( size_t i=0 ; i<n ; i++ ) printf( "Max Density[ %zu ] %g\n" , i , mx[i] );
		}
		if( verbose )
		{
			printf( "%25s: %.3f (s)\n" , "Triangulation" , epr_iso_seconds( iso , 1 ) );
			printf( "%25s: %.3f (s)\n" , "Level set" , epr_iso_seconds( iso , 2 ) );
			printf( "%25s: %.3f (s)\n" , "Oriented" , epr_iso_seconds( iso , 3 ) );
			printf( "%25s: %.3f (s)\n" , "Frames" , epr_iso_seconds( iso , 4 ) );
		}
		std::vector< std::string > props;
		for( int d=0 ; d<dim ; d++ ) props.push_back( std::string( "double " ) + Axis[d] );
		for( int c=0 ; c<2 ; c++ ) for( int d=0 ; d<dim ; d++ ) props.push_back( "double n_" + std::to_string( c ) + "_" + std::to_string( d ) );
		std::vector< unsigned char > rows( M.v.size()*sizeof(double) );
		if( !rows.empty() ) memcpy( rows.data() , M.v.data() , rows.size() );
		WritePly( out , props , M.nV , 3*dim*sizeof(double) , rows , M.nS , M.k , M.s.data() );
	}
	if( !levelSets.empty() )
	{
		// SurfaceColor (Extract.cpp:80-88), written as uchar (RGBColorFactory's default on disk)
		const unsigned char color[2][3] = { { 223 , 41 , 53 } , { 55 , 114 , 255 } };
		std::vector< std::string > props;
		for( int d=0 ; d<dim ; d++ ) props.push_back( std::string( "double " ) + Axis[d] );
		props.push_back( "uchar red" ) , props.push_back( "uchar green" ) , props.push_back( "uchar blue" );
		const size_t rowBytes = dim*sizeof(double) + 3;
		std::vector< unsigned char > rows;
		std::vector< unsigned > S;
		size_t nV = 0;
		for( int c=0 ; c<2 ; c++ )
		{
			epr_mesh *m = nullptr;
			if( epr_iso_level_set( iso , c , &m ) ) Error( epr_iso_last_error( iso ) );
			Mesh M = Take( m );
			for( size_t i=0 ; i<M.nV ; i++ )
			{
				const size_t o = rows.size();
				rows.resize( o + rowBytes );
				memcpy( &rows[o] , &M.v[ i*dim ] , dim*sizeof(double) );
				memcpy( &rows[ o + dim*sizeof(double) ] , color[c] , 3 );
			}
			for( size_t i=0 ; i<M.s.size() ; i++ ) S.push_back( M.s[i] + (unsigned)nV );
			nV += M.nV;
		}
		WritePly( levelSets , props , nV , rowBytes , rows , S.size()/dim , dim , S.data() );
	}
	epr_iso_destroy( iso );
	return EXIT_SUCCESS;
}
