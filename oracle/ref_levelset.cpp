// Golden-vector harness around the reference's simplicial extraction.
//
// Test infrastructure only: compiled against the reference's own headers where
// they lie (oracle/Makefile, target `ref`) into oracle/_ref/ref_levelset and
// run here to write fixtures under tests/golden/.  It drives
// MarchingSimplices::RegularSubGridTriangulation (Misha/MarchingSimplices.h:
// 622-706) and MarchingSimplices::LevelSet (:410-576) exactly as
// Extract/Extract.cpp:161-208 does: triangulate a list of cells, sample two
// values per mesh vertex from a node grid at the rounded vertex position, and
// extract the codimension-2 zero set; it also extracts each value's
// codimension-1 zero set (Extract.cpp:303-309).
//
// Input (binary): u32 dim, u32 res, u64 nCells, nCells*dim u32 cell offsets,
// (res+1)^dim * 2 f64 node values (node index with coordinate 0 most
// significant, two values per node).
// Output: records {u32 nameLen, name, u64 count, count*f64} (as ref_dump).

#include <stdio.h>
#include <stdlib.h>
#include <vector>
#include <string>
#include <cmath>

#include "Misha/Miscellany.h"
#include "Misha/Geometry.h"
#include "Misha/MarchingSimplices.h"

using namespace MishaK;

static FILE *dumpFile = nullptr;
static void Put( const std::string &name , const std::vector< double > &v )
{
	unsigned int l = (unsigned int)name.size();
	fwrite( &l , sizeof(l) , 1 , dumpFile );
	fwrite( name.data() , 1 , l , dumpFile );
	unsigned long long c = v.size();
	fwrite( &c , sizeof(c) , 1 , dumpFile );
	if( c ) fwrite( v.data() , sizeof(double) , c , dumpFile );
}

template< unsigned int Dim , typename Mesh >
static void PutMesh( const std::string &prefix , const Mesh &mesh , unsigned int K )
{
	std::vector< double > v , s;
	for( size_t i=0 ; i<mesh.vertices.size() ; i++ ) for( unsigned int d=0 ; d<Dim ; d++ ) v.push_back( mesh.vertices[i][d] );
	for( size_t i=0 ; i<mesh.simplexIndices.size() ; i++ ) for( unsigned int k=0 ; k<=K ; k++ ) s.push_back( (double)mesh.simplexIndices[i][k] );
	Put( prefix + "_vertices" , v );
	Put( prefix + "_simplices" , s );
}

template< unsigned int Dim >
static int Run( FILE *in , unsigned int res )
{
	unsigned long long nCells;
	if( fread( &nCells , sizeof(nCells) , 1 , in )!=1 ) return 1;
	std::vector< unsigned int > cells( nCells*Dim );
	if( nCells && fread( cells.data() , sizeof(unsigned int) , cells.size() , in )!=cells.size() ) return 1;
	size_t nNodes = 1;
	for( unsigned int d=0 ; d<Dim ; d++ ) nNodes *= res+1;
	std::vector< double > nodeValues( nNodes*2 );
	if( fread( nodeValues.data() , sizeof(double) , nodeValues.size() , in )!=nodeValues.size() ) return 1;

	// the canonical cube's simplices (positions), as RegularSubGridTriangulation takes them
	{
		HyperCubeSimplices< Dim > hcs;
		std::vector< double > s;
		for( unsigned int i=0 ; i<HyperCubeSimplices< Dim >::SimplexNum ; i++ )
		{
			Simplex< unsigned int , Dim , Dim > si = hcs.simplex( i );
			for( unsigned int j=0 ; j<=Dim ; j++ ) for( unsigned int d=0 ; d<Dim ; d++ ) s.push_back( (double)si[j][d] );
		}
		Put( "hc_simplices" , s );
	}

	auto CellIndex = [&]( size_t e )
		{
			Point< unsigned int , Dim > I;
			for( unsigned int d=0 ; d<Dim ; d++ ) I[d] = cells[ e*Dim+d ];
			return I;
		};
	MarchingSimplices::SimplicialMesh< Dim , unsigned int , Point< double , Dim > > sMesh = MarchingSimplices::RegularSubGridTriangulation< Dim >( (size_t)nCells , CellIndex , true , false );
	PutMesh< Dim >( "tri" , sMesh , Dim );

	// values at the mesh vertices: the node at the rounded position (Extract.cpp:193-201)
	std::vector< Point< double , 2 > > values( sMesh.vertices.size() );
	for( size_t v=0 ; v<sMesh.vertices.size() ; v++ )
	{
		size_t idx = 0;
		for( unsigned int d=0 ; d<Dim ; d++ ) idx = idx*(res+1) + (size_t)(int)floor( sMesh.vertices[v][d]+0.5 );
		values[v][0] = nodeValues[2*idx] , values[v][1] = nodeValues[2*idx+1];
	}
	MarchingSimplices::SimplicialMesh< Dim-2 , unsigned int , Point< double , Dim > > levelSet = MarchingSimplices::LevelSet< double >( sMesh , [&]( unsigned int idx ){ return values[idx]; } , Point< double , 2 >() );
	PutMesh< Dim >( "ls" , levelSet , Dim-2 );

	for( unsigned int c=0 ; c<2 ; c++ )
	{
		MarchingSimplices::SimplicialMesh< Dim-1 , unsigned int , Point< double , Dim > > ls1 = MarchingSimplices::LevelSet< double >( sMesh , [&]( unsigned int idx ){ return Point< double , 1 >( values[idx][c] ); } , Point< double , 1 >() );
		PutMesh< Dim >( "ls1_" + std::to_string( c ) , ls1 , Dim-1 );
	}
	return 0;
}

int main( int argc , char *argv[] )
{
	if( argc!=3 ){ fprintf( stderr , "usage: %s <input> <dump>\n" , argv[0] ) ; return 1; }
	FILE *in = fopen( argv[1] , "rb" );
	if( !in ){ fprintf( stderr , "cannot open %s\n" , argv[1] ) ; return 1; }
	dumpFile = fopen( argv[2] , "wb" );
	if( !dumpFile ){ fprintf( stderr , "cannot open %s\n" , argv[2] ) ; return 1; }
	unsigned int dim , res;
	if( fread( &dim , sizeof(dim) , 1 , in )!=1 || fread( &res , sizeof(res) , 1 , in )!=1 ) return 1;
	int rc = 1;
	if     ( dim==3 ) rc = Run< 3 >( in , res );
	else if( dim==4 ) rc = Run< 4 >( in , res );
	fclose( in );
	fclose( dumpFile );
	if( rc ) fprintf( stderr , "bad input\n" );
	return rc;
}
