// Golden-vector harness around the reference's multigrid solver.
//
// Test infrastructure only: compiled against the reference's Include/MG.h
// where it lies (oracle/Makefile, target `ref`) into oracle/_ref/ref_mg and
// run here to write fixtures under tests/golden/.  It builds
// MGSolver( M , Ps , pNum ) (Include/MG.h:83-108) and runs
// vCycle( b , x , iters , MCI ) (:146-186) with the given colour lists,
// single-threaded, dumping x after every cycle; it also dumps one
// GSRelaxer sweep of M (:62-79) and the Galerkin matrices applied to a probe.
//
// Input (binary): u32 n, u32 pNum, u32 iters, u32 cycles, the matrix M as CSC
// (u32 nnz, (n+1) u32 colptr, nnz u32 rowidx, nnz f64), then per prolongation
// i = 0..pNum-1: u32 rows, u32 cols, CSC as above; then b (n f64), then per
// level l = 1..pNum: u32 size, u32 nColours, (nColours+1) u32 starts, size u32
// indices.
// Output: records {u32 nameLen, name, u64 count, count*f64}.

#include <stdio.h>
#include <stdlib.h>
#include <vector>
#include <string>
#include <Eigen/Sparse>

#include "Misha/MultiThreading.h"
#include "Include/MG.h"

using namespace MishaK;

static FILE *dumpFile = nullptr;
static void Put( const std::string &name , const double *v , size_t n )
{
	unsigned int l = (unsigned int)name.size();
	fwrite( &l , sizeof(l) , 1 , dumpFile );
	fwrite( name.data() , 1 , l , dumpFile );
	unsigned long long c = n;
	fwrite( &c , sizeof(c) , 1 , dumpFile );
	if( n ) fwrite( v , sizeof(double) , n , dumpFile );
}

static unsigned U32( FILE *f ){ unsigned v = 0 ; if( fread( &v , sizeof(v) , 1 , f )!=1 ){ fprintf( stderr , "short input\n" ) ; exit( 1 ); } return v; }
static void F64( FILE *f , double *p , size_t n ){ if( n && fread( p , sizeof(double) , n , f )!=n ){ fprintf( stderr , "short input\n" ) ; exit( 1 ); } }

static Eigen::SparseMatrix< double > ReadCSC( FILE *f , unsigned rows , unsigned cols )
{
	unsigned nnz = U32( f );
	std::vector< unsigned > cp( cols+1 ) , ri( nnz );
	std::vector< double > v( nnz );
	for( unsigned i=0 ; i<=cols ; i++ ) cp[i] = U32( f );
	for( unsigned i=0 ; i<nnz ; i++ ) ri[i] = U32( f );
	F64( f , v.data() , nnz );
	std::vector< Eigen::Triplet< double > > t;
	for( unsigned c=0 ; c<cols ; c++ ) for( unsigned k=cp[c] ; k<cp[c+1] ; k++ ) t.emplace_back( ri[k] , c , v[k] );
	Eigen::SparseMatrix< double > M( rows , cols );
	M.setFromTriplets( t.begin() , t.end() );
	return M;
}

int main( int argc , char *argv[] )
{
	if( argc!=3 ){ fprintf( stderr , "usage: %s <input> <dump>\n" , argv[0] ) ; return 1; }
	ThreadPool::ParallelizationType = ThreadPool::ParallelType::NONE;
	FILE *in = fopen( argv[1] , "rb" );
	if( !in ) return 1;
	dumpFile = fopen( argv[2] , "wb" );
	if( !dumpFile ) return 1;
	const unsigned n = U32( in ) , pNum = U32( in ) , iters = U32( in ) , cycles = U32( in );
	Eigen::SparseMatrix< double > M = ReadCSC( in , n , n );
	std::vector< Eigen::SparseMatrix< double > > Ps( pNum );
	for( unsigned i=0 ; i<pNum ; i++ ){ const unsigned r = U32( in ) , c = U32( in ) ; Ps[i] = ReadCSC( in , r , c ); }
	Eigen::VectorXd b( n );
	F64( in , b.data() , n );
	std::vector< std::vector< std::vector< unsigned int > > > mc( pNum+1 );
	for( unsigned l=1 ; l<=pNum ; l++ )
	{
		const unsigned size = U32( in ) , nc = U32( in );
		std::vector< unsigned > st( nc+1 ) , idx( size );
		for( unsigned c=0 ; c<=nc ; c++ ) st[c] = U32( in );
		for( unsigned i=0 ; i<size ; i++ ) idx[i] = U32( in );
		mc[l].resize( nc );
		for( unsigned c=0 ; c<nc ; c++ ) mc[l][c].assign( idx.begin()+st[c] , idx.begin()+st[c+1] );
	}
	fclose( in );

	// one colour-ordered relaxation of M from x = 0
	{
		GSRelaxer relax( M );
		Eigen::VectorXd x = Eigen::VectorXd::Zero( n );
		relax( b , x , mc[pNum] );
		Put( "gs_x" , x.data() , n );
	}
	MGSolver mg( M , [&]( unsigned int i ){ return Ps[i]; } , pNum );
	Eigen::VectorXd x = Eigen::VectorXd::Zero( n );
	for( unsigned c=0 ; c<cycles ; c++ )
	{
		mg.vCycle( b , x , iters , [&]( unsigned int l ) -> const std::vector< std::vector< unsigned int > > & { return mc[l]; } );
		Put( "x" + std::to_string( c ) , x.data() , n );
		Eigen::VectorXd r = b - M*x;
		double rn = r.norm();
		Put( "res" + std::to_string( c ) , &rn , 1 );
	}
	fclose( dumpFile );
	return 0;
}
