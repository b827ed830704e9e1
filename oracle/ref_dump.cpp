// Golden-vector harness around the reference implementation.
//
// Test infrastructure only: this file is compiled against the reference's own
// headers where they lie (see oracle/Makefile) into oracle/_ref/ref_dump and is
// run here to write the fixtures under tests/golden/.  It drives the reference
// API exactly as ExteriorPoissonRecon/ExteriorPoissonRecon.cpp:201-279 does for
// the dense indexer (--aRadius -1), with threading disabled so that every sum
// is taken in one fixed order, and dumps the intermediate quantities of the
// solve path (estimator outputs, splatted field, matrices applied to a probe
// vector, per-node quadratic fits, one Gauss-Seidel sweep, the cascadic
// trajectory) as named float64 arrays.
//
// Dump format: repeated records {u32 nameLen, name, u64 count, count*f64}.

#include <stdio.h>
#include <stdlib.h>
#include <vector>
#include <string>
#include <cstring>
#include <cmath>

#include "Misha/Miscellany.h"
#include "Misha/CmdLineParser.h"
#include "Misha/Geometry.h"
#include "Misha/MarchingSimplices.h"
#include "Misha/Ply.h"
#include "Misha/PlyVertexData.h"
#include "Misha/MultiThreading.h"
#include "Misha/Streams.h"
#include "Misha/EigenStreams.h"
#include "Include/Samples.h"
#include "ExteriorPoissonRecon/ExteriorPoisson.h"

using namespace MishaK;

static const unsigned int CoDim = 2;

CmdLineParameter< std::string > In( "in" ) , Dump( "dump" ) , TreeOut( "out" );
CmdLineParameter< double > NoiseToWeightSigma( "nSigma" , 0.25 ) , Scale( "scale" , 1.1 ) , ScreeningWeight( "sWeight" , 100. ) , DirichletWeight( "dWeight" , 0.00625 ) , IterationMultiplier( "iMult" , 2. );
CmdLineParameter< unsigned int > Depth( "depth" , 5 ) , Seed( "seed" , 0 ) , TensorKernelRadius( "tRadius" , 1 ) , DensityKernelRadius( "dRadius" , 1 ) , MinSolveDepth( "minSolveDepth" , 0 ) , GSIterations( "gsIters" , 10 ) , FitLevels( "fitLevels" , 3 );
CmdLineReadable Threads( "threads" );
CmdLineReadable* params[] = { &In , &Dump , &TreeOut , &NoiseToWeightSigma , &Scale , &ScreeningWeight , &DirichletWeight , &IterationMultiplier , &Depth , &Seed , &TensorKernelRadius , &DensityKernelRadius , &MinSolveDepth , &GSIterations , &FitLevels , &Threads , NULL };

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
static void Put( const std::string &name , const std::vector< double > &v ){ Put( name , v.data() , v.size() ); }
static void Put( const std::string &name , const Eigen::VectorXd &v ){ Put( name , v.data() , (size_t)v.size() ); }
static void Put( const std::string &name , double s ){ Put( name , &s , 1 ); }

// Deterministic probe vector, reproduced by the tests: v[i] = cos(0.7*i + salt)
static Eigen::VectorXd Probe( size_t n , double salt )
{
	Eigen::VectorXd v( n );
	for( size_t i=0 ; i<n ; i++ ) v[i] = cos( 0.7*(double)i + salt );
	return v;
}

template< unsigned int Dim , typename Indexer >
struct DumpEnergy : public SingleCycleCascadicSystemEnergy< Dim , Indexer >
{
	using Base = SingleCycleCascadicSystemEnergy< Dim , Indexer >;
	using Base::Base;
	DumpEnergy( Base &&b ) : Base( std::move(b) ){}
	DumpEnergy( const Base &b ) : Base( b ){}
	const Eigen::SparseMatrix< double > &R( void ) const { return this->_R; }
	const Eigen::SparseMatrix< double > &B( void ) const { return this->_B; }
	double c( void ) const { return this->_c; }
};

template< unsigned int Dim , typename Energy >
struct DumpRecon : public ExteriorPoisson::Reconstructor< Dim , Energy >
{
	using Base = ExteriorPoisson::Reconstructor< Dim , Energy >;
	using Base::Base;
	const std::vector< Energy > &energies( void ) const { return this->_energies; }
	const std::vector< std::vector< std::vector< unsigned int > > > &mcIndices( void ) const { return this->_mcIndices; }
};

template< unsigned int Dim >
std::vector< std::pair< Point< double , Dim > , Hat::SkewSymmetricMatrix< double , Dim > > > ReadSamples( std::string fileName )
{
	std::vector< std::pair< Point< double , Dim > , Hat::SkewSymmetricMatrix< double , Dim > > > hermiteData;
	using Factory = VertexFactory::Factory< double, VertexFactory::PositionFactory< double, Dim >, VertexFactory::MatrixFactory< double , Dim , Dim > >;
	Factory factory = Factory( VertexFactory::PositionFactory< double , Dim >() , VertexFactory::MatrixFactory< double , Dim , Dim >( "skew" ) );
	std::vector< typename Factory::VertexType > vertices;
	std::vector< SimplexIndex< Dim-CoDim, unsigned int > > simplexIndices;
	PLY::ReadSimplices( fileName , factory , vertices , simplexIndices , NULL );
	hermiteData.resize( vertices.size() );
	for( size_t i=0 ; i<vertices.size() ; i++ )
	{
		hermiteData[i].first = vertices[i].template get<0>();
		hermiteData[i].second = Hat::SkewSymmetricMatrix< double , Dim >( vertices[i].template get<1>() );
	}
	return hermiteData;
}

template< unsigned int Dim >
bool ValidSampleFile( std::string fileName )
{
	VertexFactory::PositionFactory< double , Dim > pFactory;
	int fileType;
	std::vector< GregTurk::PlyProperty > plyProperties = PLY::ReadVertexHeader( fileName , fileType );
	return pFactory.plyValidReadProperties( plyProperties );
}

template< unsigned int Dim >
void Execute( std::vector< std::pair< Point< double , Dim > , Hat::SkewSymmetricMatrix< double , Dim > > > hermiteData )
{
	using namespace Hat;
	using HierarchicalIndexer = HierarchicalRegularIndexer< Dim >;
	using Indexer = typename HierarchicalIndexer::Indexer;
	using Energy = DumpEnergy< Dim , Indexer >;
	static const unsigned int CD = ( Dim*(Dim-1) )/2;

	::SquareMatrix< double , Dim+1 > unitCubeToWorld;
	{
		::SquareMatrix< double , Dim+1 > worldToUnitCube = ExteriorPoisson::ToUnitCube< Dim >( [&]( size_t idx ){ return hermiteData[idx].first; } , hermiteData.size() , Scale.value );
		for( unsigned int i=0 ; i<hermiteData.size() ; i++ ) hermiteData[i].first = worldToUnitCube( hermiteData[i].first );
		unitCubeToWorld = worldToUnitCube.inverse();
	}
	{
		std::vector< double > pos( hermiteData.size()*Dim ) , skew( hermiteData.size()*CD ) , xf( (Dim+1)*(Dim+1) );
		for( size_t i=0 ; i<hermiteData.size() ; i++ )
		{
			for( unsigned int d=0 ; d<Dim ; d++ ) pos[i*Dim+d] = hermiteData[i].first[d];
			for( unsigned int c=0 ; c<CD ; c++ ) skew[i*CD+c] = hermiteData[i].second[c];
		}
		for( unsigned int c=0 ; c<=Dim ; c++ ) for( unsigned int r=0 ; r<=Dim ; r++ ) xf[c*(Dim+1)+r] = unitCubeToWorld(c,r);
		Put( "unit_pos" , pos );
		Put( "skew_packed" , skew );
		Put( "unit_to_world" , xf );
	}

	unsigned int depth = Depth.value;
	unsigned int minSolveDepth = std::min< unsigned int >( MinSolveDepth.value , depth ) , maxSolveDepth = depth;
	HierarchicalIndexer hierarchicalIndexer( depth );

	OrderedSampler< Dim > orderedSampler( [&]( size_t idx ){ return hermiteData[idx].first; } , hermiteData.size() , 1<<depth );
	{
		std::vector< double > cells , counts;
		for( size_t i=0 ; i<orderedSampler.size() ; i++ )
		{
			for( unsigned int d=0 ; d<Dim ; d++ ) cells.push_back( orderedSampler[i].first[d] );
			counts.push_back( (double)orderedSampler[i].second.size() );
		}
		Put( "cells" , cells );
		Put( "cell_counts" , counts );
	}

	GridSamples::TreeEstimator< Dim , 2 > estimator( DensityKernelRadius.value , depth , [&]( size_t idx ){ return hermiteData[idx].first; } , orderedSampler , true );
	{
		std::vector< double > m( hermiteData.size() ) , n( hermiteData.size() ) , dd( hermiteData.size() );
		for( size_t i=0 ; i<hermiteData.size() ; i++ )
		{
			m[i] = estimator.measure( hermiteData[i].first , 0 );
			n[i] = estimator.noise( hermiteData[i].first , 0 );
			dd[i] = estimator.depth( hermiteData[i].first , 0 );
		}
		Put( "measure" , m );
		Put( "noise" , n );
		Put( "sample_depth" , dd );
	}

	// The skew splat, restated with the same reference calls as
	// ExteriorPoisson.h:175-264, to expose the per-element field S(E).
	{
		GridSamples::TreeSplatter< Dim , SkewSymmetricMatrix< double , Dim > > treeSplatter( depth );
		std::vector< std::tuple< Point< double , Dim > , SkewSymmetricMatrix< double , Dim > , double > > samples( hermiteData.size() );
		for( size_t i=0 ; i<orderedSampler.size() ; i++ ) for( size_t j : orderedSampler[i].second )
		{
			std::pair< Point< double , Dim > , SkewSymmetricMatrix< double , Dim > > s = hermiteData[j];
			s.second /= sqrt( s.second.squareNorm() );
			s.second *= estimator.measure( s.first , 0 );
			s.second *= 1<<(depth*CoDim);
			samples[j] = std::make_tuple( s.first , s.second , estimator.depth( s.first , 0 ) );
		}
		auto DepthWeight = []( unsigned int d ){ return (double)( 1<<(d*Dim) ); };
		auto __F = [&]( size_t idx ){ return std::pair< Point< double , Dim > , SkewSymmetricMatrix< double , Dim > >( std::get< 0 >( samples[idx] ) , std::get< 1 >( samples[idx] ) ); };
		auto __D = [&]( size_t idx ){ return std::get< 2 >( samples[idx] ); };
		switch( TensorKernelRadius.value )
		{
		case 0: treeSplatter.template addSamples< 0 >( __F , orderedSampler , __D , DepthWeight , false ) ; break;
		case 1: treeSplatter.template addSamples< 1 >( __F , orderedSampler , __D , DepthWeight , false ) ; break;
		case 2: treeSplatter.template addSamples< 2 >( __F , orderedSampler , __D , DepthWeight , false ) ; break;
		case 3: treeSplatter.template addSamples< 3 >( __F , orderedSampler , __D , DepthWeight , false ) ; break;
		default: treeSplatter.template addSamples< 4 >( __F , orderedSampler , __D , DepthWeight , false ) ; break;
		}
		double scale = pow( pow( 0.5 , depth ) , 1.5 * CoDim );
		treeSplatter.process( [&]( unsigned int , Hat::Index< Dim > , SkewSymmetricMatrix< double , Dim > &s ){ s *= scale; } );
		unsigned int res = 1<<depth;
		ScalarFunctions< Dim > scalars( res );
		std::vector< double > S( scalars.elementNum()*CD );
		for( size_t e=0 ; e<scalars.elementNum() ; e++ )
		{
			Hat::Index< Dim > I = scalars.elementIndex( e );
			Point< double , Dim > p;
			for( unsigned int d=0 ; d<Dim ; d++ ) p[d] = ( I[d] + 0.5 ) / res;
			SkewSymmetricMatrix< double , Dim > s = treeSplatter( p , 0 );
			for( unsigned int c=0 ; c<CD ; c++ ) S[e*CD+c] = s[c];
		}
		Put( "S_elem" , S );
	}

	srand( Seed.value );
	auto NoiseToWeight = []( double n ){ return exp( -n*n/(2.*NoiseToWeightSigma.value*NoiseToWeightSigma.value) ); };
	DumpRecon< Dim , Energy > recon
	(
		hierarchicalIndexer ,
		depth , minSolveDepth , maxSolveDepth ,
		[&]( size_t idx ){ return hermiteData[idx]; } , orderedSampler ,
		estimator ,
		NoiseToWeight ,
		DirichletWeight.value , ScreeningWeight.value ,
		TensorKernelRadius.value ,
		0 , 1e-10 ,
		false
	);
	const std::vector< Energy > &energies = recon.energies();
	Put( "c" , energies.back().c() );

	// Per-level matrices applied to probe vectors, energies at probe (x,y),
	// per-node quadratic fits and one paired Gauss-Seidel iteration.
	for( unsigned int d=0 ; d<=depth ; d++ )
	{
		size_t N = hierarchicalIndexer.numFunctions( d );
		Eigen::VectorXd v = Probe( N , 0.25 + d );
		Put( "R" + std::to_string(d) + "_v" , ( energies[d].R() * v ).eval() );
		Put( "B" + std::to_string(d) + "_v" , ( energies[d].B() * v ).eval() );
		Put( "R" + std::to_string(d) + "_nnz" , (double)energies[d].R().nonZeros() );
		Eigen::VectorXd x = Probe( N , 1.5 + d ) * 0.01 , y = Probe( N , 2.75 + 0.5*d ) * 0.01;
		std::pair< double , double > e = energies[d].energies( x , y );
		double E[2] = { e.first , e.second };
		Put( "E" + std::to_string(d) , E , 2 );
		Put( "W" + std::to_string(d) , energies[d].wedgeSquareNorm( x , y ) );
		if( d+FitLevels.value>depth || d<2 )
		{
			std::vector< double > Q( N*5 );
			for( size_t i=0 ; i<N ; i++ )
			{
				Polynomial::Polynomial2D< 2 > q = energies[d].quadraticFit( x , y , (unsigned int)i , false , 0 );
				Q[5*i+0] = q.coefficient(2,0) , Q[5*i+1] = q.coefficient(0,2) , Q[5*i+2] = q.coefficient(1,1) , Q[5*i+3] = q.coefficient(1,0) , Q[5*i+4] = q.coefficient(0,1);
			}
			Put( "Q" + std::to_string(d) , Q );
			Eigen::VectorXd gx = x , gy = y;
			ExteriorPoissonSolvers::PairedGaussSeidelSolve< Dim >( d , recon.mcIndices()[d] , energies[d] , gx , gy , 1 , false );
			Put( "gs" + std::to_string(d) + "_x" , gx );
			Put( "gs" + std::to_string(d) + "_y" , gy );
		}
		if( d<depth )
		{
			Put( "P" + std::to_string(d) + "_v" , energies[d].prolongation( v ) );
		}
	}

	// The initial guess, restated from ExteriorPoisson.h:333-341.
	{
		srand( Seed.value );
		size_t N = hierarchicalIndexer.numFunctions( minSolveDepth );
		Eigen::VectorXd x( N ) , y( N );
		for( size_t i=0 ; i<N ; i++ ) x[i] = Random< double >() , y[i] = Random< double >();
		double W0 = energies[minSolveDepth].wedgeSquareNorm( x , y );
		double scale = sqrt( sqrt( energies.back().squareNorm() / W0 ) );
		Put( "init_W" , W0 );
		Put( "init_x" , ( x*scale ).eval() );
		Put( "init_y" , ( y*scale ).eval() );
	}

	// The full cascadic solve (ExteriorPoisson.h:316-409).
	srand( Seed.value );
	recon.solve( false , IterationMultiplier.value , GSIterations.value , false , false );
	{
		Eigen::VectorXd xf = recon.x( 0 ) , yf = recon.y( 0 );
		for( unsigned int d=0 ; d<=depth ; d++ )
		{
			Put( "hx" + std::to_string(d) , recon.x( d ) );
			Put( "hy" + std::to_string(d) , recon.y( d ) );
			if( d ) xf = recon.x( d ) + energies[d-1].prolongation( xf ) , yf = recon.y( d ) + energies[d-1].prolongation( yf );
		}
		Put( "final_x" , xf );
		Put( "final_y" , yf );
		std::pair< double , double > e = energies[depth].energies( xf , yf );
		double E[2] = { e.first , e.second };
		Put( "final_E" , E , 2 );
	}

	if( TreeOut.set )
	{
		XForm< double , Dim+1 > voxelToWorld;
		{
			::SquareMatrix< double , Dim+1 > gridToCube = ::SquareMatrix< double , Dim+1 >::Identity();
			for( unsigned int d=0 ; d<Dim ; d++ ) gridToCube(d,d) = 1./(1<<depth);
			voxelToWorld = unitCubeToWorld * gridToCube;
		}
		FileStream fStream( TreeOut.value + std::string( ".tree" ) , false , true );
		unsigned int dim = Dim;
		size_t hierarchicalIndexerHashCode = typeid( hierarchicalIndexer ).hash_code();
		fStream.write( dim );
		fStream.write( hierarchicalIndexerHashCode );
		fStream.write( voxelToWorld );
		fStream.write( recon.x() );
		fStream.write( recon.y() );
		hierarchicalIndexer.write( fStream , false );
		Put( "tree_hash" , (double)(hierarchicalIndexerHashCode & 0xFFFFFFFFFFFFull) );
	}
}

int main( int argc , char* argv[] )
{
	CmdLineParse( argc-1 , argv+1 , params );
	if( !In.set || !Dump.set ){ fprintf( stderr , "usage: %s --in <ply> --dump <file> [--depth D] [--out <tree header>]\n" , argv[0] ); return 1; }
	if( !Threads.set ) ThreadPool::ParallelizationType = ThreadPool::ParallelType::NONE;
	dumpFile = fopen( Dump.value.c_str() , "wb" );
	if( !dumpFile ){ fprintf( stderr , "cannot open %s\n" , Dump.value.c_str() ); return 1; }
	if     ( ValidSampleFile< 4 >( In.value ) ) Execute< 4 >( ReadSamples< 4 >( In.value ) );
	else if( ValidSampleFile< 3 >( In.value ) ) Execute< 3 >( ReadSamples< 3 >( In.value ) );
	else { fprintf( stderr , "invalid sample file\n" ); return 1; }
	fclose( dumpFile );
	return 0;
}
