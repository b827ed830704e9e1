// CPU restatement (test infrastructure, the oracle) of the reference's Extract
// pipeline up to the PLY write, serial and in the reference's own terms:
//
//   .tree reading                       Extract.cpp:109-120, Hat.adaptive.inl:198-203, 405-414,
//                                       RegularTree.inl:616-662 (adapted tree dump)
//   cumulative (regular) values         IsoTree.inl:314-338 (_setElementCumulative)
//   zero crossings                      IsoTree.inl:135-163 (isZeroCrossingNode)
//   refinement (flood fill)             IsoTree.inl:165-222 (refineZeroCrossing),
//                                       :237-265 (addElementNeighbors: interior, face adjacent)
//   level-set nodes in tree order       IsoTree.inl:225-234, RegularTree.inl:396-414 (processNodes)
//   triangulation, vertex values        Extract.cpp:161-199 (orc_subgrid_triangulation)
//   joint zero set                      Extract.cpp:205 (orc_level_set)
//   orientation                         MarchingSimplices.h:179-246 (Orient), :84-146 (faces)
//   gradient frames                     IsoTree.inl:101-133 (gradient), RegularTree.inl:248-262
//                                       (getLeafNode), Extract.cpp:240-257
//   voxel-to-world                      Misha/Geometry.inl:413-427
//   per-function level sets             Extract.cpp:296-310
//
// Only tests/ and the CPU checks use this file; the product never links it.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <bits/hash_bytes.h>
#include <functional>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "epr_oracle.h"

namespace
{
std::string g_err;

struct Tree
{
	int dim = 0 , D = 0;
	bool adapted = false;
	std::vector< double > xf;                                  // coords[c][r], c-major
	std::vector< std::vector< double > > x , y;
	std::vector< std::vector< std::array< int , 4 > > > fIdx;  // adapted: function multi-indices per depth
	std::set< std::array< int , 4 > > indexedFinest;           // adapted: indexed elements of the finest depth
};

// typeid( Hat::Hierarchical{Regular,Adapted}Indexer< dim > ).hash_code() under libstdc++
unsigned long long IndexerHash( int dim , bool adapted )
{
	const std::string name = std::string( "N6MishaK3Hat26Hierarchical" ) + ( adapted ? "Adapted" : "Regular" ) + "IndexerILj" + std::to_string( dim ) + "EEE";
	return std::_Hash_bytes( name.data() , name.size() , 0xc70f6907UL );
}

bool ReadTree( const char *path , Tree &t )
{
	FILE *fp = fopen( path , "rb" );
	if( !fp ){ g_err = "cannot open tree" ; return false; }
	bool ok = true;
	auto Read = [&]( void *p , size_t n ){ if( ok && fread( p , 1 , n , fp )!=n ) ok = false; };
	auto U64 = [&]( void ){ unsigned long long v = 0 ; Read( &v , 8 ) ; return v; };
	unsigned dim = 0;
	Read( &dim , 4 );
	t.dim = (int)dim;
	const unsigned long long hash = U64();
	if( dim<3 || dim>4 ){ fclose( fp ) ; g_err = "bad dimension" ; return false; }
	if( hash==IndexerHash( (int)dim , false ) ) t.adapted = false;
	else if( hash==IndexerHash( (int)dim , true ) ) t.adapted = true;
	else { fclose( fp ) ; g_err = "unrecognized hash" ; return false; }
	(void)U64();
	t.xf.resize( ( dim+1 )*( dim+1 ) );
	Read( t.xf.data() , t.xf.size()*8 );
	for( int c=0 ; c<2 ; c++ )
	{
		std::vector< std::vector< double > > &v = c ? t.y : t.x;
		v.resize( U64() );
		for( auto &l : v ){ l.resize( U64() ) ; Read( l.data() , l.size()*8 ); }
	}
	t.D = (int)t.x.size() - 1;
	if( !t.adapted )
	{
		unsigned md = 0;
		Read( &md , 4 );
		fclose( fp );
		if( !ok ) g_err = "truncated tree";
		return ok;
	}
	t.fIdx.resize( t.D+1 );
	for( int d=0 ; d<=t.D ; d++ ) t.fIdx[d].resize( t.x[d].size() );
	const size_t head = ( ( 2*( dim+1 ) + 7 )/8 )*8;
	std::vector< unsigned char > rec( head + 32 );
	// RegularTreeNode::write / ReadChildren: a node with children is followed
	// by its brood, then each child's subtree
	auto Node = [&]( void )
	{
		Read( rec.data() , rec.size() );
		unsigned short h[5] = { 0 , 0 , 0 , 0 , 0 };
		memcpy( h , rec.data() , 2*( dim+1 ) );
		unsigned long long q[4];
		memcpy( q , rec.data() + head , 32 );
		std::array< int , 4 > off = { 0 , 0 , 0 , 0 };
		for( unsigned k=0 ; k<dim ; k++ ) off[k] = h[1+k];
		const int d = (int)h[0] - 1;
		if( q[2]!=~0ull && d>=0 && d<=t.D && q[2]<t.fIdx[d].size() ) t.fIdx[d][ q[2] ] = off;
		if( q[3]!=~0ull && d==t.D ) t.indexedFinest.insert( off );
		return q[1]!=0;
	};
	if( Node() )
	{
		// iterative form of ReadChildren( root )
		struct Frame { std::vector< bool > has ; int next = 0; };
		std::vector< Frame > st;
		auto Brood = [&]( void ){ Frame f ; f.has.resize( 1u<<dim ) ; for( unsigned c=0 ; c<( 1u<<dim ) ; c++ ) f.has[c] = Node() ; return f; };
		st.push_back( Brood() );
		while( !st.empty() && ok )
		{
			Frame &f = st.back();
			if( f.next==(int)f.has.size() ){ st.pop_back() ; continue; }
			const bool h = f.has[ f.next++ ];
			if( h ) st.push_back( Brood() );
		}
	}
	// the active sizes and node tables that follow are not needed here
	fclose( fp );
	if( !ok ) g_err = "truncated tree";
	return ok;
}

using Idx = std::array< int , 4 >;

struct Iso
{
	int dim , D , res;
	std::vector< double > V[2];                                // finest cumulative values, function index order
	bool adapted;
	const std::set< Idx > *indexed;

	long long F( const Idx &I ) const { long long f = 0 ; for( int d=0 ; d<dim ; d++ ) f = f*( res+1 ) + I[d] ; return f; }
	void values( const Idx &E , double v[16][2] ) const        // element corner values in window order
	{
		for( int i=0 ; i<( 1<<dim ) ; i++ )
		{
			Idx I = E;
			for( int d=0 ; d<dim ; d++ ) I[d] += ( i>>( dim-1-d ) ) & 1;
			v[i][0] = V[0][ F( I ) ] , v[i][1] = V[1][ F( I ) ];
		}
	}
	bool zeroCrossing( const Idx &E , bool all ) const
	{
		double v[16][2];
		values( E , v );
		unsigned char flags[2] = { 0 , 0 };
		for( int i=0 ; i<( 1<<dim ) ; i++ ) for( int c=0 ; c<2 ; c++ )
			if     ( v[i][c]<0 ) flags[c] |= 1;
			else if( v[i][c]>0 ) flags[c] |= 2;
			else                 flags[c] |= 3;
		return all ? ( flags[0]==3 && flags[1]==3 ) : ( flags[0]==3 || flags[1]==3 );
	}
};

// IsoTree::_setElementCumulative's recursion, level by level over every function
void Cumulative( const Tree &t , Iso &iso )
{
	const int dim = t.dim;
	std::vector< double > prev[2];
	for( int d=0 ; d<=t.D ; d++ )
	{
		const int r = 1<<d;
		long long n = 1 ; for( int k=0 ; k<dim ; k++ ) n *= r+1;
		std::vector< double > cur[2] = { std::vector< double >( n , 0. ) , std::vector< double >( n , 0. ) };
		for( size_t i=0 ; i<t.x[d].size() ; i++ )
		{
			long long f = (long long)i;
			if( t.adapted ){ f = 0 ; for( int k=0 ; k<dim ; k++ ) f = f*( r+1 ) + t.fIdx[d][i][k]; }
			cur[0][f] = t.x[d][i] , cur[1][f] = t.y[d][i];
		}
		if( d>0 ) for( long long f=0 ; f<n ; f++ )
		{
			Idx Fi{} ;
			{ long long q = f ; for( int k=dim-1 ; k>=0 ; k-- ){ Fi[k] = (int)( q % ( r+1 ) ) ; q /= r+1; } }
			// the parent element of an element with corner Fi; its 2^dim corners in window order
			Idx PE{};
			for( int k=0 ; k<dim ; k++ ) PE[k] = std::min( Fi[k] , r-1 ) / 2;
			double cum[2] = { 0 , 0 };
			for( int i=0 ; i<( 1<<dim ) ; i++ )
			{
				Idx P = PE;
				double w = 1.;
				for( int k=0 ; k<dim ; k++ )
				{
					P[k] += ( i>>( dim-1-k ) ) & 1;
					const int I2 = 2*P[k];
					if( I2==Fi[k] ) w *= 1.0;
					else if( I2==Fi[k]+1 || I2==Fi[k]-1 ) w *= 0.5;
					else w *= 0.0;
				}
				if( !w ) continue;
				long long p = 0 ; for( int k=0 ; k<dim ; k++ ) p = p*( r/2+1 ) + P[k];
				cum[0] += prev[0][p]*w , cum[1] += prev[1][p]*w;
			}
			cur[0][f] += cum[0] , cur[1][f] += cum[1];
		}
		prev[0].swap( cur[0] ) , prev[1].swap( cur[1] );
	}
	iso.V[0].swap( prev[0] ) , iso.V[1].swap( prev[1] );
}

// IsoTree::refineZeroCrossing + levelSetNodes: returns the level-set nodes in
// tree order and the valid finest elements
void LevelSetNodes( const Iso &iso , bool all , std::vector< Idx > &nodes , std::set< Idx > &valid )
{
	const int dim = iso.dim , res = iso.res;
	const unsigned char FINALIZED = 1 , ADDED = 2;
	std::map< Idx , unsigned char > state;
	valid.clear();
	std::vector< Idx > active;
	auto Each = [&]( auto f )
	{
		Idx E{};
		long long n = 1 ; for( int k=0 ; k<dim ; k++ ) n *= res;
		for( long long e=0 ; e<n ; e++ ){ long long q = e ; for( int k=dim-1 ; k>=0 ; k-- ){ E[k] = (int)( q % res ) ; q /= res; } f( E ); }
	};
	Each( [&]( const Idx &E ){ if( !iso.adapted || iso.indexed->count( E ) ) valid.insert( E ); } );
	for( const Idx &E : valid )
		if( iso.zeroCrossing( E , all ) ) active.push_back( E ) , state[E] |= ADDED;
		else state[E] |= FINALIZED;
	while( !active.empty() )
	{
		const Idx E = active.back();
		active.pop_back();
		if( state[E] & FINALIZED ) continue;
		for( int d=0 ; d<dim ; d++ ) for( int s=-1 ; s<=1 ; s+=2 )
		{
			Idx N = E;
			N[d] += s;
			if( N[d]<0 || N[d]>=res ) continue;
			valid.insert( N );
			unsigned char &st = state[N];
			if( !( st & ADDED ) && !( st & FINALIZED ) )
			{
				if( iso.zeroCrossing( N , all ) ) active.push_back( N ) , st |= ADDED;
				else st |= FINALIZED;
			}
		}
		state[E] |= FINALIZED;
	}
	nodes.clear();
	for( const Idx &E : valid ) if( iso.zeroCrossing( E , all ) ) nodes.push_back( E );
	// depth-first order: child index bit d = coordinate d, coarser bits first
	auto Key = [&]( const Idx &E )
	{
		unsigned long long k = 0;
		for( int b=iso.D-1 ; b>=0 ; b-- ) for( int d=dim-1 ; d>=0 ; d-- ) k = ( k<<1 ) | ( ( E[d]>>b ) & 1 );
		return k;
	};
	std::sort( nodes.begin() , nodes.end() , [&]( const Idx &a , const Idx &b ){ return Key( a )<Key( b ); } );
}

// MarchingSimplices::Orient over (K+1)-vertex simplices
void Orient( int K , std::vector< unsigned > &S )
{
	const size_t n = S.size()/( K+1 );
	auto Face = [&]( size_t s , int v , std::vector< unsigned > &key , bool &even )
	{
		const unsigned *si = &S[ s*( K+1 ) ];
		std::vector< unsigned > f;
		if( K==1 ){ f = { si[ v==0 ? 1 : 0 ] } ; even = v==0; }
		else
		{
			for( int d=0 ; d<K ; d++ ) f.push_back( si[ ( v+d+1 )%( K+1 ) ] );
			if( ( K&1 ) && ( v&1 ) ) std::swap( f[0] , f[1] );
			// bubble-sort parity
			std::vector< unsigned > g = f;
			int count = 0;
			for( bool again=true ; again ; )
			{
				again = false;
				for( int i=0 ; i+1<(int)g.size() ; i++ ) if( g[i]>g[i+1] ){ std::swap( g[i] , g[i+1] ) ; count++ ; again = true ; break; }
			}
			even = ( count&1 )==0;
		}
		key = f;
		std::sort( key.begin() , key.end() );
	};
	struct FD { long long s1 = -1 , s2 = -1 ; bool o1 = false , o2 = false; };
	std::map< std::vector< unsigned > , FD > fMap;
	for( size_t i=0 ; i<n ; i++ ) for( int v=0 ; v<=K ; v++ )
	{
		std::vector< unsigned > key ; bool e;
		Face( i , v , key , e );
		FD &fd = fMap[key];
		if( fd.s1==-1 ) fd.s1 = (long long)i , fd.o1 = e;
		else if( fd.s2==-1 ) fd.s2 = (long long)i , fd.o2 = e;
		else throw std::string( "Both faces occupied" );
	}
	std::vector< bool > oriented( n , false );
	for( size_t idx=0 ; idx<n ; idx++ )
	{
		if( oriented[idx] ) continue;
		std::vector< std::pair< size_t , bool > > queue;
		oriented[idx] = true;
		queue.push_back( { idx , true } );
		while( !queue.empty() )
		{
			std::pair< size_t , bool > s = queue.back();
			queue.pop_back();
			for( int v=0 ; v<=K ; v++ )
			{
				std::vector< unsigned > key ; bool e;
				Face( s.first , v , key , e );
				const FD &fd = fMap[key];
				const long long o = (long long)s.first==fd.s1 ? fd.s2 : fd.s1;
				if( o==-1 || oriented[o] ) continue;
				oriented[o] = true;
				if( s.second!=( fd.o1==fd.o2 ) ) queue.push_back( { (size_t)o , true } );
				else
				{
					std::swap( S[ o*( K+1 ) ] , S[ o*( K+1 )+1 ] );
					queue.push_back( { (size_t)o , false } );
				}
			}
		}
	}
}

void ToWorld( const Tree &t , const double *p , double *out )
{
	const int Dim1 = t.dim+1;
	auto C = [&]( int c , int r ){ return t.xf[ c*Dim1+r ]; };
	double scale = 1;
	for( int i=0 ; i<t.dim ; i++ )
	{
		out[i] = 0;
		for( int j=0 ; j<t.dim ; j++ ) out[i] += p[j]*C( j , i );
		out[i] += C( t.dim , i );
		scale += C( i , t.dim );
	}
	for( int i=0 ; i<t.dim ; i++ ) out[i] /= scale;
}

// IsoTree::gradient at p in [0,1]^dim (grid position / res)
bool Gradient( const Iso &iso , const std::set< Idx > &valid , const double *pIn , double grad[2][4] )
{
	const int dim = iso.dim;
	double p[4];
	for( int d=0 ; d<dim ; d++ ) p[d] = pIn[d];
	// getLeafNode from the space root: centre 0.5, width 1, strict '>' to the upper child
	Idx leaf{};
	{
		double center[4] , width = 1.;
		for( int d=0 ; d<dim ; d++ ) center[d] = 0.5;
		for( int l=0 ; l<iso.D ; l++ )
		{
			int c = 0;
			for( int d=0 ; d<dim ; d++ ) if( p[d]>center[d] ) c |= 1<<d;
			for( int d=0 ; d<dim ; d++ ) leaf[d] = 2*leaf[d] + ( ( c>>d ) & 1 );
			width /= 2;
			for( int d=0 ; d<dim ; d++ ) if( c&( 1<<d ) ) center[d] += width/2 ; else center[d] -= width/2;
		}
	}
	if( !valid.count( leaf ) ) return false;
	const unsigned res = 1u<<iso.D;
	Idx E{};
	for( int d=0 ; d<dim ; d++ ) E[d] = std::max< int >( 0 , std::min< int >( (int)floor( p[d]*res ) , (int)res-1 ) );
	double v[16][2];
	iso.values( leaf , v );
	for( int d=0 ; d<dim ; d++ ) p[d] = p[d]*res - E[d];
	for( int c=0 ; c<2 ; c++ ) for( int d=0 ; d<4 ; d++ ) grad[c][d] = 0;
	for( int i=0 ; i<( 1<<dim ) ; i++ )
	{
		int I[4];
		for( int d=0 ; d<dim ; d++ ) I[d] = ( i>>( dim-1-d ) ) & 1;
		for( int d=0 ; d<dim ; d++ )
		{
			double weight = res;
			for( int dd=0 ; dd<dim ; dd++ )
			{
				if( dd==d ) weight *= ( I[dd]==1 ) ? 1. : -1.;
				else        weight *= ( I[dd]==1 ) ? p[dd] : 1.-p[dd];
			}
			for( int c=0 ; c<2 ; c++ ) grad[c][d] += v[i][c] * weight;
		}
	}
	return true;
}

// ---- density trimming (Extract.cpp:153-159, 217-238) ----
// TreeEstimator read (GridSamples.inl:700-738), RegularTreeNode::read
// (RegularTree.inl:643-662); nodes found by (depth, offset).
struct Density
{
	int dim = 0 , maxDepth = 0;
	double target = 0;
	std::map< std::pair< int , Idx > , double > w;           // weightSum of every node
	std::map< std::pair< int , Idx > , bool > hasChildren;
};

bool ReadDensity( const char *path , int dim , Density &T )
{
	FILE *fp = fopen( path , "rb" );
	if( !fp ){ g_err = "cannot open density" ; return false; }
	bool ok = true;
	auto Read = [&]( void *p , size_t n ){ if( ok && fread( p , 1 , n , fp )!=n ) ok = false; };
	std::vector< unsigned char > skip( 8 + 8*( dim+1 )*( dim+1 ) );
	Read( skip.data() , skip.size() );
	const size_t head = ( ( 2*( dim+1 ) + 7 )/8 )*8 , rec = head + 16 + 8*( 2 + dim + dim*dim );
	std::vector< unsigned char > b( rec );
	T.dim = dim;
	int height = 0;
	auto Node = [&]( void )
	{
		Read( b.data() , rec );
		unsigned short h[5] = { 0 , 0 , 0 , 0 , 0 };
		memcpy( h , b.data() , 2*( dim+1 ) );
		unsigned long long ch;
		double w;
		memcpy( &ch , b.data()+head+8 , 8 );
		memcpy( &w , b.data()+head+16 , 8 );
		Idx o{};
		for( int k=0 ; k<dim ; k++ ) o[k] = h[1+k];
		T.w[ { (int)h[0] , o } ] = w;
		T.hasChildren[ { (int)h[0] , o } ] = ch!=0;
		height = std::max( height , (int)h[0] );
		return ch!=0;
	};
	std::function< void ( bool ) > Read_ = [&]( bool has )
	{
		if( !has || !ok ) return;
		std::vector< bool > c( 1u<<dim );
		for( unsigned i=0 ; i<( 1u<<dim ) ; i++ ) c[i] = Node();
		for( unsigned i=0 ; i<( 1u<<dim ) ; i++ ) Read_( c[i] );
	};
	Read_( Node() );
	Read( &T.target , 8 );
	fclose( fp );
	T.maxDepth = height - 1;
	if( !ok ) g_err = "truncated density";
	return ok;
}

double SamplesPerCell( const Density &T , const double *pIn )
{
	const int dim = T.dim;
	double p[4];
	for( int d=0 ; d<dim ; d++ ) p[d] = pIn[d];
	// getLeafNode from the space root (depth 1, offset 0)
	std::vector< Idx > anc( 2 );                          // anc[t]: ancestor offset at depth t
	anc[1] = Idx{};
	{
		double center[4] , width = 1.;
		for( int d=0 ; d<dim ; d++ ) center[d] = 0.5;
		int t = 1;
		while( true )
		{
			auto it = T.hasChildren.find( { t , anc[t] } );
			if( it==T.hasChildren.end() || !it->second ) break;
			int c = 0;
			for( int d=0 ; d<dim ; d++ ) if( p[d]>center[d] ) c |= 1<<d;
			Idx o = anc[t];
			for( int d=0 ; d<dim ; d++ ) o[d] = 2*o[d] + ( ( c>>d ) & 1 );
			anc.push_back( o );
			width /= 2;
			for( int d=0 ; d<dim ; d++ ) if( c&( 1<<d ) ) center[d] += width/2 ; else center[d] -= width/2;
			t++;
		}
	}
	const int leafDepth = (int)anc.size() - 1;
	auto SPC = [&]( int depth )
	{
		const unsigned res = 1u<<( depth-1 );
		double values[4][2];
		for( int d=0 ; d<dim ; d++ )
		{
			const int e = (int)floor( p[d]*res );
			const double x = p[d]*res - e;
			values[d][1] = 1.;
			values[d][0] = values[d][1] * ( 1.-x ) * 1.;
			values[d][1] *= x * 1.;
		}
		double ws = 0;
		for( int i=0 ; i<( 1<<dim ) ; i++ )
		{
			Idx o = anc[depth];
			int f[4];
			for( int d=0 ; d<dim ; d++ ) f[d] = ( i>>( dim-1-d ) ) & 1 , o[d] += f[d];
			auto it = T.w.find( { depth , o } );
			if( it==T.w.end() ) continue;
			double scale = 1.;
			for( int d=0 ; d<dim ; d++ ) scale *= values[d][ f[d] ];
			ws += it->second * scale;
		}
		return ws / ( 1<<( ( depth-1 )*( dim-2 ) ) );
	};
	// TreeEstimator::_depth with the leaf's depth
	double dep;
	{
		const unsigned maxD = (unsigned)leafDepth;
		unsigned d1 = maxD , d2 = maxD+1;
		while( d2>1 && SPC( d1 )<T.target ) d1-- , d2--;
		if( d2==1 || d2==maxD+1 )
		{
			unsigned d = d2==1 ? 1 : maxD;
			dep = d + log( SPC( d )/T.target ) / log( 2. );
		}
		else
		{
			const double s1 = SPC( d1 ) , s2 = SPC( d2 );
			dep = !s2 ? d1 : d1 + log( T.target/s1 ) / log( s2/s1 );
		}
	}
	dep -= 1.;
	// pow( 0.5 , y ) in GridSamples.inl:849; the reference's build (clang -O3)
	// lowers it to exp2( -y ) (LLVM's power-of-two-base pow simplification), and
	// glibc's exp2 and pow differ in the last bit for some y, so the oracle
	// follows the reference binary
	return T.target * exp2( -( (double)( dim-2 ) * ( T.maxDepth - dep ) ) );
}

// MarchingSimplices::FaceAdjacentConnectedComponents, literally
std::vector< std::vector< size_t > > Components( int K , const std::vector< unsigned > &S )
{
	const size_t n = S.size()/( K+1 );
	auto Face = [&]( size_t s , int v )
	{
		const unsigned *si = &S[ s*( K+1 ) ];
		std::vector< unsigned > f;
		if( K==1 ) f = { si[ v==0 ? 1 : 0 ] };
		else for( int d=0 ; d<K ; d++ ) f.push_back( si[ ( v+d+1 )%( K+1 ) ] );
		std::sort( f.begin() , f.end() );
		return f;
	};
	std::map< std::vector< unsigned > , std::pair< long long , long long > > fData;
	for( size_t i=0 ; i<n ; i++ ) for( int v=0 ; v<=K ; v++ )
	{
		auto it = fData.find( Face( i , v ) );
		if( it==fData.end() ) fData[ Face( i , v ) ] = { (long long)i , -1 };
		else if( it->second.second==-1 ) it->second.second = (long long)i;
		else throw std::string( "Face occuppied" );
	}
	std::vector< std::vector< size_t > > components;
	std::vector< bool > processed( n , false );
	std::vector< size_t > un( n );
	for( size_t i=0 ; i<n ; i++ ) un[i] = i;
	while( un.size() )
	{
		std::vector< size_t > component , Q;
		Q.push_back( un.back() );
		processed[ un.back() ] = true;
		un.pop_back();
		while( Q.size() )
		{
			const size_t i = Q.back();
			Q.pop_back();
			component.push_back( i );
			for( int v=0 ; v<=K ; v++ )
			{
				const auto &fd = fData[ Face( i , v ) ];
				const long long opposite = fd.first==(long long)i ? fd.second : fd.first;
				if( opposite!=-1 && !processed[opposite] ) Q.push_back( (size_t)opposite ) , processed[opposite] = true;
			}
		}
		components.push_back( component );
		for( size_t i=un.size() ; i>0 ; i-- ) if( processed[ un[i-1] ] )
		{
			un[i-1] = un.back();
			un.pop_back();
		}
	}
	return components;
}

template< typename T > T *Dup( const std::vector< T > &v )
{
	T *p = (T *)malloc( std::max< size_t >( 1 , v.size() )*sizeof(T) );
	if( p && !v.empty() ) memcpy( p , v.data() , v.size()*sizeof(T) );
	return p;
}

bool Prepare( const char *path , Tree &t , Iso &iso )
{
	if( !ReadTree( path , t ) ) return false;
	iso.dim = t.dim , iso.D = t.D , iso.res = 1<<t.D , iso.adapted = t.adapted , iso.indexed = &t.indexedFinest;
	Cumulative( t , iso );
	return true;
}

void SubGrid( const Iso &iso , bool all , std::vector< Idx > &nodes , std::set< Idx > &valid , size_t &nV , double *&V , size_t &nS , unsigned *&S ,
              std::vector< double > &vals )
{
	const int dim = iso.dim;
	LevelSetNodes( iso , all , nodes , valid );
	std::vector< unsigned > cells;
	for( const Idx &E : nodes ) for( int d=0 ; d<dim ; d++ ) cells.push_back( (unsigned)E[d] );
	orc_subgrid_triangulation( dim , nodes.size() , cells.data() , 1 , &nV , &V , &nS , &S );
	vals.resize( 2*nV );
	for( size_t v=0 ; v<nV ; v++ )
	{
		Idx F{};
		for( int d=0 ; d<dim ; d++ ) F[d] = (int)floor( V[ v*dim+d ]+0.5 );
		vals[2*v] = iso.V[0][ iso.F( F ) ] , vals[2*v+1] = iso.V[1][ iso.F( F ) ];
	}
}
} // namespace

extern "C" {

const char *orc_extract_error( void ){ return g_err.c_str(); }

unsigned long long orc_indexer_hash( int dim , int adapted ){ return IndexerHash( dim , adapted!=0 ); }

// samples per cell (SamplesPerCell above) of a density file at n unit-cube points
int orc_density_at( const char *densityPath , int dim , size_t n , const double *points , double *out )
{
	Density T;
	if( !ReadDensity( densityPath , dim , T ) ) return -4;
	for( size_t i=0 ; i<n ; i++ ) out[i] = SamplesPerCell( T , points + i*dim );
	return 0;
}

int orc_orient( int k , size_t nS , unsigned *simplices )
{
	try
	{
		std::vector< unsigned > S( simplices , simplices + nS*k );
		Orient( k-1 , S );
		memcpy( simplices , S.data() , S.size()*sizeof(unsigned) );
		return 0;
	}
	catch( const std::string &e ){ g_err = e ; return -3; }
}

int orc_extract_trimmed( const char *treePath , int all , const char *densityPath , double trim ,
                         size_t *nCells , unsigned **cells , double **values /* (2^D+1)^dim x 2 */ ,
                         size_t *nV , double **V /* nV x 3 dim */ , size_t *nS , unsigned **S /* nS x (dim-1) */ ,
                         size_t *nComp , double **compMax );

int orc_extract( const char *treePath , int all ,
                 size_t *nCells , unsigned **cells , double **values /* (2^D+1)^dim x 2 */ ,
                 size_t *nV , double **V /* nV x 3 dim */ , size_t *nS , unsigned **S /* nS x (dim-1) */ )
{
	size_t nc = 0;
	double *cm = nullptr;
	const int rc = orc_extract_trimmed( treePath , all , nullptr , 0. , nCells , cells , values , nV , V , nS , S , &nc , &cm );
	free( cm );
	return rc;
}

int orc_extract_trimmed( const char *treePath , int all , const char *densityPath , double trim ,
                         size_t *nCells , unsigned **cells , double **values ,
                         size_t *nV , double **V , size_t *nS , unsigned **S ,
                         size_t *nComp , double **compMax )
{
	try
	{
		Tree t;
		Iso iso;
		if( !Prepare( treePath , t , iso ) ) return -1;
		const int dim = t.dim;
		std::vector< Idx > nodes;
		std::set< Idx > valid;
		size_t tnV = 0 , tnS = 0;
		double *TV = nullptr;
		unsigned *TS = nullptr;
		std::vector< double > vals;
		SubGrid( iso , all!=0 , nodes , valid , tnV , TV , tnS , TS , vals );
		std::vector< unsigned > c;
		for( const Idx &E : nodes ) for( int d=0 ; d<dim ; d++ ) c.push_back( (unsigned)E[d] );
		*nCells = nodes.size() , *cells = Dup( c );
		std::vector< double > vv( 2*iso.V[0].size() );
		for( size_t i=0 ; i<iso.V[0].size() ; i++ ) vv[2*i] = iso.V[0][i] , vv[2*i+1] = iso.V[1][i];
		*values = Dup( vv );
		size_t nLV = 0 , nLS = 0;
		double *LV = nullptr;
		unsigned *LS = nullptr;
		const double zero[2] = { 0 , 0 };
		orc_level_set( dim , 2 , dim , tnV , TV , tnS , TS , vals.data() , zero , &nLV , &LV , &nLS , &LS );
		free( TV ) , free( TS );
		std::vector< unsigned > simp( LS , LS + nLS*( dim-1 ) );
		free( LS );
		Orient( dim-2 , simp );
		std::vector< double > cmax;
		if( densityPath && trim>0 )
		{
			Density T;
			if( !ReadDensity( densityPath , dim , T ) ){ free( LV ) ; return -4; }
			const int K = dim-2;
			const double res = (double)( 1u<<t.D );
			std::vector< std::vector< size_t > > comps = Components( K , simp );
			std::vector< unsigned > temp;
			for( const std::vector< size_t > &c : comps )
			{
				double maxDensity = 0;
				for( size_t j : c ) for( int v=0 ; v<=K ; v++ )
				{
					double p[4];
					for( int d=0 ; d<dim ; d++ ) p[d] = LV[ simp[ j*( K+1 )+v ]*dim+d ] / res;
					maxDensity = std::max< double >( maxDensity , SamplesPerCell( T , p ) );
				}
				cmax.push_back( maxDensity );
				if( maxDensity>trim ) for( size_t j : c ) for( int v=0 ; v<=K ; v++ ) temp.push_back( simp[ j*( K+1 )+v ] );
			}
			simp.swap( temp );
		}
		*nComp = cmax.size() , *compMax = Dup( cmax );
		double scale = 1.;
		for( int d=0 ; d<dim ; d++ ) scale *= t.xf[ d*( dim+1 )+d ];
		scale = pow( scale , 1./dim );
		const double res = (double)( 1u<<t.D );
		std::vector< double > out( nLV*3*dim );
		for( size_t i=0 ; i<nLV ; i++ )
		{
			double p[4] , grad[2][4];
			for( int d=0 ; d<dim ; d++ ) p[d] = LV[ i*dim+d ] * ( 1./res );
			if( !Gradient( iso , valid , p , grad ) ){ free( LV ) ; g_err = "Leaf node is not a valid element node" ; return -2; }
			ToWorld( t , LV + i*dim , &out[ i*3*dim ] );
			for( int c=0 ; c<2 ; c++ ) for( int d=0 ; d<dim ; d++ ) out[ i*3*dim + dim*( 1+c ) + d ] = ( grad[c][d] * ( 1./res ) ) * ( 1./scale );
		}
		free( LV );
		*nV = nLV , *V = Dup( out );
		*nS = simp.size()/( dim-1 ) , *S = Dup( simp );
		return 0;
	}
	catch( const std::string &e ){ g_err = e ; return -3; }
}

int orc_extract_level_set( const char *treePath , int component ,
                           size_t *nV , double **V /* nV x dim */ , size_t *nS , unsigned **S /* nS x dim */ )
{
	try
	{
		Tree t;
		Iso iso;
		if( !Prepare( treePath , t , iso ) ) return -1;
		const int dim = t.dim;
		std::vector< Idx > nodes;
		std::set< Idx > valid;
		size_t tnV = 0 , tnS = 0;
		double *TV = nullptr;
		unsigned *TS = nullptr;
		std::vector< double > vals;
		SubGrid( iso , false , nodes , valid , tnV , TV , tnS , TS , vals );
		std::vector< double > col( tnV );
		for( size_t v=0 ; v<tnV ; v++ ) col[v] = vals[ 2*v+component ];
		size_t nLV = 0 , nLS = 0;
		double *LV = nullptr;
		unsigned *LS = nullptr;
		const double zero[2] = { 0 , 0 };
		orc_level_set( dim , 1 , dim , tnV , TV , tnS , TS , col.data() , zero , &nLV , &LV , &nLS , &LS );
		free( TV ) , free( TS );
		std::vector< double > out( nLV*dim );
		for( size_t i=0 ; i<nLV ; i++ ) ToWorld( t , LV + i*dim , &out[ i*dim ] );
		free( LV );
		*nV = nLV , *V = Dup( out );
		*nS = nLS , *S = LS;
		return 0;
	}
	catch( const std::string &e ){ g_err = e ; return -3; }
}

}
