// Host construction of the adapted hierarchy (the default --aRadius path):
// which functions and elements each depth indexes, their numbering, and the
// tree that the .tree file serialises.  Include/Hat.adaptive.inl:349-729.
//
// The reference grows a pointer tree (Misha/RegularTree.inl) from the sample
// cells and numbers its nodes; everything it produces is a function of the
// sample cells, the depth and the radius, and is computed here from sorted
// offset lists and dense flag grids instead:
//
//  * internal depth e = external depth + 1: the root (e = 0) has 2^Dim
//    children whose first one is the unit cube, so a function F at external
//    depth d is the node of offset F at internal depth d+1;
//  * active functions at depth d: the box A + [-K, K+1]^Dim around the depth-d
//    ancestor A of every sample cell, clipped to [0, 2^d]^Dim
//    (Hat.adaptive.inl:596-614);
//  * they are numbered in the tree's depth-first order (:617-641), which for
//    nodes of one depth is the Morton order with coordinate 0 in the lowest bit
//    of each level (child c has offset bit d = (c>>d)&1, RegularTree.inl:155-172);
//  * then, depth by depth and for the active functions in that order, the
//    3^Dim window (coordinate 0 slowest) numbers the functions not yet indexed
//    (padding) and the elements E <= F not yet indexed, in discovery order
//    (:658-697);
//  * tree nodes exist as whole broods: a node has children if it lies on a
//    sample cell's root path (insertPoint, RegularTree.inl:206-223), or if it is
//    the parent of a node of a neighbour window that was gathered with node
//    creation: the [-K, K+1] window of every sample cell's ancestor at every
//    depth (:596) and the [-1, 1] window of every active function's ancestor at
//    every depth (:658).
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace epr
{
template< int Dim >
struct AdaptedHierarchy
{
	struct Level
	{
		int res = 0;                           // 2^d
		std::vector< unsigned char > fmask;    // (res+1)^Dim functions: 0 not indexed, 1 active, 2 padding
		std::vector< unsigned char > emask;    // res^Dim elements (coordinate 0 least significant): 1 indexed
		std::vector< long long > fOrder;       // dense index of every indexed function, in the reference's numbering
		std::vector< long long > eOrder;       // dense index of every indexed element, in the reference's numbering
		size_t nActive = 0;
	};
	int D = 0 , K = 0;
	std::vector< Level > levels;                          // external depth 0..D
	std::vector< std::vector< unsigned char > > expanded; // internal depth 0..D: (2^e)^Dim flags of nodes with children

	// ---- index helpers: functions with coordinate 0 most significant (the dense
	// levels), elements with coordinate 0 least significant (the per-element arrays)
	static long long Flat( const int *F , long long n ){ long long i = 0 ; for( int d=0 ; d<Dim ; d++ ) i = i*n + F[d] ; return i; }
	static long long EFlat( const int *E , long long n ){ long long i = 0 ; for( int d=Dim-1 ; d>=0 ; d-- ) i = i*n + E[d] ; return i; }
	static void Unflat( long long i , long long n , int *F ){ for( int d=Dim-1 ; d>=0 ; d-- ){ F[d] = (int)( i % n ) ; i /= n; } }
	static unsigned long long Pack( const int *F ){ unsigned long long k = 0 ; for( int d=0 ; d<Dim ; d++ ) k = ( k<<16 ) | (unsigned long long)F[d] ; return k; }
	static void Unpack( unsigned long long k , int *F ){ for( int d=Dim-1 ; d>=0 ; d-- ){ F[d] = (int)( k & 0xFFFF ) ; k >>= 16; } }
	static unsigned long long Morton( const int *F , int bits )
	{
		unsigned long long k = 0;
		for( int l=0 ; l<bits ; l++ ) for( int d=0 ; d<Dim ; d++ ) k |= (unsigned long long)( ( F[d]>>l ) & 1 ) << ( l*Dim + d );
		return k;
	}
	// calls f( G ) for every G in the box [lo, hi] (inclusive, per coordinate)
	template< typename Fn > static void Box( const int *lo , const int *hi , Fn &&f )
	{
		for( int d=0 ; d<Dim ; d++ ) if( lo[d]>hi[d] ) return;
		int G[Dim];
		for( int d=0 ; d<Dim ; d++ ) G[d] = lo[d];
		while( true )
		{
			f( (const int *)G );
			int d = Dim-1;
			for( ; d>=0 ; d-- ){ if( ++G[d]<=hi[d] ) break ; G[d] = lo[d]; }
			if( d<0 ) break;
		}
	}
	static void SortUnique( std::vector< unsigned long long > &v ){ std::sort( v.begin() , v.end() ) ; v.erase( std::unique( v.begin() , v.end() ) , v.end() ); }

	static void CheckArgs( int depth , int radius )
	{
		if( radius<0 || radius>5 ) throw std::string( "Only kernel radius in the range [0,5] supported: " + std::to_string( radius ) );
		if( depth>14 ) throw std::string( "adapted depth must be at most 14" );
	}
	// unique ancestors of the sample cells at every depth (packed offsets, sorted)
	static std::vector< std::vector< unsigned long long > > Ancestors( const int *cells , size_t ncell , int D )
	{
		std::vector< std::vector< unsigned long long > > anc( D+1 );
		for( size_t c=0 ; c<ncell ; c++ ) anc[D].push_back( Pack( cells + c*Dim ) );
		SortUnique( anc[D] );
		for( int d=D-1 ; d>=0 ; d-- )
		{
			anc[d].reserve( anc[d+1].size() );
			for( unsigned long long k : anc[d+1] ){ int A[Dim] ; Unpack( k , A ) ; for( int dd=0 ; dd<Dim ; dd++ ) A[dd] >>= 1 ; anc[d].push_back( Pack( A ) ); }
			SortUnique( anc[d] );
		}
		return anc;
	}

	// cells: ncell x Dim cell indices at depth D (the ordered sampler's cells)
	void build( const int *cells , size_t ncell , int depth , int radius )
	{
		buildLevels( cells , ncell , depth , radius );
		buildTree( cells , ncell );
	}

	// the per-depth function and element sets and their numbering
	void buildLevels( const int *cells , size_t ncell , int depth , int radius )
	{
		CheckArgs( depth , radius );
		D = depth , K = radius;
		levels.assign( D+1 , Level() );
		std::vector< std::vector< unsigned long long > > anc = Ancestors( cells , ncell , D );
		// active functions, their depth-first order, then padding and elements by discovery
		for( int d=0 ; d<=D ; d++ )
		{
			Level &L = levels[d];
			L.res = 1<<d;
			const long long n1 = L.res+1;
			long long nf = 1 , ne = 1;
			for( int dd=0 ; dd<Dim ; dd++ ) nf *= n1 , ne *= L.res;
			L.fmask.assign( nf , 0 ) , L.emask.assign( ne , 0 );
			for( unsigned long long k : anc[d] )
			{
				int A[Dim] , lo[Dim] , hi[Dim];
				Unpack( k , A );
				for( int dd=0 ; dd<Dim ; dd++ ) lo[dd] = std::max( A[dd]-K , 0 ) , hi[dd] = std::min( A[dd]+K+1 , L.res );
				Box( lo , hi , [&]( const int *G ){ L.fmask[ Flat( G , n1 ) ] = 1; } );
			}
			std::vector< std::pair< unsigned long long , long long > > act;
			for( long long i=0 ; i<nf ; i++ ) if( L.fmask[i] ){ int F[Dim] ; Unflat( i , n1 , F ) ; act.emplace_back( Morton( F , d+1 ) , i ); }
			std::sort( act.begin() , act.end() );
			L.nActive = act.size();
			L.fOrder.resize( act.size() );
			for( size_t i=0 ; i<act.size() ; i++ ) L.fOrder[i] = act[i].second;
			for( size_t i=0 ; i<L.nActive ; i++ )
			{
				int F[Dim];
				Unflat( L.fOrder[i] , n1 , F );
				int lo[Dim] , hi[Dim];
				for( int dd=0 ; dd<Dim ; dd++ ) lo[dd] = F[dd]-1 , hi[dd] = F[dd]+1;
				Box( lo , hi , [&]( const int *G )
				{
					bool inF = true , inE = true;
					for( int dd=0 ; dd<Dim ; dd++ )
					{
						inF = inF && G[dd]>=0 && G[dd]<=L.res;
						inE = inE && G[dd]>=0 && G[dd]<L.res && G[dd]<=F[dd];
					}
					if( inF ){ const long long g = Flat( G , n1 ) ; if( !L.fmask[g] ){ L.fmask[g] = 2 ; L.fOrder.push_back( g ); } }
					if( inE ){ const long long e = EFlat( G , L.res ) ; if( !L.emask[e] ){ L.emask[e] = 1 ; L.eOrder.push_back( e ); } }
				} );
			}
		}
	}

	// the tree's shape (nodes with children) from the levels (however they were
	// built: D, K and levels[] set) and the sample cells
	void buildTree( const int *cells , size_t ncell )
	{
		std::vector< std::vector< unsigned long long > > anc = Ancestors( cells , ncell , D );
		expanded.assign( D+1 , std::vector< unsigned char >() );
		// nodes with children, by internal depth e (offsets in [0, 2^e))
		for( int e=0 ; e<=D ; e++ ){ long long n = 1 ; for( int dd=0 ; dd<Dim ; dd++ ) n *= 1LL<<e ; expanded[e].assign( n , 0 ); }
		expanded[0][0] = 1;
		auto MarkParents = [&]( int w , const int *lo , const int *hi )   // window [lo, hi] at internal depth w
		{
			const int top = ( 1<<w ) - 1;
			int plo[Dim] , phi[Dim];
			for( int dd=0 ; dd<Dim ; dd++ ) plo[dd] = std::max( lo[dd] , 0 )>>1 , phi[dd] = std::min( hi[dd] , top )>>1;
			std::vector< unsigned char > &X = expanded[w-1];
			Box( plo , phi , [&]( const int *P ){ X[ Flat( P , 1LL<<(w-1) ) ] = 1; } );
		};
		for( int w=1 ; w<=D+1 ; w++ ) for( unsigned long long k : anc[w-1] )
		{
			int A[Dim] , lo[Dim] , hi[Dim];
			Unpack( k , A );
			if( w<=D ) expanded[w][ Flat( A , 1LL<<w ) ] = 1;          // on the root path of a sample cell
			for( int dd=0 ; dd<Dim ; dd++ ) lo[dd] = A[dd]-K , hi[dd] = A[dd]+K+1;
			MarkParents( w , lo , hi );
		}
		// ancestors at internal depth w of the active functions of depths >= w-1
		std::vector< unsigned long long > S;
		for( int w=D+1 ; w>=1 ; w-- )
		{
			std::vector< unsigned long long > T;
			T.reserve( S.size() + levels[w-1].nActive );
			for( unsigned long long k : S ){ int A[Dim] ; Unpack( k , A ) ; for( int dd=0 ; dd<Dim ; dd++ ) A[dd] >>= 1 ; T.push_back( Pack( A ) ); }
			const Level &L = levels[w-1];
			for( size_t i=0 ; i<L.nActive ; i++ ){ int F[Dim] ; Unflat( L.fOrder[i] , L.res+1 , F ) ; T.push_back( Pack( F ) ); }
			SortUnique( T );
			S.swap( T );
			for( unsigned long long k : S )
			{
				int A[Dim] , lo[Dim] , hi[Dim];
				Unpack( k , A );
				for( int dd=0 ; dd<Dim ; dd++ ) lo[dd] = A[dd]-1 , hi[dd] = A[dd]+1;
				MarkParents( w , lo , hi );
			}
		}
	}

	size_t numFunctions( int d ) const { return levels[d].fOrder.size(); }
	size_t numElements( int d ) const { return levels[d].eOrder.size(); }

	// RegularTreeNode< Dim , NodeData , unsigned short > as the reference writes it
	// raw (Misha/RegularTree.h:95-105, Streams.h): depth and offsets, padding to 8,
	// parent and children pointers, then NodeData{ functionIndex , elementIndex }.
	static constexpr size_t HeadBytes = ( ( 2*(Dim+1) + 7 ) / 8 ) * 8;
	static constexpr size_t NodeBytes = HeadBytes + 32;

	// The indexer part of a .tree (HierarchicalAdaptedIndexer::write,
	// Hat.adaptive.inl:408-414): the tree (RegularTree.inl:616-641, unserialised:
	// the root, then depth first every node's brood of children), the active
	// counts and the per-depth function and element node tables.  Pointer values
	// are meaningless in a file (the reader re-links them); they are written as 0,
	// except that a node with children carries a non-zero children field, which is
	// what the reader tests.
	void write( FILE *fp ) const
	{
		std::vector< std::vector< unsigned int > > fIdx( D+1 ) , eIdx( D+1 );
		for( int d=0 ; d<=D ; d++ )
		{
			const Level &L = levels[d];
			fIdx[d].assign( L.fmask.size() , 0xFFFFFFFFu ) , eIdx[d].assign( L.emask.size() , 0xFFFFFFFFu );
			for( size_t i=0 ; i<L.fOrder.size() ; i++ ) fIdx[d][ L.fOrder[i] ] = (unsigned int)i;
			for( size_t i=0 ; i<L.eOrder.size() ; i++ ) eIdx[d][ L.eOrder[i] ] = (unsigned int)i;
		}
		std::vector< unsigned char > buf;
		buf.reserve( 1<<22 );
		auto flush = [&]( bool force ){ if( force || buf.size()>( 1u<<22 ) ){ if( buf.size() && fwrite( buf.data() , 1 , buf.size() , fp )!=buf.size() ) throw std::string( "short write" ) ; buf.clear(); } };
		auto node = [&]( int e , const int *O )
		{
			unsigned char rec[NodeBytes];
			memset( rec , 0 , sizeof(rec) );
			unsigned short h[Dim+1];
			h[0] = (unsigned short)e;
			for( int d=0 ; d<Dim ; d++ ) h[d+1] = (unsigned short)O[d];
			memcpy( rec , h , sizeof(h) );
			unsigned long long children = ( e<=D && expanded[e][ Flat( O , 1LL<<e ) ] ) ? 1 : 0;
			unsigned long long fi = ~0ULL , ei = ~0ULL;
			if( e>0 )
			{
				const int d = e-1 , res = 1<<d;
				bool inF = true , inE = true;
				for( int dd=0 ; dd<Dim ; dd++ ) inF = inF && O[dd]<=res , inE = inE && O[dd]<res;
				if( inF ){ unsigned int v = fIdx[d][ Flat( O , res+1 ) ] ; if( v!=0xFFFFFFFFu ) fi = v; }
				if( inE ){ unsigned int v = eIdx[d][ EFlat( O , res ) ] ; if( v!=0xFFFFFFFFu ) ei = v; }
			}
			memcpy( rec + HeadBytes + 8 , &children , 8 );
			memcpy( rec + HeadBytes + 16 , &fi , 8 );
			memcpy( rec + HeadBytes + 24 , &ei , 8 );
			buf.insert( buf.end() , rec , rec+NodeBytes );
			flush( false );
		};
		// depth-first: a node's brood, then each child's subtree
		std::vector< int > off( (size_t)( D+2 )*Dim , 0 );
		struct Frame { int e , c; };
		{
			int O[Dim] = {};
			node( 0 , O );
		}
		std::vector< Frame > st;
		auto pushBrood = [&]( int e )   // node at depth e with offset off[e]; writes its children
		{
			const int *P = &off[ (size_t)e*Dim ];
			for( int c=0 ; c<(1<<Dim) ; c++ )
			{
				int O[Dim];
				for( int d=0 ; d<Dim ; d++ ) O[d] = ( P[d]<<1 ) | ( ( c>>d ) & 1 );
				node( e+1 , O );
			}
			st.push_back( { e , 0 } );
		};
		pushBrood( 0 );
		while( !st.empty() )
		{
			Frame &f = st.back();
			if( f.c==( 1<<Dim ) ){ st.pop_back() ; continue; }
			const int e = f.e , c = f.c++;
			int *O = &off[ (size_t)( e+1 )*Dim ];
			const int *P = &off[ (size_t)e*Dim ];
			for( int d=0 ; d<Dim ; d++ ) O[d] = ( P[d]<<1 ) | ( ( c>>d ) & 1 );
			if( e+1<=D && expanded[e+1][ Flat( O , 1LL<<(e+1) ) ] ) pushBrood( e+1 );
		}
		auto u64 = [&]( unsigned long long v ){ const unsigned char *p = (const unsigned char *)&v ; buf.insert( buf.end() , p , p+8 ); };
		u64( D+1 );
		for( int d=0 ; d<=D ; d++ ) u64( levels[d].nActive );
		for( int pass=0 ; pass<2 ; pass++ )
		{
			u64( D+1 );
			for( int d=0 ; d<=D ; d++ )
			{
				const size_t cnt = pass==0 ? numFunctions( d ) : numElements( d );
				u64( cnt );
				buf.insert( buf.end() , cnt*8 , 0 );
				flush( false );
			}
		}
		flush( true );
	}
};
} // namespace epr
