// Minimal PLY vertex reader for oriented samples: x,y,z[,w] and the full
// frame matrix skew_c_r (the properties ExteriorPoissonRecon.cpp:282-303
// reads through Misha's VertexFactory; PlyVertexData.inl:228-234 naming).
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace epr
{
struct PlySamples
{
	int dim = 0;
	std::vector< double > pos , skew;   // n*dim , n*dim*dim
	size_t size( void ) const { return dim ? pos.size()/dim : 0; }
};

inline int PlyTypeSize( const std::string &t )
{
	if( t=="char" || t=="uchar" || t=="int8" || t=="uint8" ) return 1;
	if( t=="short" || t=="ushort" || t=="int16" || t=="uint16" ) return 2;
	if( t=="int" || t=="uint" || t=="int32" || t=="uint32" || t=="float" || t=="float32" ) return 4;
	if( t=="double" || t=="float64" ) return 8;
	return 0;
}

inline double PlyRead( const unsigned char *p , const std::string &t , bool swap )
{
	unsigned char b[8];
	int s = PlyTypeSize( t );
	for( int i=0 ; i<s ; i++ ) b[i] = swap ? p[s-1-i] : p[i];
	if( t=="char" || t=="int8" ){ int8_t v ; memcpy( &v , b , 1 ) ; return v; }
	if( t=="uchar" || t=="uint8" ){ uint8_t v ; memcpy( &v , b , 1 ) ; return v; }
	if( t=="short" || t=="int16" ){ int16_t v ; memcpy( &v , b , 2 ) ; return v; }
	if( t=="ushort" || t=="uint16" ){ uint16_t v ; memcpy( &v , b , 2 ) ; return v; }
	if( t=="int" || t=="int32" ){ int32_t v ; memcpy( &v , b , 4 ) ; return v; }
	if( t=="uint" || t=="uint32" ){ uint32_t v ; memcpy( &v , b , 4 ) ; return v; }
	if( t=="float" || t=="float32" ){ float v ; memcpy( &v , b , 4 ) ; return v; }
	double v ; memcpy( &v , b , 8 ) ; return v;
}

// Returns an empty set (dim 0) and a message on failure.
inline PlySamples ReadPlySamples( const std::string &path , std::string &err )
{
	PlySamples out;
	std::ifstream f( path , std::ios::binary );
	if( !f ){ err = "cannot open " + path ; return out; }
	std::string line , format;
	std::getline( f , line );
	if( line.rfind( "ply" , 0 )!=0 ){ err = "not a PLY file: " + path ; return out; }
	size_t nv = 0;
	bool inVertex = false , vertexFirst = true , seenElement = false;
	std::vector< std::pair< std::string , std::string > > props;
	while( std::getline( f , line ) )
	{
		if( !line.empty() && line.back()=='\r' ) line.pop_back();
		std::istringstream ss( line );
		std::string k ; ss >> k;
		if( k=="format" ) ss >> format;
		else if( k=="element" )
		{
			std::string name ; size_t c ; ss >> name >> c;
			inVertex = ( name=="vertex" );
			if( inVertex ){ nv = c ; if( seenElement ) vertexFirst = false; }
			seenElement = true;
		}
		else if( k=="property" && inVertex )
		{
			std::string t , name ; ss >> t;
			if( t=="list" ){ err = "list property in vertex element" ; return out; }
			ss >> name;
			props.emplace_back( name , t );
		}
		else if( k=="end_header" ) break;
	}
	if( !vertexFirst ){ err = "the vertex element must come first" ; return out; }
	std::map< std::string , int > col;
	for( size_t i=0 ; i<props.size() ; i++ ) col[ props[i].first ] = (int)i;
	int dim = col.count( "w" ) ? 4 : 3;
	const char *names[4] = { "x" , "y" , "z" , "w" };
	for( int d=0 ; d<dim ; d++ ) if( !col.count( names[d] ) ){ err = "missing position property" ; return out; }
	for( int c=0 ; c<dim ; c++ ) for( int r=0 ; r<dim ; r++ )
		if( !col.count( "skew_" + std::to_string( c ) + "_" + std::to_string( r ) ) ){ err = "missing skew property" ; return out; }
	out.dim = dim;
	out.pos.resize( nv*dim ) , out.skew.resize( nv*dim*dim );
	std::vector< double > row( props.size() );
	if( format=="ascii" )
	{
		for( size_t v=0 ; v<nv ; v++ ) for( size_t i=0 ; i<props.size() ; i++ ) if( !( f >> row[i] ) ){ err = "truncated PLY" ; out.dim = 0 ; return out; }
			else if( i+1==props.size() )
			{
				for( int d=0 ; d<dim ; d++ ) out.pos[v*dim+d] = row[ col[ names[d] ] ];
				for( int c=0 ; c<dim ; c++ ) for( int r=0 ; r<dim ; r++ ) out.skew[ (v*dim+c)*dim+r ] = row[ col[ "skew_" + std::to_string( c ) + "_" + std::to_string( r ) ] ];
			}
		return out;
	}
	bool swap = ( format=="binary_big_endian" );
	if( !swap && format!="binary_little_endian" ){ err = "unknown PLY format " + format ; out.dim = 0 ; return out; }
	size_t stride = 0;
	std::vector< size_t > off( props.size() );
	for( size_t i=0 ; i<props.size() ; i++ ){ off[i] = stride ; stride += PlyTypeSize( props[i].second ); }
	std::vector< unsigned char > buf( stride*nv );
	f.read( (char*)buf.data() , buf.size() );
	if( (size_t)f.gcount()!=buf.size() ){ err = "truncated PLY" ; out.dim = 0 ; return out; }
	for( size_t v=0 ; v<nv ; v++ )
	{
		const unsigned char *b = &buf[v*stride];
		for( size_t i=0 ; i<props.size() ; i++ ) row[i] = PlyRead( b+off[i] , props[i].second , swap );
		for( int d=0 ; d<dim ; d++ ) out.pos[v*dim+d] = row[ col[ names[d] ] ];
		for( int c=0 ; c<dim ; c++ ) for( int r=0 ; r<dim ; r++ ) out.skew[ (v*dim+c)*dim+r ] = row[ col[ "skew_" + std::to_string( c ) + "_" + std::to_string( r ) ] ];
	}
	return out;
}
} // namespace epr
