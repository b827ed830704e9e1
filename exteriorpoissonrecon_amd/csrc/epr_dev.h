// Device-memory helpers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <exception>
#include <string>
#include <utility>

namespace epr
{
struct HipError { std::string msg; };
#define HIPCHECK( call ) do { hipError_t _e = ( call ) ; if( _e!=hipSuccess ) throw ::epr::HipError{ std::string( #call ) + ": " + hipGetErrorString( _e ) }; } while( 0 )

template< typename T >
struct DevBuf
{
	T *p = nullptr;
	size_t n = 0;
	DevBuf(){}
	DevBuf( const DevBuf & ) = delete;
	DevBuf &operator=( const DevBuf & ) = delete;
	DevBuf( DevBuf &&o ) noexcept : p( o.p ) , n( o.n ) { o.p = nullptr , o.n = 0; }
	DevBuf &operator=( DevBuf &&o ) noexcept { std::swap( p , o.p ) ; std::swap( n , o.n ) ; return *this; }
	~DevBuf(){ if( p ) (void)hipFree( p ); }
	void ensure( size_t count ){ if( count>n ) alloc( count ); }
	void alloc( size_t count )
	{
		if( p ){ (void)hipFree( p ) ; p = nullptr; }
		n = count;
		if( count ) HIPCHECK( hipMalloc( (void**)&p , count*sizeof(T) ) );
	}
	void zero( hipStream_t s ){ if( n ) HIPCHECK( hipMemsetAsync( p , 0 , n*sizeof(T) , s ) ); }
	void upload( const T *h , size_t count , hipStream_t s ){ if( count>n ) alloc( count ) ; if( count ) HIPCHECK( hipMemcpyAsync( p , h , count*sizeof(T) , hipMemcpyHostToDevice , s ) ); }
	void download( T *h , size_t count , hipStream_t s ) const { if( count ) HIPCHECK( hipMemcpyAsync( h , p , count*sizeof(T) , hipMemcpyDeviceToHost , s ) ); }
};

static inline unsigned int Blocks( long long n , int bs ){ return (unsigned int)( ( n + bs - 1 ) / bs ); }

// the calling thread's last error of a context-free entry point (epr_last_error(NULL))
std::string &ThreadError( void );

// runs f, mapping exceptions to status codes and messages (EPR_* of epr.h)
template< typename F >
int ContextFree( F f )
{
	try { f() ; ThreadError().clear() ; return 0; }
	catch( const HipError &e ){ ThreadError() = e.msg ; return -3; }
	catch( const std::string &s ){ ThreadError() = s ; return -4; }
	catch( const std::exception &e ){ ThreadError() = e.what() ; return -2; }
}
} // namespace epr
