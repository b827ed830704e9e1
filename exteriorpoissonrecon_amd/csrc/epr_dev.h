// Device-memory helpers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <exception>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace epr
{
struct HipError { std::string msg; };
#define HIPCHECK( call ) do { hipError_t _e = ( call ) ; if( _e!=hipSuccess ) throw ::epr::HipError{ std::string( #call ) + ": " + hipGetErrorString( _e ) }; } while( 0 )

// Large device blocks are kept for reuse instead of being freed: the extraction
// allocates several gigabytes of temporaries per call, and fresh allocations of
// that size pay the driver's mapping cost every time (4D: 0.5 s on the first
// call, 3 s on later ones without the cache).  A cached block is handed out
// again only after a device synchronisation, which is what hipFree waits for.
// Blocks stay cached until the process ends (or an allocation fails).
struct BlockCache
{
	static constexpr size_t MinBytes = size_t( 32 ) << 20 , MaxHeld = size_t( 64 ) << 30;
	struct Block { void *p ; size_t bytes ; int dev; };
	std::mutex m;
	std::vector< Block > blocks;
	size_t held = 0;
	static BlockCache &Get( void ){ static BlockCache *c = new BlockCache ; return *c; }   // never destroyed: blocks outlive static teardown
	// a cached block of the current device with between bytes and 5/4 bytes (its size in got), or null
	void *take( size_t bytes , size_t &got )
	{
		if( bytes<MinBytes ) return nullptr;
		int dev = 0;
		if( hipGetDevice( &dev )!=hipSuccess ) return nullptr;
		Block b;
		{
			std::lock_guard< std::mutex > lock( m );
			size_t best = blocks.size();
			for( size_t i=0 ; i<blocks.size() ; i++ )
				if( blocks[i].dev==dev && blocks[i].bytes>=bytes && blocks[i].bytes<=bytes+bytes/4 && ( best==blocks.size() || blocks[i].bytes<blocks[best].bytes ) ) best = i;
			if( best==blocks.size() ) return nullptr;
			b = blocks[best];
			held -= b.bytes;
			blocks[best] = blocks.back() , blocks.pop_back();
		}
		// the block is this caller's now; the wait for the work that may still
		// use it runs outside the lock, so other threads' takes and gives do not
		// queue behind a device synchronisation
		if( hipDeviceSynchronize()!=hipSuccess ){ (void)hipGetLastError() ; (void)hipFree( b.p ) ; return nullptr; }
		got = b.bytes;
		return b.p;
	}
	// keeps a block (true) or leaves it to the caller to free
	bool give( void *p , size_t bytes )
	{
		if( bytes<MinBytes ) return false;
		int dev = 0;
		if( hipGetDevice( &dev )!=hipSuccess ) return false;
		std::lock_guard< std::mutex > lock( m );
		if( held+bytes>MaxHeld ) return false;
		blocks.push_back( { p , bytes , dev } ) , held += bytes;
		return true;
	}
	// frees the current device's cached blocks
	void flush( void )
	{
		int dev = 0;
		if( hipGetDevice( &dev )!=hipSuccess ) return;
		std::lock_guard< std::mutex > lock( m );
		for( size_t i=0 ; i<blocks.size() ; )
			if( blocks[i].dev==dev ){ (void)hipFree( blocks[i].p ) ; held -= blocks[i].bytes ; blocks[i] = blocks.back() ; blocks.pop_back(); }
			else i++;
	}
};

template< typename T >
struct DevBuf
{
	T *p = nullptr;
	size_t n = 0 , bytes = 0;
	DevBuf(){}
	DevBuf( const DevBuf & ) = delete;
	DevBuf &operator=( const DevBuf & ) = delete;
	DevBuf( DevBuf &&o ) noexcept : p( o.p ) , n( o.n ) , bytes( o.bytes ) { o.p = nullptr , o.n = 0 , o.bytes = 0; }
	DevBuf &operator=( DevBuf &&o ) noexcept { std::swap( p , o.p ) ; std::swap( n , o.n ) ; std::swap( bytes , o.bytes ) ; return *this; }
	~DevBuf(){ release(); }
	void release( void )
	{
		if( p && !BlockCache::Get().give( p , bytes ) ) (void)hipFree( p );
		p = nullptr , n = 0 , bytes = 0;
	}
	void ensure( size_t count ){ if( count>n ) alloc( count ); }
	void alloc( size_t count )
	{
		release();
		if( !count ) return;
		const size_t want = count*sizeof(T);
		size_t got = want;
		void *q = BlockCache::Get().take( want , got );   // may be larger than asked for; it returns to the cache at its own size
		if( !q )
		{
			hipError_t e = hipMalloc( &q , want );
			if( e!=hipSuccess ){ (void)hipGetLastError() ; BlockCache::Get().flush() ; e = hipMalloc( &q , want ); }
			if( e!=hipSuccess ) throw HipError{ std::string( "hipMalloc( " ) + std::to_string( want ) + " bytes ): " + hipGetErrorString( e ) };
		}
		p = (T *)q , n = count , bytes = got;
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
