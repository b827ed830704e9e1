// Multigrid V-cycle on the device: the reference's Include/MG.h GSRelaxer
// (:37-77) and MGSolver (:79-195) for an explicit sparse matrix and a
// hierarchy of prolongation matrices.
//
//   setup   M_{pNum} = M,  M_i = P_i^T M_{i+1} P_i  (sparse products on the
//           device, (R M) P in the reference's evaluation order, terms of
//           every entry summed in increasing inner index), coarse M_0
//           inverted by Gauss-Jordan in one workgroup (symmetric positive
//           definite: no pivoting) in place of SimplicialLDLT.
//   cycle   multicolour Gauss-Seidel x_i += (b_i - sum_j M_ji x_j) / D_i over
//           the column of i in storage order, colour after colour; residual
//           b - M x and restriction P^T r as row gathers in increasing column
//           order (the order of Eigen's column-major sparse * dense product),
//           coarse solve as a dense product, prolongation x += P X.
//
// Sparse matrices are CSC as in Eigen (column pointers, sorted row indices).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/epr.h"
#include "epr_dev.h"

namespace epr
{
namespace mg
{
static inline unsigned int Grid( long long n ){ return (unsigned int)std::max< long long >( 1 , ( n + 255 )/256 ); }

// compressed sparse matrix on the device: ptr (outer+1), idx (inner), val
struct DCompressed
{
	size_t outer = 0 , inner = 0 , nnz = 0;
	DevBuf< unsigned > ptr , idx;
	DevBuf< double > val;
};

#pragma clang fp contract(off)

// x_i += ( b_i - sum_k M(row_k, i) x_{row_k} ) / D_i for the indices of one colour
__global__ void __launch_bounds__( 256 ) KRelaxColour( const unsigned *cp , const unsigned *ri , const double *v , const double *D ,
                                                       const unsigned *indices , unsigned count , const double *b , double *x )
{
	const unsigned t = blockIdx.x*blockDim.x + threadIdx.x;
	if( t>=count ) return;
	const unsigned i = indices[t];
	double value = 0;
	for( unsigned k=cp[i] ; k<cp[i+1] ; k++ ) value += v[k] * x[ ri[k] ];
	x[i] += ( b[i] - value ) / D[i];
}

// y = A x for A given by rows (ptr over rows, idx = columns ascending); y = a - A x when a != nullptr
__global__ void __launch_bounds__( 256 ) KRowGather( const unsigned *rp , const unsigned *ci , const double *v , unsigned rows , const double *x , const double *a , double *y )
{
	const unsigned i = blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=rows ) return;
	double s = 0;
	for( unsigned k=rp[i] ; k<rp[i+1] ; k++ ) s += v[k] * x[ ci[k] ];
	y[i] = a ? a[i] - s : s;
}

__global__ void __launch_bounds__( 256 ) KAddTo( double *x , const double *t , unsigned n )
{
	const unsigned i = blockIdx.x*blockDim.x + threadIdx.x;
	if( i<n ) x[i] += t[i];
}

// dense y = Minv b (coarse solve)
__global__ void __launch_bounds__( 256 ) KDenseMv( const double *A , unsigned n , const double *x , double *y )
{
	const unsigned i = blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=n ) return;
	double s = 0;
	for( unsigned j=0 ; j<n ; j++ ) s += A[ (size_t)i*n+j ] * x[j];
	y[i] = s;
}

// products a_ik * b_kj of C = A B (A, B by rows), one thread per row of A
__global__ void __launch_bounds__( 256 ) KExpandCount( const unsigned *ap , const unsigned *ai , const unsigned *bp , unsigned rows , unsigned *cnt )
{
	const unsigned i = blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=rows ) return;
	unsigned c = 0;
	for( unsigned k=ap[i] ; k<ap[i+1] ; k++ ) c += bp[ ai[k]+1 ] - bp[ ai[k] ];
	cnt[i] = c;
}

__global__ void __launch_bounds__( 256 ) KExpand( const unsigned *ap , const unsigned *ai , const double *av , const unsigned *bp , const unsigned *bi , const double *bv ,
                                                  unsigned rows , unsigned long long cols , const unsigned *off , unsigned long long *key , double *prod )
{
	const unsigned i = blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=rows ) return;
	unsigned o = off[i];
	for( unsigned k=ap[i] ; k<ap[i+1] ; k++ )
	{
		const unsigned r = ai[k];
		const double a = av[k];
		for( unsigned q=bp[r] ; q<bp[r+1] ; q++ ){ key[o] = (unsigned long long)i*cols + bi[q] ; prod[o] = a * bv[q] ; o++; }
	}
}

// runs of equal keys (sorted, stable): the first position of a run sums it in order
__global__ void __launch_bounds__( 256 ) KRunHeads( const unsigned long long *key , unsigned n , unsigned *head )
{
	const unsigned j = blockIdx.x*blockDim.x + threadIdx.x;
	if( j<n ) head[j] = ( j==0 || key[j]!=key[j-1] ) ? 1u : 0u;
}

__global__ void __launch_bounds__( 256 ) KRunSums( const unsigned long long *key , const double *prod , const unsigned *order , unsigned n , const unsigned *head , const unsigned *pos ,
                                                   unsigned long long cols , unsigned *outRow , unsigned *outCol , double *outVal )
{
	const unsigned j = blockIdx.x*blockDim.x + threadIdx.x;
	if( j>=n || !head[j] ) return;
	double s = 0;
	unsigned t = j;
	for( ; t<n && key[t]==key[j] ; t++ ) s += prod[ order[t] ];
	const unsigned p = pos[j];
	outRow[p] = (unsigned)( key[j] / cols ) , outCol[p] = (unsigned)( key[j] % cols ) , outVal[p] = s;
}
#pragma clang fp contract(on)

__global__ void __launch_bounds__( 256 ) KIota( unsigned *p , unsigned n ){ const unsigned i = blockIdx.x*blockDim.x + threadIdx.x ; if( i<n ) p[i] = i; }
__global__ void __launch_bounds__( 256 ) KZeroU( unsigned *p , size_t n ){ const size_t i = (size_t)blockIdx.x*blockDim.x + threadIdx.x ; if( i<n ) p[i] = 0; }
__global__ void __launch_bounds__( 256 ) KZeroD( double *p , size_t n ){ const size_t i = (size_t)blockIdx.x*blockDim.x + threadIdx.x ; if( i<n ) p[i] = 0; }

// row pointer from sorted row indices: ptr[r] = first position with row >= r
__global__ void __launch_bounds__( 256 ) KRowPtr( const unsigned *row , unsigned nnz , unsigned rows , unsigned *ptr )
{
	const unsigned j = blockIdx.x*blockDim.x + threadIdx.x;
	if( j>nnz ) return;
	const unsigned lo = j==0 ? 0 : row[j-1]+1 , hi = j==nnz ? rows : row[j];
	for( unsigned r=lo ; r<=hi && r<=rows ; r++ ) ptr[r] = j;
}

__global__ void __launch_bounds__( 256 ) KDiagonal( const unsigned *cp , const unsigned *ri , const double *v , unsigned n , double *D )
{
	const unsigned i = blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=n ) return;
	for( unsigned k=cp[i] ; k<cp[i+1] ; k++ ) if( ri[k]==i ) D[i] = v[k];
}

__global__ void __launch_bounds__( 256 ) KDenseFromCsc( const unsigned *cp , const unsigned *ri , const double *v , unsigned n , double *A , double *I )
{
	const unsigned c = blockIdx.x*blockDim.x + threadIdx.x;
	if( c>=n ) return;
	for( unsigned k=cp[c] ; k<cp[c+1] ; k++ ) A[ (size_t)ri[k]*n+c ] = v[k];
	I[ (size_t)c*n+c ] = 1.;
}

// Gauss-Jordan inversion of a symmetric positive definite n x n matrix in one workgroup
__global__ void __launch_bounds__( 1024 ) KGaussJordan( double *A , double *I , unsigned n )
{
	for( unsigned k=0 ; k<n ; k++ )
	{
		const double piv = A[ (size_t)k*n+k ];
		__syncthreads();
		for( unsigned j=threadIdx.x ; j<n ; j+=blockDim.x ){ A[ (size_t)k*n+j ] /= piv ; I[ (size_t)k*n+j ] /= piv; }
		__syncthreads();
		for( size_t e=threadIdx.x ; e<(size_t)n*n ; e+=blockDim.x )
		{
			const unsigned r = (unsigned)( e / n ) , j = (unsigned)( e % n );
			if( r==k ) continue;
			const double f = A[ (size_t)r*n+k ];
			if( j!=k ) A[e] -= f*A[ (size_t)k*n+j ];
			I[e] -= f*I[ (size_t)k*n+j ];
		}
		__syncthreads();
		for( unsigned r=threadIdx.x ; r<n ; r+=blockDim.x ) if( r!=k ) A[ (size_t)r*n+k ] = 0;
		__syncthreads();
	}
}

// (row, col, val) triplets sorted by (row, col) -> compressed by rows
static void TripletsToRows( DevBuf< unsigned > &row , DevBuf< unsigned > &col , DevBuf< double > &val , unsigned nnz , unsigned rows , unsigned cols , DCompressed &out )
{
	out.outer = rows , out.inner = cols , out.nnz = nnz;
	out.ptr.alloc( rows+1 );
	KRowPtr<<< Grid( nnz+1 ) , 256 >>>( row.p , nnz , rows , out.ptr.p );
	out.idx = std::move( col ) , out.val = std::move( val );
}

// transpose of a compressed matrix (rows <-> columns), inner indices ascending
// (index bookkeeping of the setup: a counting sort on the host)
static void Transpose( const DCompressed &A , DCompressed &T )
{
	const size_t nnz = A.nnz;
	std::vector< unsigned > ap( A.outer+1 ) , ai( nnz );
	std::vector< double > av( nnz );
	HIPCHECK( hipMemcpy( ap.data() , A.ptr.p , ap.size()*sizeof(unsigned) , hipMemcpyDeviceToHost ) );
	if( nnz ) HIPCHECK( hipMemcpy( ai.data() , A.idx.p , nnz*sizeof(unsigned) , hipMemcpyDeviceToHost ) );
	if( nnz ) HIPCHECK( hipMemcpy( av.data() , A.val.p , nnz*sizeof(double) , hipMemcpyDeviceToHost ) );
	std::vector< unsigned > tp( A.inner+1 , 0 ) , ti( nnz );
	std::vector< double > tv( nnz );
	for( size_t k=0 ; k<nnz ; k++ ) tp[ ai[k]+1 ]++;
	for( size_t r=0 ; r<A.inner ; r++ ) tp[r+1] += tp[r];
	std::vector< unsigned > fill( tp.begin() , tp.end()-1 );
	for( size_t o=0 ; o<A.outer ; o++ ) for( unsigned k=ap[o] ; k<ap[o+1] ; k++ ){ const unsigned q = fill[ ai[k] ]++ ; ti[q] = (unsigned)o ; tv[q] = av[k]; }
	T.outer = A.inner , T.inner = A.outer , T.nnz = nnz;
	T.ptr.upload( tp.data() , tp.size() , nullptr );
	T.idx.alloc( nnz ) , T.val.alloc( nnz );
	if( nnz ){ T.idx.upload( ti.data() , nnz , nullptr ) ; T.val.upload( tv.data() , nnz , nullptr ); }
	HIPCHECK( hipStreamSynchronize( nullptr ) );
}

// C = A B, all by rows (A: rows x inner, B: inner x cols)
static void Multiply( const DCompressed &A , const DCompressed &B , DCompressed &C )
{
	const unsigned rows = (unsigned)A.outer;
	const unsigned long long cols = B.inner;
	DevBuf< unsigned > cnt , off;
	DevBuf< unsigned char > tmp;
	cnt.alloc( rows+1 ) , off.alloc( rows+1 );
	KZeroU<<< 1 , 256 >>>( cnt.p + rows , 1 );
	KExpandCount<<< Grid( rows ) , 256 >>>( A.ptr.p , A.idx.p , B.ptr.p , rows , cnt.p );
	size_t bytes = 0;
	HIPCHECK( hipcub::DeviceScan::ExclusiveSum( nullptr , bytes , cnt.p , off.p , (int)rows+1 ) );
	tmp.ensure( bytes );
	HIPCHECK( hipcub::DeviceScan::ExclusiveSum( tmp.p , bytes , cnt.p , off.p , (int)rows+1 ) );
	unsigned m = 0;
	HIPCHECK( hipMemcpy( &m , off.p + rows , sizeof(unsigned) , hipMemcpyDeviceToHost ) );
	C.outer = rows , C.inner = cols;
	if( !m ){ C.nnz = 0 ; C.ptr.alloc( rows+1 ) ; KZeroU<<< Grid( rows+1 ) , 256 >>>( C.ptr.p , rows+1 ) ; C.idx.alloc( 0 ) , C.val.alloc( 0 ) ; return; }
	DevBuf< unsigned long long > key , skey;
	DevBuf< double > prod;
	DevBuf< unsigned > ord , sord , head , pos;
	key.alloc( m ) , skey.alloc( m ) , prod.alloc( m ) , ord.alloc( m ) , sord.alloc( m ) , head.alloc( m+1 ) , pos.alloc( m+1 );
	KExpand<<< Grid( rows ) , 256 >>>( A.ptr.p , A.idx.p , A.val.p , B.ptr.p , B.idx.p , B.val.p , rows , cols , off.p , key.p , prod.p );
	KIota<<< Grid( m ) , 256 >>>( ord.p , m );
	bytes = 0;
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( nullptr , bytes , key.p , skey.p , ord.p , sord.p , (int)m ) );
	tmp.ensure( bytes );
	HIPCHECK( hipcub::DeviceRadixSort::SortPairs( tmp.p , bytes , key.p , skey.p , ord.p , sord.p , (int)m ) );
	KRunHeads<<< Grid( m ) , 256 >>>( skey.p , m , head.p );
	KZeroU<<< 1 , 256 >>>( head.p + m , 1 );
	bytes = 0;
	HIPCHECK( hipcub::DeviceScan::ExclusiveSum( nullptr , bytes , head.p , pos.p , (int)m+1 ) );
	tmp.ensure( bytes );
	HIPCHECK( hipcub::DeviceScan::ExclusiveSum( tmp.p , bytes , head.p , pos.p , (int)m+1 ) );
	unsigned nnz = 0;
	HIPCHECK( hipMemcpy( &nnz , pos.p + m , sizeof(unsigned) , hipMemcpyDeviceToHost ) );
	DevBuf< unsigned > r , c;
	DevBuf< double > v;
	r.alloc( nnz ) , c.alloc( nnz ) , v.alloc( nnz );
	KRunSums<<< Grid( m ) , 256 >>>( skey.p , prod.p , sord.p , m , head.p , pos.p , cols , r.p , c.p , v.p );
	TripletsToRows( r , c , v , nnz , rows , (unsigned)cols , C );
}

static void Upload( size_t outer , size_t inner , const unsigned *ptr , const unsigned *idx , const double *val , DCompressed &A )
{
	A.outer = outer , A.inner = inner , A.nnz = ptr[outer];
	for( size_t o=0 ; o<outer ; o++ ) if( ptr[o+1]<ptr[o] ) throw std::string( "column pointers decrease" );
	for( size_t k=0 ; k<A.nnz ; k++ ) if( idx[k]>=inner ) throw std::string( "row index out of range" );
	A.ptr.upload( ptr , outer+1 , nullptr ) , A.idx.upload( idx , A.nnz , nullptr ) , A.val.upload( val , A.nnz , nullptr );
}

struct Level
{
	unsigned n = 0;
	DCompressed cols , rows;        // M_l by columns (CSC) and by rows
	DevBuf< double > D , x , b , t;
	std::vector< unsigned > cstart;
	DevBuf< unsigned > cidx;
};

struct Solver
{
	int pNum = 0;
	std::vector< Level > L;          // 0 = coarsest, pNum = the given matrix
	std::vector< DCompressed > Prows , Pcols;   // P_i: level i -> i+1, by rows and by columns
	DevBuf< double > coarseInv;

	void relax( int l , const double *b , double *x , unsigned iters )
	{
		Level &lv = L[l];
		if( lv.cstart.empty() ) throw std::string( "no colour lists for level " + std::to_string( l ) );
		for( unsigned it=0 ; it<iters ; it++ ) for( size_t c=0 ; c+1<lv.cstart.size() ; c++ )
		{
			const unsigned cnt = lv.cstart[c+1]-lv.cstart[c];
			if( cnt ) KRelaxColour<<< Grid( cnt ) , 256 >>>( lv.cols.ptr.p , lv.cols.idx.p , lv.cols.val.p , lv.D.p , lv.cidx.p + lv.cstart[c] , cnt , b , x );
		}
	}

	// MG.h:146-186
	void vCycle( double *b , double *x , unsigned iters )
	{
		for( int l=pNum ; l>0 ; l-- )
		{
			double *bl = l==pNum ? b : L[l].b.p , *xl = l==pNum ? x : L[l].x.p;
			relax( l , bl , xl , iters );
			KRowGather<<< Grid( L[l].n ) , 256 >>>( L[l].rows.ptr.p , L[l].rows.idx.p , L[l].rows.val.p , L[l].n , xl , bl , L[l].t.p );
			// B_{l-1} = P^T r: rows of P^T are the columns of P
			KRowGather<<< Grid( L[l-1].n ) , 256 >>>( Pcols[l-1].ptr.p , Pcols[l-1].idx.p , Pcols[l-1].val.p , L[l-1].n , L[l].t.p , nullptr , L[l-1].b.p );
			KZeroD<<< Grid( L[l-1].n ) , 256 >>>( L[l-1].x.p , L[l-1].n );
		}
		double *b0 = pNum ? L[0].b.p : b , *x0 = pNum ? L[0].x.p : x;
		KDenseMv<<< Grid( L[0].n ) , 256 >>>( coarseInv.p , L[0].n , b0 , x0 );
		for( int l=1 ; l<=pNum ; l++ )
		{
			double *bl = l==pNum ? b : L[l].b.p , *xl = l==pNum ? x : L[l].x.p;
			KRowGather<<< Grid( L[l].n ) , 256 >>>( Prows[l-1].ptr.p , Prows[l-1].idx.p , Prows[l-1].val.p , L[l].n , L[l-1].x.p , nullptr , L[l].t.p );
			KAddTo<<< Grid( L[l].n ) , 256 >>>( xl , L[l].t.p , L[l].n );
			relax( l , bl , xl , iters );
		}
	}
};

// the level's matrix by columns given, by rows and its diagonal derived
static void FinishLevel( Level &lv )
{
	lv.n = (unsigned)lv.cols.outer;
	lv.D.alloc( lv.n ) , lv.x.alloc( lv.n ) , lv.b.alloc( lv.n ) , lv.t.alloc( lv.n );
	KZeroD<<< Grid( lv.n ) , 256 >>>( lv.D.p , lv.n );
	KDiagonal<<< Grid( lv.n ) , 256 >>>( lv.cols.ptr.p , lv.cols.idx.p , lv.cols.val.p , lv.n , lv.D.p );
	KZeroD<<< Grid( lv.n ) , 256 >>>( lv.x.p , lv.n );
}
} // namespace mg
} // namespace epr

struct epr_mg { epr::mg::Solver s; };

extern "C" {

int epr_mg_create( size_t n , const unsigned *colptr , const unsigned *rowidx , const double *values ,
                   int pNum , const size_t *pRows , const size_t *pCols ,
                   const unsigned *const *pColptr , const unsigned *const *pRowidx , const double *const *pValues , epr_mg **out )
{
	using namespace epr;
	using namespace epr::mg;
	if( !out || !colptr || !n || pNum<0 || ( pNum && ( !pRows || !pCols || !pColptr || !pRowidx || !pValues ) ) ){ epr::ThreadError() = "bad argument" ; return EPR_ERR_ARG; }
	*out = nullptr;
	epr_mg *mg = new epr_mg;
	const int rc = epr::ContextFree( [&]()
	{
		Solver &s = mg->s;
		s.pNum = pNum;
		s.L.resize( pNum+1 ) , s.Prows.resize( pNum ) , s.Pcols.resize( pNum );
		Upload( n , n , colptr , rowidx , values , s.L[pNum].cols );
		Transpose( s.L[pNum].cols , s.L[pNum].rows );
		FinishLevel( s.L[pNum] );
		size_t size = n;
		for( int i=pNum-1 ; i>=0 ; i-- )
		{
			if( pRows[i]!=size ) throw std::string( "prolongation " + std::to_string( i ) + " has the wrong number of rows" );
			Upload( pCols[i] , pRows[i] , pColptr[i] , pRowidx[i] , pValues[i] , s.Pcols[i] );
			Transpose( s.Pcols[i] , s.Prows[i] );
			// M_i = ( P^T M_{i+1} ) P_i, by rows; rows of P^T are the columns of P
			DCompressed T;
			Multiply( s.Pcols[i] , s.L[i+1].rows , T );
			Multiply( T , s.Prows[i] , s.L[i].rows );
			Transpose( s.L[i].rows , s.L[i].cols );
			FinishLevel( s.L[i] );
			size = pCols[i];
		}
		const unsigned n0 = s.L[0].n;
		if( n0>4096 ) throw std::string( "coarsest level too large for the dense solve (" + std::to_string( n0 ) + " > 4096)" );
		DevBuf< double > A;
		A.alloc( (size_t)n0*n0 ) , s.coarseInv.alloc( (size_t)n0*n0 );
		KZeroD<<< Grid( (long long)n0*n0 ) , 256 >>>( A.p , (size_t)n0*n0 );
		KZeroD<<< Grid( (long long)n0*n0 ) , 256 >>>( s.coarseInv.p , (size_t)n0*n0 );
		KDenseFromCsc<<< Grid( n0 ) , 256 >>>( s.L[0].cols.ptr.p , s.L[0].cols.idx.p , s.L[0].cols.val.p , n0 , A.p , s.coarseInv.p );
		KGaussJordan<<< 1 , 1024 >>>( A.p , s.coarseInv.p , n0 );
		HIPCHECK( hipDeviceSynchronize() );
	} );
	if( rc ){ delete mg ; return rc; }
	*out = mg;
	return EPR_OK;
}

int epr_mg_num_levels( const epr_mg *mg ){ return mg ? mg->s.pNum+1 : 0; }
size_t epr_mg_level_size( const epr_mg *mg , int level ){ return ( mg && level>=0 && level<=mg->s.pNum ) ? mg->s.L[level].n : 0; }

int epr_mg_set_colours( epr_mg *mg , int level , int nColours , const unsigned *starts , const unsigned *indices )
{
	if( !mg || level<0 || level>mg->s.pNum || nColours<1 || !starts || !indices ){ epr::ThreadError() = "bad argument" ; return EPR_ERR_ARG; }
	return epr::ContextFree( [&]()
	{
		epr::mg::Level &lv = mg->s.L[level];
		const unsigned total = starts[nColours];
		for( int c=0 ; c<nColours ; c++ ) if( starts[c+1]<starts[c] ) throw std::string( "colour starts decrease" );
		for( unsigned k=0 ; k<total ; k++ ) if( indices[k]>=lv.n ) throw std::string( "colour index out of range" );
		lv.cstart.assign( starts , starts+nColours+1 );
		lv.cidx.upload( indices , total , nullptr );
		HIPCHECK( hipStreamSynchronize( nullptr ) );
	} );
}

int epr_mg_vcycle( epr_mg *mg , const double *b , double *x , unsigned iters )
{
	if( !mg || !b || !x ){ epr::ThreadError() = "bad argument" ; return EPR_ERR_ARG; }
	return epr::ContextFree( [&]()
	{
		epr::mg::Level &top = mg->s.L[ mg->s.pNum ];
		top.b.upload( b , top.n , nullptr ) , top.x.upload( x , top.n , nullptr );
		mg->s.vCycle( top.b.p , top.x.p , iters );
		top.x.download( x , top.n , nullptr );
		HIPCHECK( hipStreamSynchronize( nullptr ) );
	} );
}

int epr_mg_relax( epr_mg *mg , int level , const double *b , double *x , unsigned iters )
{
	if( !mg || level<0 || level>mg->s.pNum || !b || !x ){ epr::ThreadError() = "bad argument" ; return EPR_ERR_ARG; }
	return epr::ContextFree( [&]()
	{
		epr::mg::Level &lv = mg->s.L[level];
		epr::DevBuf< double > db , dx;
		db.upload( b , lv.n , nullptr ) , dx.upload( x , lv.n , nullptr );
		mg->s.relax( level , db.p , dx.p , iters );
		dx.download( x , lv.n , nullptr );
		HIPCHECK( hipStreamSynchronize( nullptr ) );
	} );
}

void epr_mg_destroy( epr_mg *mg ){ delete mg; }

}
