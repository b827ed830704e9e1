// Misha's square matrices (Misha/Geometry.inl) as the reference computes
// with them: the product, transpose, Gauss-Jordan inverse and the affine map
// of a point, with the same operations in the same order, so transforms
// written to files (the .tree's voxel-to-world map) and samples mapped
// through them are the reference's bit for bit.  Host code.
#pragma once
#include <cmath>
#include <cstring>
#include <string>

namespace epr
{
namespace sample
{
// Misha's square matrix: c[col][row], m(c,r) = c[c][r]
template< int N >
struct Mat
{
	double c[N][N];
	Mat( void ){ memset( c , 0 , sizeof( c ) ); }
	static Mat Identity( void ){ Mat m ; for( int i=0 ; i<N ; i++ ) m.c[i][i] = 1 ; return m; }
	double &operator()( int col , int row ){ return c[col][row]; }
	double operator()( int col , int row ) const { return c[col][row]; }
	// Geometry.inl:217-226: (A B)[i][j] = sum_k B[i][k] A[k][j]
	Mat operator*( const Mat &m ) const
	{
		Mat n;
		for( int i=0 ; i<N ; i++ ) for( int j=0 ; j<N ; j++ ) for( int k=0 ; k<N ; k++ ) n.c[i][j] += m.c[i][k]*c[k][j];
		return n;
	}
	Mat transpose( void ) const { Mat o ; for( int i=0 ; i<N ; i++ ) for( int j=0 ; j<N ; j++ ) o.c[j][i] = c[i][j] ; return o; }
	// Geometry.inl:277-304 (Gauss-Jordan with the row of the largest |entry| of column i)
	Mat inverse( void ) const
	{
		Mat xf = *this , ixf = Identity();
		for( int i=0 ; i<N ; i++ )
		{
			int p = i;
			double v = fabs( xf( i , i ) );
			for( int j=i+1 ; j<N ; j++ ) if( fabs( xf( i , j ) )>v ) p = j , v = fabs( xf( i , j ) );
			if( v==0. ) throw std::string( "singular unit-cube transform" );
			Mat t = Identity();
			const double scl = 1./xf( i , p );
			for( int j=0 ; j<N ; j++ ) t( p , j ) = -xf( i , j )*scl;
			t( i , p ) = 1.;
			t( p , p ) = -xf( i , i )*scl;
			t( i , i ) = 0.;
			t( p , i ) = scl;
			xf = t*xf , ixf = t*ixf;
		}
		return ixf;
	}
};
// Geometry.inl:412-424: the affine map of an (N+1)x(N+1) matrix applied to a point
template< int N >
void Apply( const Mat< N+1 > &m , const double *v , double *out )
{
	double scale = 1;
	for( int i=0 ; i<N ; i++ )
	{
		double o = 0;
		for( int j=0 ; j<N ; j++ ) o += v[j]*m.c[j][i];
		o += m.c[N][i];
		scale += m.c[i][N];
		out[i] = o;
	}
	for( int i=0 ; i<N ; i++ ) out[i] /= scale;
}
} // namespace sample
} // namespace epr
