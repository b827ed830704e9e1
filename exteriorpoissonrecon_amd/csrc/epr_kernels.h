// HIP kernels of the dense ExteriorPoissonRecon solve path (gfx950).
//
// Layout in HBM (one reconstruction, level l has resolution r = 2^l):
//   * coefficients x,y: N = (r+1)^Dim doubles each, function index with
//     coordinate 0 most significant (reference Hat.scalars.inl:475-481);
//   * R and B: N x 3^Dim doubles, column f holds M(f+o, f) for the window
//     offset o in {-1,0,1}^Dim (window index base 3, first coordinate most
//     significant, Misha/Window.h:79-83);
//   * K: 3^Dim boundary cases x 3^Dim dense rows of the stiffness stencil;
//   * per-element target field S: r^Dim x C(Dim,2) doubles, element index with
//     coordinate 0 fastest (Hat.scalars.inl:455-473).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>
#include "epr_wedge_tables.h"

namespace epr
{
template< int Dim > struct DimC;
template<> struct DimC< 3 > { static constexpr int W = 27 , C = 8 , CD = 3 , NDNI = 13; };
template<> struct DimC< 4 > { static constexpr int W = 81 , C = 16 , CD = 6 , NDNI = 21; };

constexpr int MAXT = 16;     // tree depths supported (depth <= 14)

template< int Dim > __device__ __host__ constexpr int Bit( int c , int d ){ return ( c>>( Dim-1-d ) ) & 1; }

template< int Dim > __device__ __forceinline__ double WT( int j , int p ){ if constexpr( Dim==3 ) return wedge3::WT[j][p]; else return wedge4::WT[j][p]; }
template< int Dim > __device__ __forceinline__ double WA( int j , int k ){ if constexpr( Dim==3 ) return wedge3::WA[j][k]; else return wedge4::WA[j][k]; }
template< int Dim > __device__ __forceinline__ int PK( int p ){ if constexpr( Dim==3 ) return wedge3::PK[p]; else return wedge4::PK[p]; }
template< int Dim > __device__ __forceinline__ int PL( int p ){ if constexpr( Dim==3 ) return wedge3::PL[p]; else return wedge4::PL[p]; }

// Column layout of R and B: the columns (3^Dim values each) of the nodes of
// one colour (parity class, ExteriorPoisson.h:161-169) are contiguous, colour
// after colour, in ColourSet order within a colour -- a colour phase of the
// sweep reads one contiguous block.
template< int Dim >
struct ColLayout
{
	int res;
	long long cstart[ (1<<Dim)+1 ];
	__device__ __host__ long long col( const int *F ) const
	{
		int c = 0;
		long long k = 0;
		for( int d=0 ; d<Dim ; d++ ) c |= ( F[d]&1 )<<d;
		for( int d=0 ; d<Dim ; d++ ){ const int cnt = ( res - ( ( c>>d )&1 ) )/2 + 1 ; k = k*cnt + ( F[d]>>1 ); }
		return cstart[c] + k;
	}
	void set( int r )
	{
		res = r , cstart[0] = 0;
		for( int c=0 ; c<(1<<Dim) ; c++ )
		{
			long long t = 1;
			for( int d=0 ; d<Dim ; d++ ) t *= ( r - ( ( c>>d )&1 ) )/2 + 1;
			cstart[c+1] = cstart[c] + t;
		}
	}
};

// One level as the node kernels see it.
template< int Dim , typename Real = double >
struct LevelView
{
	int res;
	long long N;
	Real *x , *y;
	const Real *R , *B , *K;
	Real sW;      // wedge scale res^(4-Dim)
	Real dW;      // Dirichlet weight
	long long stride[Dim];
	ColLayout< Dim > lay;   // column position of a node in R and B
	const unsigned char *fmask;   // adapted levels: 1 = active (updated by the sweeps); null: every node is active
	int separate;                 // --separate: x then y per node (SeparateGaussSeidelSolve) instead of the paired step
};

enum { MODE_GS = 0 , MODE_QFIT = 1 , MODE_ENERGY = 2 };

// Compile-time access to the corner-0 wedge tables of a dimension.
template< int Dim > struct WTab;
template<> struct WTab< 3 >
{
	static constexpr int C = wedge3::C , P = wedge3::P;
	static constexpr double T( int j , int p ){ return wedge3::WT[j][p]; }
	static constexpr double A( int j , int k ){ return wedge3::WA[j][k]; }
	static constexpr int K( int p ){ return wedge3::PK[p]; }
	static constexpr int L( int p ){ return wedge3::PL[p]; }
};
template<> struct WTab< 4 >
{
	static constexpr int C = wedge4::C , P = wedge4::P;
	static constexpr double T( int j , int p ){ return wedge4::WT[j][p]; }
	static constexpr double A( int j , int k ){ return wedge4::WA[j][k]; }
	static constexpr int K( int p ){ return wedge4::PK[p]; }
	static constexpr int L( int p ){ return wedge4::PL[p]; }
};

// z[j] += T(j,p) * w for the non-zero entries of column p (expanded at compile time)
template< int Dim , int p , typename Real , int ... js >
__device__ __forceinline__ void WedgeColumn( Real *z , Real w , std::integer_sequence< int , js... > )
{
	( [&]{ if constexpr( WTab< Dim >::T( js , p )!=0.0 ) z[js] = fma( (Real)WTab< Dim >::T( js , p ) , w , z[js] ); }() , ... );
}
template< int Dim , typename Real , int ... ps >
__device__ __forceinline__ void WedgeContract( const Real *xp , const Real *yp , Real *z , std::integer_sequence< int , ps... > )
{
	( [&]{
		constexpr int k = WTab< Dim >::K( ps ) , l = WTab< Dim >::L( ps );
		const Real w = xp[k]*yp[l] - xp[l]*yp[k];
		WedgeColumn< Dim , ps >( z , w , std::make_integer_sequence< int , WTab< Dim >::C-1 >{} );
	}() , ... );
}
// u[j] = sum_k A(j,k) yp[k+1], v[j] = sum_k A(j,k) xp[k+1]
template< int Dim , int j , typename Real , int ... ks >
__device__ __forceinline__ void QuadRow( const Real *xp , const Real *yp , Real &u , Real &v , std::integer_sequence< int , ks... > )
{
	( [&]{ if constexpr( WTab< Dim >::A( j , ks )!=0.0 ){ u = fma( (Real)WTab< Dim >::A( j , ks ) , yp[ks+1] , u ) ; v = fma( (Real)WTab< Dim >::A( j , ks ) , xp[ks+1] , v ); } }() , ... );
}
template< int Dim , typename Real , int ... js >
__device__ __forceinline__ void QuadForms( const Real *xp , const Real *yp , const Real *z , Real &a20 , Real &a02 , Real &a11 , Real &b10 , Real &b01 , std::integer_sequence< int , js... > )
{
	( [&]{
		Real u = 0 , v = 0;
		QuadRow< Dim , js >( xp , yp , u , v , std::make_integer_sequence< int , WTab< Dim >::C-1 >{} );
		a20 = fma( yp[js+1] , u , a20 );
		a02 = fma( xp[js+1] , v , a02 );
		a11 = fma( xp[js+1] , u , a11 );
		b10 = fma( yp[js+1] , z[js] , b10 );
		b01 = fma( xp[js+1] , z[js] , b01 );
	}() , ... );
}

// Wedge-energy terms of one element for the node at its corner 0 after the
// corner relabelling c -> c^beta (the element integrals are invariant under
// reflections).  xp/yp hold the relabelled corner coefficients.  Returns the
// unscaled coefficients of E(x+s e_i, y+t e_i) restricted to the element.
// The quadratic terms follow SystemEnergy.h:623-639, the cubic terms :616-621.
template< int Dim , typename Real >
__device__ __forceinline__ void ElementWedge( const Real *xp , const Real *yp , Real &q20 , Real &q02 , Real &q11 , Real &q10 , Real &q01 , Real &zy )
{
	constexpr int C = 1<<Dim , P = ( C*(C-1) )/2;
	Real z[C-1];
#pragma unroll
	for( int j=0 ; j<C-1 ; j++ ) z[j] = 0;
	WedgeContract< Dim , Real >( xp , yp , z , std::make_integer_sequence< int , P >{} );
	Real a20 = 0 , a02 = 0 , a11 = 0 , b10 = 0 , b01 = 0;
	QuadForms< Dim , Real >( xp , yp , z , a20 , a02 , a11 , b10 , b01 , std::make_integer_sequence< int , C-1 >{} );
	q20 = a20 , q02 = a02 , q11 = Real( -2 )*a11 , q10 = Real( 2 )*b10 , q01 = Real( -2 )*b01 , zy = b10;
}

// The wedge energy of one element for one coordinate pair (i, j):
// (1/2) int_e F_ij^2, F_ij = d_i x d_j y - d_j x d_i y, at unit resolution,
// for multilinear x, y with corner values xc, yc whose index bit q stands for
// coordinate order[q], order = (i, j, the other coordinates ascending).  The
// element's wedge energy (the quartic form of the corner-0 tables, summed
// over the corners) is the sum of these over the pairs: the reference's
// 2-form inner product of grad x ^ grad y with itself is (1/2) sum_{i<j}
// F_ij^2 (Include/Hat.inl:32-37, 358-370).  d_i x is multilinear in the other
// coordinates, so F_ij = sum_{a,b,m} c[a,b,m] psi_a(t_i) psi_b(t_j) prod_q
// e_{m_q}(t_q) in the hat basis psi_0 = 1-t, psi_1 = t and the products
// e_0 = psi_0^2, e_1 = psi_0 psi_1, e_2 = psi_1^2 for the other coordinates;
// int F^2 = c^T (G (x) G (x) H (x) ...) c with the 1D Gram matrices
// G = [[1/3,1/6],[1/6,1/3]] and H_{mm'} = (m+m')!(4-m-m')!/5!, evaluated as
// |U c|^2 with U the Kronecker product of their upper Cholesky factors,
// applied in place along each axis.  About 4 3^(Dim-2) (8 + 2 Dim) flops per
// pair instead of the corner-0 contractions of every corner.
constexpr double UG00 = 0.5773502691896257 , UG01 = 0.28867513459481287 , UG11 = 0.5;
constexpr double UH00 = 0.4472135954999579 , UH01 = 0.1118033988749895 , UH02 = 0.07453559924999299;
constexpr double UH11 = 0.14433756729740643 , UH12 = 0.2886751345948129 , UH22 = 0.3333333333333333;
template< int Dim >
__device__ __forceinline__ double PairWedgeEnergy( const double *xc , const double *yc )
{
	constexpr int NO = Dim-2 , R = 1<<NO , M3 = NO==1 ? 3 : 9;
	// differences along i (indexed by b_j and the other bits) and along j (by b_i and the others)
	double dxi[2][R] , dyi[2][R] , dxj[2][R] , dyj[2][R];
#pragma unroll
	for( int b=0 ; b<2 ; b++ )
#pragma unroll
		for( int r=0 ; r<R ; r++ )
		{
			dxi[b][r] = xc[ 1 | (b<<1) | (r<<2) ] - xc[ (b<<1) | (r<<2) ];
			dyi[b][r] = yc[ 1 | (b<<1) | (r<<2) ] - yc[ (b<<1) | (r<<2) ];
			dxj[b][r] = xc[ b | 2 | (r<<2) ] - xc[ b | (r<<2) ];
			dyj[b][r] = yc[ b | 2 | (r<<2) ] - yc[ b | (r<<2) ];
		}
	// c[(a + 2b) M3 + sum_q m_q 3^q]
	double c[4*M3];
#pragma unroll
	for( int k=0 ; k<4*M3 ; k++ ) c[k] = 0;
#pragma unroll
	for( int r=0 ; r<R ; r++ )
#pragma unroll
		for( int rp=0 ; rp<R ; rp++ )
		{
			int m = 0 , p3 = 1;
#pragma unroll
			for( int q=0 ; q<NO ; q++ ){ m += ( ( (r>>q)&1 ) + ( (rp>>q)&1 ) )*p3 ; p3 *= 3; }
#pragma unroll
			for( int a=0 ; a<2 ; a++ )
#pragma unroll
				for( int b=0 ; b<2 ; b++ ) c[ (a+2*b)*M3 + m ] += dxi[b][r]*dyj[a][rp] - dxj[a][r]*dyi[b][rp];
		}
	// U along a (stride M3), b (stride 2 M3), then each other coordinate's digit
#pragma unroll
	for( int k=0 ; k<M3 ; k++ )
#pragma unroll
		for( int b=0 ; b<2 ; b++ )
		{
			double &v0 = c[ (2*b)*M3 + k ] , &v1 = c[ (1+2*b)*M3 + k ];
			v0 = UG00*v0 + UG01*v1 , v1 = UG11*v1;
		}
#pragma unroll
	for( int k=0 ; k<M3 ; k++ )
#pragma unroll
		for( int a=0 ; a<2 ; a++ )
		{
			double &v0 = c[ a*M3 + k ] , &v1 = c[ (a+2)*M3 + k ];
			v0 = UG00*v0 + UG01*v1 , v1 = UG11*v1;
		}
#pragma unroll
	for( int q=0 , p3=1 ; q<NO ; q++ , p3*=3 )
#pragma unroll
		for( int k=0 ; k<4*M3 ; k++ )
			if( ( ( k%M3 )/p3 )%3==0 )
			{
				double &v0 = c[k] , &v1 = c[k+p3] , &v2 = c[k+2*p3];
				v0 = UH00*v0 + UH01*v1 + UH02*v2 , v1 = UH11*v1 + UH12*v2 , v2 = UH22*v2;
			}
	double s = 0;
#pragma unroll
	for( int k=0 ; k<4*M3 ; k++ ) s += c[k]*c[k];
	return 0.5*s;
}

// The omega=1/2 block-Newton step of ExteriorPoissonSolvers.h:173-181 with the
// 2x2 inverse of Misha/Geometry.inl:319-331.
template< typename Real >
__device__ __forceinline__ bool NodeStep( Real q20 , Real q02 , Real q11 , Real q10 , Real q01 , Real &dx , Real &dy )
{
	const Real H00 = 2*q20 , H11 = 2*q02 , H01 = q11 , H10 = q11;
	const Real det = H00*H11 - H01*H10;
	if( !( fabs( det )>Real( 1e-20 ) ) ) return false;
	const Real d = Real( 1 ) / det;
	const Real m00 = -( H11*d ) , m11 = -( H00*d ) , m01 = -( (-H01)*d ) , m10 = -( (-H10)*d );
	Real o0 = q10*m00 , o1 = q10*m01;
	o0 = o0 + q01*m10 , o1 = o1 + q01*m11;
	dx = o0*Real( 0.5 ) , dy = o1*Real( 0.5 );
	return true;
}

// The --separate step of ExteriorPoissonSolvers.h:198-238: x with y fixed,
// then y with the new x.  The reference guards each with fabs( H(k,k)>1e-20 ),
// the absolute value of a comparison, i.e. H(k,k)>1e-20; the update is always
// added (zero when a guard fails).
template< typename Real >
__device__ __forceinline__ void NodeStepSeparate( Real q20 , Real q02 , Real q11 , Real q10 , Real q01 , Real &dx , Real &dy )
{
	const Real H00 = 2*q20 , H11 = 2*q02 , H01 = q11 , H10 = q11;
	Real p0 = 0 , p1 = 0;
	if( H00>Real( 1e-20 ) ) p0 = -( ( H10 + H01 )*p1 + q10 ) / ( 2*H00 );
	if( H11>Real( 1e-20 ) ) p1 = -( ( H10 + H01 )*p0 + q01 ) / ( 2*H11 );
	dx = p0 , dy = p1;
}

// The step a level view selects: paired (false: leave the node) or separate.
template< int Dim , typename Real >
__device__ __forceinline__ bool LevelStep( const LevelView< Dim , Real > &L , const Real q[5] , Real &dx , Real &dy )
{
	if( L.separate ){ NodeStepSeparate( q[0] , q[1] , q[2] , q[3] , q[4] , dx , dy ) ; return true; }
	return NodeStep( q[0] , q[1] , q[2] , q[3] , q[4] , dx , dy );
}

// v + v[lane'] for the DPP lane pattern CTRL within a row of 16 lanes; a source
// lane outside the row or inactive contributes 0 (bound_ctrl: no old value to
// initialise, and the FP32 add takes the DPP operand directly)
template< int CTRL , typename Real >
__device__ __forceinline__ Real DppAddT( Real v ){ return v + __builtin_amdgcn_mov_dpp( v , CTRL , 0xF , 0xF , true ); }
// sum over an aligned group of G = 2, 4, 8 or 16 lanes: xor 1, xor 2, then the
// half-row and row mirrors (after the quad stage every lane of a quad holds
// the quad's sum, so pairing lane i with 7-i equals pairing it with i^4, and
// likewise 15-i with i^8: the same sums in the same order as an xor butterfly)
template< int G , typename Real >
__device__ __forceinline__ Real GroupSum( Real v )
{
	static_assert( G==2 || G==4 || G==8 || G==16 , "group width" );
	v = DppAddT< 0xB1 >( v );
	if constexpr( G>=4 ) v = DppAddT< 0x4E >( v );
	if constexpr( G>=8 ) v = DppAddT< 0x141 >( v );
	if constexpr( G>=16 ) v = DppAddT< 0x140 >( v );
	return v;
}

// Processes node i (multi-index F) with the C lanes of its group: lane = one
// incident element (wedge terms) + a share of the 3^Dim window (R, K, B terms).
// All lanes return the group totals.
template< typename T > struct NoDeduce { using type = T; };
template< int Dim , int MODE , bool PREFETCH = true , typename Real = double , typename Idx = long long >
__device__ __forceinline__ void NodeGroup( const LevelView< Dim , Real > &L , const int *F , typename NoDeduce< Idx >::type i , typename NoDeduce< Idx >::type col , int lane , Real acc[5] )
{
	constexpr int C = DimC< Dim >::C , W = DimC< Dim >::W;
	Real q20 = 0 , q02 = 0 , q11 = 0 , q10 = 0 , q01 = 0;
	Real eW = 0 , eB = 0 , eR = 0 , eK = 0;
	// ---- the lane's share of the 3^Dim window: issue its loads first
	constexpr int NWL = ( W + C - 1 ) / C;
	int cse = 0;
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) cse = cse*3 + ( F[d]<=0 ? 0 : ( F[d]>=L.res ? 2 : 1 ) );
	Real wx[NWL] , wy[NWL] , wR[NWL] , wB[NWL] , wK[NWL];
#define EPR_LOAD_WINDOW( m ) \
	{ \
		const int w_ = lane + (m)*C; \
		int t_ = w_; \
		Idx r_ = i; \
		bool in_ = w_<W; \
		_Pragma( "unroll" ) \
		for( int d=Dim-1 ; d>=0 ; d-- ) \
		{ \
			const int o_ = t_%3 - 1 ; t_ /= 3; \
			const int g_ = F[d]+o_; \
			in_ = in_ && g_>=0 && g_<=L.res; \
			r_ += o_*(Idx)L.stride[d]; \
		} \
		if( in_ ) wx[m] = L.x[r_] , wy[m] = L.y[r_] , wR[m] = L.R[ col*W+w_ ] , wB[m] = L.B[ col*W+w_ ] , wK[m] = L.K[ cse*W + w_ ]; \
		else      wx[m] = wy[m] = wR[m] = wB[m] = wK[m] = 0; \
	}
	if constexpr( PREFETCH )
	{
#pragma unroll
		for( int m=0 ; m<NWL ; m++ ) EPR_LOAD_WINDOW( m )
	}
	const Real xi = L.x[i] , yi = L.y[i];
	// ---- the incident element of this lane (the energy's wedge term is KEnergyWedge's)
	if constexpr( MODE!=MODE_ENERGY )
	{
		bool valid = true;
		Idx sig[Dim];
#pragma unroll
		for( int d=0 ; d<Dim ; d++ )
		{
			const int b = Bit< Dim >( lane , d );
			const int e = F[d] - 1 + b;
			valid = valid && e>=0 && e<L.res;
			sig[d] = b ? (Idx)L.stride[d] : -(Idx)L.stride[d];
		}
		if( valid )
		{
			Real xp[C] , yp[C];
#pragma unroll
			for( int c=0 ; c<C ; c++ )
			{
				Idx o = i;
#pragma unroll
				for( int d=0 ; d<Dim ; d++ ) if( Bit< Dim >( c , d ) ) o += sig[d];
				xp[c] = L.x[o] , yp[c] = L.y[o];
			}
			Real a20 , a02 , a11 , a10 , a01 , zy;
			ElementWedge< Dim , Real >( xp , yp , a20 , a02 , a11 , a10 , a01 , zy );
			q20 = a20*L.sW , q02 = a02*L.sW , q11 = a11*L.sW , q10 = a10*L.sW , q01 = a01*L.sW;
			(void)zy;
		}
	}
#pragma unroll
	for( int m=0 ; m<NWL ; m++ )
	{
		if constexpr( !PREFETCH ) EPR_LOAD_WINDOW( m )
		const int w = lane + m*C;
		const Real xr = wx[m] , yr = wy[m] , Rv = wR[m] , Bv = wB[m] , Kv = wK[m];
		if constexpr( MODE==MODE_ENERGY )
		{
			eR += xi*Rv*xr + yi*Rv*yr;
			eK += xi*Kv*xr + yi*Kv*yr;
			eB += xi*Bv*yr;
		}
		else
		{
			const Real Dv = Kv*L.dW;
			if( w==W/2 ){ q20 += Rv + Dv ; q02 += Rv + Dv; }
			q10 += Real( 2 )*Rv*xr + Real( 2 )*Dv*xr;
			q01 += Real( 2 )*Rv*yr + Real( 2 )*Dv*yr;
			if( w!=W/2 ){ q10 -= Real( 2 )*Bv*yr ; q01 += Real( 2 )*Bv*xr; }
		}
	}
#undef EPR_LOAD_WINDOW
	if constexpr( MODE==MODE_ENERGY ){ q20 = eW , q02 = eB , q11 = eR , q10 = eK , q01 = 0; }
	// ---- group reduction (DPP inside the C-lane group)
	q20 = GroupSum< C >( q20 ) , q02 = GroupSum< C >( q02 ) , q11 = GroupSum< C >( q11 ) , q10 = GroupSum< C >( q10 ) , q01 = GroupSum< C >( q01 );
	acc[0] = q20 , acc[1] = q02 , acc[2] = q11 , acc[3] = q10 , acc[4] = q01;
}

// Unsigned division by a fixed divisor d >= 1 of numbers below 2^31: the
// high half of n * ceil(2^32/d), one down when that overshoots (exact).
struct FastDiv
{
	unsigned int d , m;
	__host__ void set( unsigned int _d ){ d = _d , m = (unsigned int)( ( 0xFFFFFFFFull )/_d + 1 ); }
	__device__ __forceinline__ unsigned int div( unsigned int n ) const { unsigned int q = d==1 ? n : __umulhi( n , m ) ; if( q*d>n ) q-- ; return q; }
};

// Workgroups are dispatched to the 8 XCDs round robin (workgroup b to XCD
// b mod 8), each XCD with its own L2.  This maps the workgroups of one XCD to
// one contiguous range of the work, so neighbouring rows and planes -- which
// share their x, y cache lines -- stay in the same L2.  A bijection of [0, n).
__device__ __forceinline__ unsigned XcdBlock( unsigned b , unsigned n )
{
	constexpr unsigned X = 8;
	const unsigned q = n / X , r = n % X , xcd = b % X , local = b / X;
	return xcd<r ? xcd*( q+1 ) + local : r*( q+1 ) + ( xcd-r )*q + local;
}

// Per-dimension counts of the nodes of a colour and the k-th node's index.
template< int Dim >
struct ColourSet
{
	int cnt[Dim] , par[Dim];
	long long total;
	__device__ __host__ void set( int res , int colour )
	{
		total = 1;
		for( int d=0 ; d<Dim ; d++ )
		{
			par[d] = ( colour>>d ) & 1;   // bit d of the colour is the parity of coordinate d (ExteriorPoisson.h:161-169)
			cnt[d] = ( res - par[d] )/2 + 1;
			total *= cnt[d];
		}
	}
	__device__ __forceinline__ void node( long long k64 , int *F ) const
	{
		unsigned int k = (unsigned int)k64;   // colour sets hold fewer than 2^32 nodes
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ const unsigned int c = (unsigned int)cnt[d] , q = k / c ; F[d] = 2*(int)( k - q*c ) + par[d] ; k = q; }
	}
};

template< int Dim > __device__ __forceinline__ long long FlatIndex( const int *F , const long long *stride ){ long long i = 0 ; for( int d=0 ; d<Dim ; d++ ) i += F[d]*stride[d] ; return i; }
template< int Dim > __device__ __forceinline__ void Unflatten( long long i , int res , int *F ){ for( int d=Dim-1 ; d>=0 ; d-- ){ F[d] = (int)( i % (res+1) ) ; i /= res+1; } }

// One colour phase of the paired Gauss-Seidel sweep (one launch per colour);
// block b of the phase.
template< int Dim , typename Real = double , typename Idx = long long >
__device__ __forceinline__ void GsPhase( const LevelView< Dim , Real > &L , const ColourSet< Dim > &cs , long long colBase , unsigned b )
{
	constexpr int C = DimC< Dim >::C;
	const long long g = (long long)b*blockDim.x + threadIdx.x;
	const long long k = g / C;
	const int lane = (int)( g % C );
	if( k>=cs.total ) return;
	int F[Dim];
	cs.node( k , F );
	const Idx i = (Idx)FlatIndex< Dim >( F , L.stride );
	if( L.fmask && L.fmask[i]!=1 ) return;   // the whole group leaves together
	Real q[5];
	NodeGroup< Dim , MODE_GS , Dim==3 , Real , Idx >( L , F , i , (Idx)( colBase+k ) , lane , q );
	if( lane==0 )
	{
		Real dx , dy;
		if( LevelStep< Dim , Real >( L , q , dx , dy ) ) L.x[i] += dx , L.y[i] += dy;
	}
}
// Idx = int when N 3^Dim < 2^31 (every index of the level fits), else long long
template< int Dim , typename Real = double , typename Idx = long long >
__global__ void __launch_bounds__( 256 ) KGsPhase( LevelView< Dim , Real > L , ColourSet< Dim > cs , long long colBase ){ GsPhase< Dim , Real , Idx >( L , cs , colBase , XcdBlock( blockIdx.x , gridDim.x ) ); }

// ---- The colour phase of a dense level with per-lane constants: the same
// node groups and arithmetic as GsPhase, with the lane's window and element
// offsets as precomputed byte offsets and boundary masks, and x, y read at
// 32-bit byte offsets from the level's base pointers (exact while
// N sizeof(Real) < 2^32: offsets relative to the node are added modulo 2^32).
// Several nodes per thread (amortising the lane constants) measured slower.

// A colour's node decomposition with fixed-divisor divisions.
template< int Dim >
struct ColourDiv
{
	FastDiv cnt[Dim];
	int par[Dim];
	unsigned int total;
	__host__ void set( const ColourSet< Dim > &cs ){ for( int d=0 ; d<Dim ; d++ ) cnt[d].set( (unsigned int)cs.cnt[d] ) , par[d] = cs.par[d] ; total = (unsigned int)cs.total; }
	__device__ __forceinline__ void node( unsigned int k , int *F ) const
	{
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ const unsigned int q = cnt[d].div( k ) ; F[d] = 2*(int)( k - q*cnt[d].d ) + par[d] ; k = q; }
	}
};

// The lane's share of the 3^Dim window (entries lane + m C) and its incident
// element: byte offsets relative to the node and the coordinates (bit d) in
// which they step below / above it.
template< int Dim , typename Real >
struct GsLane
{
	static constexpr int C = DimC< Dim >::C , W = DimC< Dim >::W , NWL = ( W + C - 1 )/C;
	unsigned int wOff[NWL] , cOff[C];
	int wNeg[NWL] , wPos[NWL] , eNeg , ePos;
	bool lastIn;   // lane + (NWL-1) C < W (the earlier entries always are)
	__device__ __forceinline__ void set( int lane , const int *strideI )
	{
#pragma unroll
		for( int m=0 ; m<NWL ; m++ )
		{
			int t = lane + m*C , o = 0 , neg = 0 , pos = 0;
#pragma unroll
			for( int d=Dim-1 ; d>=0 ; d-- ){ const int od = t%3 - 1 ; t /= 3 ; o += od*strideI[d] ; if( od<0 ) neg |= 1<<d ; if( od>0 ) pos |= 1<<d; }
			wOff[m] = (unsigned int)o*(unsigned int)sizeof(Real) , wNeg[m] = neg , wPos[m] = pos;
		}
		lastIn = lane + (NWL-1)*C < W;
		int sig[Dim];
		eNeg = ePos = 0;
#pragma unroll
		for( int d=0 ; d<Dim ; d++ ){ const int b = Bit< Dim >( lane , d ) ; sig[d] = b ? strideI[d] : -strideI[d] ; if( b ) ePos |= 1<<d ; else eNeg |= 1<<d; }
#pragma unroll
		for( int c=0 ; c<C ; c++ )
		{
			int o = 0;
#pragma unroll
			for( int d=0 ; d<Dim ; d++ ) if( Bit< Dim >( c , d ) ) o += sig[d];
			cOff[c] = (unsigned int)o*(unsigned int)sizeof(Real);
		}
	}
};
template< typename Real > __device__ __forceinline__ Real LoadAt( const Real *base , unsigned int byteOff ){ return *(const Real *)( (const char *)base + byteOff ); }

// NodeGroup< Dim , MODE_GS > of node F (flat index i, column col) with the
// lane's constants.
template< int Dim , typename Real >
__device__ __forceinline__ void NodeGroupGs( const LevelView< Dim , Real > &L , const GsLane< Dim , Real > &ln , const int *F , unsigned int i , long long col , int lane , Real acc[5] )
{
	constexpr int C = DimC< Dim >::C , W = DimC< Dim >::W , NWL = GsLane< Dim , Real >::NWL;
	// the window's loads are issued where they are used, after the element's
	// (issued before the element stage they measured no faster in FP64 and 6 %
	// slower in FP32 at the finest 3D level, 12 % in 4D)
	constexpr bool PREFETCH = false;
	Real q20 = 0 , q02 = 0 , q11 = 0 , q10 = 0 , q01 = 0;
	int cse = 0 , low = 0 , high = 0;
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ){ cse = cse*3 + ( F[d]<=0 ? 0 : ( F[d]>=L.res ? 2 : 1 ) ) ; if( F[d]==0 ) low |= 1<<d ; if( F[d]==L.res ) high |= 1<<d; }
	const unsigned int ib = i*(unsigned int)sizeof(Real);
	const Real *Rn = L.R + col*W + lane , *Bn = L.B + col*W + lane , *Kn = L.K + cse*W + lane;
	Real wx[NWL] , wy[NWL] , wR[NWL] , wB[NWL] , wK[NWL];
#define EPR_LOAD_WINDOW_GS( m ) \
	{ \
		const bool in_ = ( (m)<NWL-1 || ln.lastIn ) && !( ln.wNeg[m] & low ) && !( ln.wPos[m] & high ); \
		const unsigned int ob_ = ib + ln.wOff[m]; \
		if( in_ ) wx[m] = LoadAt( L.x , ob_ ) , wy[m] = LoadAt( L.y , ob_ ) , wR[m] = Rn[(m)*C] , wB[m] = Bn[(m)*C] , wK[m] = Kn[(m)*C]; \
		else      wx[m] = wy[m] = wR[m] = wB[m] = wK[m] = 0; \
	}
	if constexpr( PREFETCH )
	{
#pragma unroll
		for( int m=0 ; m<NWL ; m++ ) EPR_LOAD_WINDOW_GS( m )
	}
	if( !( ln.eNeg & low ) && !( ln.ePos & high ) )
	{
		Real xp[C] , yp[C];
#pragma unroll
		for( int c=0 ; c<C ; c++ ){ const unsigned int ob = ib + ln.cOff[c] ; xp[c] = LoadAt( L.x , ob ) , yp[c] = LoadAt( L.y , ob ); }
		Real a20 , a02 , a11 , a10 , a01 , zy;
		ElementWedge< Dim , Real >( xp , yp , a20 , a02 , a11 , a10 , a01 , zy );
		q20 = a20*L.sW , q02 = a02*L.sW , q11 = a11*L.sW , q10 = a10*L.sW , q01 = a01*L.sW;
		(void)zy;
	}
#pragma unroll
	for( int m=0 ; m<NWL ; m++ )
	{
		if constexpr( !PREFETCH ) EPR_LOAD_WINDOW_GS( m )
		const int w = lane + m*C;
		const Real xr = wx[m] , yr = wy[m] , Rv = wR[m] , Bv = wB[m] , Kv = wK[m];
		const Real Dv = Kv*L.dW;
		if( w==W/2 ){ q20 += Rv + Dv ; q02 += Rv + Dv; }
		q10 += Real( 2 )*Rv*xr + Real( 2 )*Dv*xr;
		q01 += Real( 2 )*Rv*yr + Real( 2 )*Dv*yr;
		if( w!=W/2 ){ q10 -= Real( 2 )*Bv*yr ; q01 += Real( 2 )*Bv*xr; }
	}
#undef EPR_LOAD_WINDOW_GS
	q20 = GroupSum< C >( q20 ) , q02 = GroupSum< C >( q02 ) , q11 = GroupSum< C >( q11 ) , q10 = GroupSum< C >( q10 ) , q01 = GroupSum< C >( q01 );
	acc[0] = q20 , acc[1] = q02 , acc[2] = q11 , acc[3] = q10 , acc[4] = q01;
}

// One node group per C lanes (the lane's constants computed once per
// thread); block b of the phase.  Dense levels only (no active-node mask).
template< int Dim , typename Real >
__device__ __forceinline__ void GsPhase32( const LevelView< Dim , Real > &L , const ColourDiv< Dim > &cs , long long colBase , unsigned b )
{
	constexpr int C = DimC< Dim >::C;
	const unsigned int k = ( b*blockDim.x + threadIdx.x )/C;
	if( k>=cs.total ) return;
	const int lane = threadIdx.x & (C-1);
	int strideI[Dim];
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) strideI[d] = (int)L.stride[d];
	GsLane< Dim , Real > ln ; ln.set( lane , strideI );
	int F[Dim];
	cs.node( k , F );
	unsigned int i = 0;
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) i += (unsigned int)F[d]*(unsigned int)strideI[d];
	Real q[5];
	NodeGroupGs< Dim , Real >( L , ln , F , i , colBase+k , lane , q );
	if( lane==0 )
	{
		Real dx , dy;
		if( LevelStep< Dim , Real >( L , q , dx , dy ) ) L.x[i] += dx , L.y[i] += dy;
	}
}
template< int Dim , typename Real >
__global__ void __launch_bounds__( 256 ) KGsPhase32( LevelView< Dim , Real > L , ColourDiv< Dim > cs , long long colBase ){ GsPhase32< Dim , Real >( L , cs , colBase , XcdBlock( blockIdx.x , gridDim.x ) ); }

// One colour phase over a list of nodes (the active nodes of one colour of an
// adapted level, dense indices in increasing order); block b of the phase.
template< int Dim >
__device__ __forceinline__ void GsPhaseList( const LevelView< Dim > &L , const unsigned int *list , long long count , unsigned b )
{
	constexpr int C = DimC< Dim >::C;
	const long long g = (long long)b*blockDim.x + threadIdx.x;
	const long long k = g / C;
	const int lane = (int)( g % C );
	if( k>=count ) return;
	const unsigned int i = list[k];
	int F[Dim];
	{
		const unsigned int n1 = (unsigned int)L.res + 1;
		unsigned int t = i;
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ const unsigned int q = t / n1 ; F[d] = (int)( t - q*n1 ) ; t = q; }
	}
	double q[5];
	NodeGroup< Dim , MODE_GS , Dim==3 >( L , F , i , L.lay.col( F ) , lane , q );
	if( lane==0 )
	{
		double dx , dy;
		if( LevelStep< Dim >( L , q , dx , dy ) ) L.x[i] += dx , L.y[i] += dy;
	}
}
template< int Dim >
__global__ void __launch_bounds__( 256 ) KGsPhaseList( LevelView< Dim > L , const unsigned int *list , long long count ){ GsPhaseList< Dim >( L , list , count , XcdBlock( blockIdx.x , gridDim.x ) ); }

// All iterations of a small level in one workgroup (barrier between colours).
template< int Dim >
__global__ void __launch_bounds__( 512 ) KGsSmall( LevelView< Dim > L , int iters )
{
	constexpr int C = DimC< Dim >::C;
	for( int it=0 ; it<iters ; it++ ) for( int c=0 ; c<C ; c++ )
	{
		ColourSet< Dim > cs ; cs.set( L.res , c );
		for( long long g=threadIdx.x ; g<cs.total*C ; g+=blockDim.x )
		{
			const long long k = g / C;
			const int lane = (int)( g % C );
			int F[Dim];
			cs.node( k , F );
			const long long i = FlatIndex< Dim >( F , L.stride );
			double q[5];
			NodeGroup< Dim , MODE_GS >( L , F , i , L.lay.cstart[c]+k , lane , q );
			if( lane==0 && ( !L.fmask || L.fmask[i]==1 ) )
			{
				double dx , dy;
				if( LevelStep< Dim >( L , q , dx , dy ) ) L.x[i] += dx , L.y[i] += dy;
			}
		}
		__syncthreads();
	}
}

// Per-node quadratic fits of all nodes (test hook).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KQFit( LevelView< Dim > L , double *q )
{
	constexpr int C = DimC< Dim >::C;
	const long long g = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	const long long i = g / C;
	const int lane = (int)( g % C );
	if( i>=L.N ) return;
	int F[Dim];
	Unflatten< Dim >( i , L.res , F );
	double a[5];
	NodeGroup< Dim , MODE_QFIT , Dim==3 >( L , F , i , L.lay.col( F ) , lane , a );
	if( lane==0 ) for( int k=0 ; k<5 ; k++ ) q[5*i+k] = a[k];
}

// Energy terms as per-block partial-sum rows (W, y^T B x, x^T R x + y^T R y,
// x^T K x + y^T K y), summed in a fixed order by KSumRows< 4 >: KEnergy writes
// the window terms' rows (W = 0), KEnergyWedge the wedge term's rows after them.
// KEnergy: the nodes in the column order of R and B (colour by colour), so
// the node groups of a wavefront read consecutive columns.
// Idx: the index type, int where N 3^Dim < 2^31.
template< int Dim , typename Idx = long long >
__global__ void __launch_bounds__( 256 ) KEnergy( LevelView< Dim > L , double *partials )
{
	constexpr int C = DimC< Dim >::C , G = 256/C;
	__shared__ double sh[G][4];
	const Idx g = (Idx)blockIdx.x*blockDim.x + threadIdx.x;
	const Idx k = g / C;   // column
	const int lane = (int)( g % C );
	double a[5] = { 0 , 0 , 0 , 0 , 0 };
	if( k<L.N )
	{
		int c = 0;
#pragma unroll
		for( int q=1 ; q<C ; q++ ) c += k>=L.lay.cstart[q] ? 1 : 0;
		ColourSet< Dim > cs ; cs.set( L.res , c );
		int F[Dim];
		cs.node( k - L.lay.cstart[c] , F );
		const Idx i = (Idx)FlatIndex< Dim >( F , L.stride );
		NodeGroup< Dim , MODE_ENERGY , Dim==3 , double , Idx >( L , F , i , k , lane , a );
	}
	if( lane==0 ) for( int k=0 ; k<4 ; k++ ) sh[ threadIdx.x/C ][k] = a[k];
	__syncthreads();
	if( threadIdx.x==0 )
	{
		double s[4] = { 0 , 0 , 0 , 0 };
		for( int gidx=0 ; gidx<G ; gidx++ ) for( int k=0 ; k<4 ; k++ ) s[k] += sh[gidx][k];
		for( int k=0 ; k<4 ; k++ ) partials[ (long long)blockIdx.x*4 + k ] = s[k];
	}
}

// W: every element's wedge energy, one (element, coordinate pair) per thread
// (PairWedgeEnergy at unit resolution, scaled by the level's r^(4-Dim)); the
// element index runs with the last coordinate fastest, the element's lowest
// corner is its node of the same multi-index.  Rows (W, 0, 0, 0).  Idx: int
// where the (element, pair) count and N are below 2^31.
template< int Dim , typename Idx = long long >
__global__ void __launch_bounds__( 256 ) KEnergyWedge( LevelView< Dim > L , double *partials )
{
	constexpr int C = DimC< Dim >::C , NP = Dim*( Dim-1 )/2;
	__shared__ double sh[256];
	Idx ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= L.res;
	const Idx g = (Idx)blockIdx.x*blockDim.x + threadIdx.x;
	const Idx e = g / NP;
	double w = 0;
	if( e<ne )
	{
		int F[Dim];
		{
			Idx t = e;
#pragma unroll
			for( int d=Dim-1 ; d>=0 ; d-- ){ const Idx q = t / L.res ; F[d] = (int)( t - q*L.res ) ; t = q; }
		}
		const Idx i = (Idx)FlatIndex< Dim >( F , L.stride );
		Idx st[Dim];
		{
			int p = (int)( g - e*NP ) , pi = 0 , pj = 1;
#pragma unroll
			for( int a=0 ; a<Dim ; a++ )
#pragma unroll
				for( int b=a+1 ; b<Dim ; b++ ){ if( p==0 ) pi = a , pj = b ; p--; }
			st[0] = L.stride[pi] , st[1] = L.stride[pj];
			int q = 2;
#pragma unroll
			for( int d=0 ; d<Dim ; d++ ) if( d!=pi && d!=pj ) st[q++] = L.stride[d];
		}
		double xc[C] , yc[C];
#pragma unroll
		for( int c=0 ; c<C ; c++ )
		{
			Idx o = i;
#pragma unroll
			for( int q=0 ; q<Dim ; q++ ) if( (c>>q)&1 ) o += st[q];
			xc[c] = L.x[o] , yc[c] = L.y[o];
		}
		w = PairWedgeEnergy< Dim >( xc , yc )*L.sW;
	}
	sh[threadIdx.x] = w;
	__syncthreads();
	for( int h=128 ; h>0 ; h>>=1 ){ if( threadIdx.x<h ) sh[threadIdx.x] += sh[threadIdx.x+h] ; __syncthreads(); }
	if( threadIdx.x<4 ) partials[ (long long)blockIdx.x*4 + threadIdx.x ] = threadIdx.x==0 ? sh[0] : 0.;
}

// Pair p (in (i, j) order, i < j) of Dim coordinates; corner c' of the
// pair's order (bit q: coordinate order[q], order = (i, j, the others
// ascending)) as a corner of the element's own order (bit d: coordinate d).
template< int Dim > __host__ __device__ constexpr int PairFirst( int p ){ for( int a=0 ; a<Dim ; a++ ) for( int b=a+1 ; b<Dim ; b++ ){ if( p==0 ) return a ; p--; } return -1; }
template< int Dim > __host__ __device__ constexpr int PairSecond( int p ){ for( int a=0 ; a<Dim ; a++ ) for( int b=a+1 ; b<Dim ; b++ ){ if( p==0 ) return b ; p--; } return -1; }
template< int Dim > __host__ __device__ constexpr int PairCorner( int p , int cp )
{
	const int a = PairFirst< Dim >( p ) , b = PairSecond< Dim >( p );
	int ord[Dim] = {};
	ord[0] = a , ord[1] = b;
	int q = 2;
	for( int d=0 ; d<Dim ; d++ ) if( d!=a && d!=b ) ord[q++] = d;
	int c = 0;
	for( int k=0 ; k<Dim ; k++ ) if( ( cp>>k )&1 ) c |= 1<<ord[k];
	return c;
}
template< int Dim , int p , int ... cps >
__device__ __forceinline__ double ElementPairEnergy( const double *xs , const double *ys , std::integer_sequence< int , cps... > )
{
	constexpr int C = DimC< Dim >::C;
	const double xc[C] = { xs[ PairCorner< Dim >( p , cps ) ]... } , yc[C] = { ys[ PairCorner< Dim >( p , cps ) ]... };
	return PairWedgeEnergy< Dim >( xc , yc );
}
template< int Dim , int ... ps >
__device__ __forceinline__ double ElementWedgeEnergy( const double *xs , const double *ys , double sW , std::integer_sequence< int , ps... > )
{
	constexpr int C = DimC< Dim >::C;
	double w = 0;
	( ( w += ElementPairEnergy< Dim , ps >( xs , ys , std::make_integer_sequence< int , C >{} )*sW ) , ... );
	return w;
}

// KEnergyWedge with one thread per element and all its coordinate pairs: the
// element's 2^Dim corner pairs are loaded once instead of once per pair (the
// per-thread sum adds the pairs in order, then the block sums as KEnergyWedge
// does).  Rows (W, 0, 0, 0).  Idx: int where the element count and N are
// below 2^31.
template< int Dim , typename Idx = long long >
__global__ void __launch_bounds__( 256 , 2 ) KEnergyWedgeElem( LevelView< Dim > L , double *partials )
{
	constexpr int C = DimC< Dim >::C , NP = Dim*( Dim-1 )/2;
	__shared__ double sh[256];
	Idx ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= L.res;
	const Idx e = (Idx)blockIdx.x*blockDim.x + threadIdx.x;
	double w = 0;
	if( e<ne )
	{
		int F[Dim];
		{
			Idx t = e;
#pragma unroll
			for( int d=Dim-1 ; d>=0 ; d-- ){ const Idx q = t / L.res ; F[d] = (int)( t - q*L.res ) ; t = q; }
		}
		const Idx i = (Idx)FlatIndex< Dim >( F , L.stride );
		double xs[C] , ys[C];
#pragma unroll
		for( int c=0 ; c<C ; c++ )
		{
			Idx o = i;
#pragma unroll
			for( int d=0 ; d<Dim ; d++ ) if( (c>>d)&1 ) o += (Idx)L.stride[d];
			xs[c] = L.x[o] , ys[c] = L.y[o];
		}
		w = ElementWedgeEnergy< Dim >( xs , ys , L.sW , std::make_integer_sequence< int , NP >{} );
	}
	sh[threadIdx.x] = w;
	__syncthreads();
	for( int h=128 ; h>0 ; h>>=1 ){ if( threadIdx.x<h ) sh[threadIdx.x] += sh[threadIdx.x+h] ; __syncthreads(); }
	if( threadIdx.x<4 ) partials[ (long long)blockIdx.x*4 + threadIdx.x ] = threadIdx.x==0 ? sh[0] : 0.;
}

// Deterministic sum of nb rows of nv values (one block).
template< int NV >
__global__ void __launch_bounds__( 256 ) KSumRows( const double *in , long long nb , double *out )
{
	__shared__ double sh[256][NV];
	double s[NV];
	for( int k=0 ; k<NV ; k++ ) s[k] = 0;
	for( long long r=threadIdx.x ; r<nb ; r+=blockDim.x ) for( int k=0 ; k<NV ; k++ ) s[k] += in[ r*NV + k ];
	for( int k=0 ; k<NV ; k++ ) sh[threadIdx.x][k] = s[k];
	__syncthreads();
	for( int h=128 ; h>0 ; h>>=1 )
	{
		if( threadIdx.x<h ) for( int k=0 ; k<NV ; k++ ) sh[threadIdx.x][k] += sh[threadIdx.x+h][k];
		__syncthreads();
	}
	if( threadIdx.x==0 ) for( int k=0 ; k<NV ; k++ ) out[k] = sh[0][k];
}

// First pass of a fixed-order sum of nb rows of NV values over gridDim.x
// blocks: block b sums the contiguous rows [b per, (b+1) per) (per =
// ceil(nb / gridDim.x)) into out row b; KSumRows over the gridDim.x rows
// finishes it.
template< int NV >
__global__ void __launch_bounds__( 256 ) KSumRowsPart( const double *in , long long nb , double *out )
{
	__shared__ double sh[256][NV];
	const long long per = ( nb + gridDim.x - 1 ) / gridDim.x;
	const long long r0 = (long long)blockIdx.x*per , r1 = r0+per<nb ? r0+per : nb;
	double s[NV];
	for( int k=0 ; k<NV ; k++ ) s[k] = 0;
	for( long long r=r0+threadIdx.x ; r<r1 ; r+=blockDim.x ) for( int k=0 ; k<NV ; k++ ) s[k] += in[ r*NV + k ];
	for( int k=0 ; k<NV ; k++ ) sh[threadIdx.x][k] = s[k];
	__syncthreads();
	for( int h=128 ; h>0 ; h>>=1 )
	{
		if( threadIdx.x<h ) for( int k=0 ; k<NV ; k++ ) sh[threadIdx.x][k] += sh[threadIdx.x+h][k];
		__syncthreads();
	}
	if( threadIdx.x==0 ) for( int k=0 ; k<NV ; k++ ) out[ (long long)blockIdx.x*NV + k ] = sh[0][k];
}

// Block partial sums of a strided array (first pass of a deterministic reduction).
__global__ void __launch_bounds__( 256 ) KPartialSum( const double *in , long long n , double *partials )
{
	__shared__ double sh[256];
	double s = 0;
	const long long per = ( n + gridDim.x - 1 ) / gridDim.x;
	const long long b = (long long)blockIdx.x*per , e = b+per<n ? b+per : n;
	for( long long r=b+threadIdx.x ; r<e ; r+=blockDim.x ) s += in[r];
	sh[threadIdx.x] = s;
	__syncthreads();
	for( int h=128 ; h>0 ; h>>=1 ){ if( threadIdx.x<h ) sh[threadIdx.x] += sh[threadIdx.x+h] ; __syncthreads(); }
	if( threadIdx.x==0 ) partials[blockIdx.x] = sh[0];
}

// x_f (=, +=, -=) sum_c P(f,c) x_c, summed over coarse c in increasing index
// order (Eigen's column-major SpMV of SystemEnergy.h:162; weights of
// Hat.scalars.inl:373-386).
template< int Dim >
__device__ __forceinline__ void Prolong( const double *xc , double *xf , int cres , int mode , const unsigned char *fmaskF , long long f )
{
	const int fres = 2*cres;
	long long nf = 1 ; for( int d=0 ; d<Dim ; d++ ) nf *= fres+1;
	if( f>=nf ) return;
	// adapted levels: the prolongation has rows only for the fine functions that are indexed (Hat.scalars.inl:1694-1744)
	if( fmaskF && !fmaskF[f] ){ if( mode==0 ) xf[f] = 0 ; return; }
	int F[Dim];
	Unflatten< Dim >( f , fres , F );
	int nodd = 0 , oddDims[Dim];
	for( int d=0 ; d<Dim ; d++ ) if( F[d]&1 ) oddDims[nodd++] = d;
	double s = 0;
	for( int m=0 ; m<(1<<nodd) ; m++ )
	{
		long long c = 0;
		double w = 1;
		int oi = 0;
		for( int d=0 ; d<Dim ; d++ )
		{
			int cd = F[d]>>1;
			if( F[d]&1 ){ cd += ( m>>( nodd-1-oi ) ) & 1 ; oi++ ; w *= 0.5; }
			c = c*(cres+1) + cd;
		}
		s += w*xc[c];
	}
	if( mode==0 ) xf[f] = s;
	else if( mode==1 ) xf[f] = xf[f] + s;
	else xf[f] = xf[f] - s;
}
template< int Dim >
__global__ void __launch_bounds__( 256 ) KProlong( const double *xc , double *xf , int cres , int mode , const unsigned char *fmaskF ){ Prolong< Dim >( xc , xf , cres , mode , fmaskF , (long long)blockIdx.x*blockDim.x + threadIdx.x ); }

// Prolong of x and y together (out (=, +=, -=) P in) for levels whose
// fine functions number below 2^31: 32-bit indices, the fine multi-index by
// one fixed-divisor division per coordinate, the coarse index as a base plus
// the coarse strides of the odd coordinates; the same terms in the same order
// as Prolong (the weight of every term is 1/2 per odd coordinate), so the
// same bits.
template< int Dim >
__global__ void __launch_bounds__( 256 ) KProlong2( const double *xc , const double *yc , double *xf , double *yf , int cres , int mode , const unsigned char *fmaskF , FastDiv n1 )
{
	const unsigned int fres = 2*cres;
	unsigned int nf = 1 ; for( int d=0 ; d<Dim ; d++ ) nf *= fres+1;
	const unsigned int f = blockIdx.x*blockDim.x + threadIdx.x;
	if( f>=nf ) return;
	if( fmaskF && !fmaskF[f] ){ if( mode==0 ) xf[f] = 0 , yf[f] = 0 ; return; }
	int F[Dim];
	{
		unsigned int t = f;
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ const unsigned int q = n1.div( t ) ; F[d] = (int)( t - q*n1.d ) ; t = q; }
	}
	unsigned int cs[Dim] , c0 = 0;
	{
		unsigned int s = 1;
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ cs[d] = s ; c0 += (unsigned int)( F[d]>>1 )*s ; s *= cres+1; }
	}
	int nodd = 0;
	unsigned int oddStride[Dim];
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) if( F[d]&1 ) oddStride[nodd++] = cs[d];
	double w = 1;
	for( int k=0 ; k<nodd ; k++ ) w *= 0.5;
	double sx = 0 , sy = 0;
	for( int m=0 ; m<(1<<nodd) ; m++ )
	{
		unsigned int c = c0;
		for( int oi=0 ; oi<nodd ; oi++ ) if( ( m>>( nodd-1-oi ) ) & 1 ) c += oddStride[oi];
		sx += w*xc[c] , sy += w*yc[c];
	}
	if( mode==0 ) xf[f] = sx , yf[f] = sy;
	else if( mode==1 ) xf[f] = xf[f] + sx , yf[f] = yf[f] + sy;
	else xf[f] = xf[f] - sx , yf[f] = yf[f] - sy;
}

// out = a + b
// y = M x for a level operator stored by columns (column f holds M(f+o_w, f)):
// y_r = sum over the window of r of M(r, f) x_f with f = r - o_w, taken in
// increasing f (the accumulation order of a column-major sparse product).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KApplyCols( const double *M , const double *x , double *y , LevelView< Dim > L )
{
	constexpr int W = DimC< Dim >::W;
	const long long r = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( r>=L.N ) return;
	int F[Dim];
	Unflatten< Dim >( r , L.res , F );
	double s = 0;
	for( int w=W-1 ; w>=0 ; w-- )
	{
		int t = w , Ff[Dim];
		long long f = r;
		bool in = true;
		for( int d=Dim-1 ; d>=0 ; d-- ){ const int o = t%3 - 1 ; t /= 3 ; Ff[d] = F[d]-o ; in = in && Ff[d]>=0 && Ff[d]<=L.res ; f -= o*L.stride[d]; }
		if( !in ) continue;
		const double v = M[ L.lay.col( Ff )*W + w ];
		if( v!=0 ) s = __dadd_rn( s , __dmul_rn( v , x[f] ) );
	}
	y[r] = s;
}

// Plain device copy (16 B per lane per iteration), the calibration kernel for HBM counters.
__global__ void __launch_bounds__( 256 ) KCopy( const double2 *__restrict__ a , double2 *__restrict__ b , long long n )
{
	for( long long k=(long long)blockIdx.x*blockDim.x + threadIdx.x ; k<n ; k+=(long long)gridDim.x*blockDim.x ) b[k] = a[k];
}

// Up to 64 device ranges for one launch (blockIdx.y selects the range).
struct Spans { void *p[64] ; unsigned long long n[64]; };
// zero-fills n[j] bytes at p[j] (16-byte aligned starts)
__global__ void __launch_bounds__( 256 ) KZeroSpans( Spans s )
{
	unsigned char *p = (unsigned char *)s.p[ blockIdx.y ];
	const unsigned long long nb = s.n[ blockIdx.y ] , nw = nb/16;
	for( unsigned long long i=(unsigned long long)blockIdx.x*blockDim.x + threadIdx.x ; i<nw ; i+=(unsigned long long)gridDim.x*blockDim.x ) ( (uint4 *)p )[i] = make_uint4( 0 , 0 , 0 , 0 );
	if( blockIdx.x==0 && threadIdx.x<nb-nw*16 ) p[ nw*16 + threadIdx.x ] = 0;
}
// scales n[j] doubles at p[j] by a
__global__ void __launch_bounds__( 256 ) KScaleSpans( Spans s , double a )
{
	double *p = (double *)s.p[ blockIdx.y ];
	const unsigned long long n = s.n[ blockIdx.y ];
	for( unsigned long long i=(unsigned long long)blockIdx.x*blockDim.x + threadIdx.x ; i<n ; i+=(unsigned long long)gridDim.x*blockDim.x ) p[i] *= a;
}

__global__ void KAddInto( const double *a , const double *b , double *out , long long n )
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( i<n ) out[i] = a[i] + b[i];
}

__global__ void KScale2( double *x , double *y , long long n , const double *s )
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( i<n ){ x[i] *= s[0] ; y[i] *= s[0]; }
}


//////////////////////////////////////////////////////////////////////////
// Build phase
//////////////////////////////////////////////////////////////////////////

// Uniform B-spline weights of degree Deg at x in [0,1] (GridSamples.inl:32-48).
template< int Deg >
__device__ __forceinline__ void BSpline( double x , double *v )
{
	if constexpr( Deg==0 ) v[0] = 1.;
	else
	{
		const double Scale = 1./Deg;
		BSpline< Deg-1 >( x , v+1 );
		v[0] = v[1] * (1.-x) * Scale;
#pragma unroll
		for( int i=1 ; i<Deg ; i++ )
		{
			const double x1 = ( x-i+Deg ) , x2 = ( -x+i+1 );
			v[i] = ( v[i]*x1 + v[i+1]*x2 ) * Scale;
		}
		v[Deg] *= x * Scale;
	}
}

template< int Dim >
struct TreeGrids
{
	double *data[MAXT];            // data[t]: (2^(t-1)+1)^Dim corners x NV values
	unsigned char *kids[MAXT];     // kids[t]: (2^t)^Dim flags, t<=maxDepth
	int maxDepth;                  // D (tree depths 1..D+1)
};

template< int Dim > __device__ __forceinline__ long long CornerIndex( int t , const int *o ){ const int r = 1<<(t-1) ; long long i = 0 ; for( int d=0 ; d<Dim ; d++ ) i = i*(r+1) + o[d] ; return i; }
template< int Dim > __device__ __forceinline__ long long TreeNodeIndex( int t , const int *o ){ long long i = 0 ; for( int d=0 ; d<Dim ; d++ ) i = i*( 1LL<<t ) + o[d] ; return i; }

// Per occupied cell: the mean position and the summed density statistics of its
// samples, in the cell's sample order (GridSamples.inl:381-387).
template< int Dim >
__global__ void KCellStats( const double *pos , const int *cellStart , const int *cellSamples , int ncell , double *cellP , double *cellData )
{
	constexpr int NV = DimC< Dim >::NDNI;
	const int c = blockIdx.x*blockDim.x + threadIdx.x;
	if( c>=ncell ) return;
	double p[Dim] , dat[NV];
	for( int d=0 ; d<Dim ; d++ ) p[d] = 0;
	for( int k=0 ; k<NV ; k++ ) dat[k] = 0;
	for( int s=cellStart[c] ; s<cellStart[c+1] ; s++ )
	{
		const double *q = pos + (long long)cellSamples[s]*Dim;
		for( int d=0 ; d<Dim ; d++ ) p[d] += q[d];
		dat[0] += 1.;
		for( int d=0 ; d<Dim ; d++ ) dat[1+d] += q[d]*1.;
		for( int a=0 ; a<Dim ; a++ ) for( int b=0 ; b<Dim ; b++ ) dat[1+Dim+a*Dim+b] += q[a]*q[b]*1.;
	}
	const double inv = 1. / (double)( cellStart[c+1]-cellStart[c] );
	for( int d=0 ; d<Dim ; d++ ) cellP[ (long long)c*Dim+d ] = p[d]*inv;
	for( int k=0 ; k<NV ; k++ ) cellData[ (long long)c*NV+k ] = dat[k];
}

// Node creation of the tree splatters (Misha/RegularTree.inl:762-800): the
// parents (C_t+i)>>1, i in [-K,K+1], of every occupied cell's ancestors get
// children.  One thread per (cell, depth).
template< int Dim >
__global__ void KTouch( const int *cellIdx , int ncell , TreeGrids< Dim > T , int K )
{
	const long long g = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	const int nt = T.maxDepth;   // depths t = 2 .. D+1
	if( g>=(long long)ncell*nt ) return;
	const int c = (int)( g / nt ) , t = (int)( g % nt ) + 2;
	int lo[Dim] , hi[Dim] , P[Dim];
	for( int d=0 ; d<Dim ; d++ )
	{
		const int Cd = cellIdx[ (long long)c*Dim+d ]>>( T.maxDepth+1-t );
		lo[d] = max( 0 , ( Cd-K )>>1 );
		hi[d] = min( (1<<(t-1))-1 , ( Cd+K+1 )>>1 );
		P[d] = lo[d];
	}
	while( true )
	{
		T.kids[t-1][ TreeNodeIndex< Dim >( t-1 , P ) ] = 1;
		int d = Dim-1;
		for( ; d>=0 ; d-- ){ if( ++P[d]<=hi[d] ) break ; P[d] = lo[d]; }
		if( d<0 ) break;
	}
}

// rows of a row-major array by index (negative: a row of zeros)
__global__ void __launch_bounds__( 256 ) KGatherRows( const double *src , const long long *idx , long long n , int w , double *dst )
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( i>=n ) return;
	const long long s = idx[i];
	for( int k=0 ; k<w ; k++ ) dst[ i*w+k ] = s>=0 ? src[ s*w+k ] : 0.;
}

// Estimator splat (GridSamples.inl:436-491, merged cells): one thread per
// (cell, tree depth, tap); density statistics added with FP64 atomics.
// Tap `tap` of cell c at depth t: the corner offset and the B-spline weight;
// false when the corner lies outside the depth's grid.
template< int Dim , int KR >
__device__ __forceinline__ bool EstTap( const double *p , int t , int tap , int *o , double &scl )
{
	constexpr int Deg = 2*KR+1 , NT = Deg+1;
	const int res = 1<<(t-1);
	scl = 1.;
	int tt = tap , ii[Dim];
	for( int d=Dim-1 ; d>=0 ; d-- ){ ii[d] = tt%NT ; tt /= NT; }
	for( int d=0 ; d<Dim ; d++ )
	{
		const int E = (int)floor( p[d]*res );
		double v[NT];
		BSpline< Deg >( p[d]*res - E , v );
		scl = scl * v[ ii[d] ];
		o[d] = E - KR + ii[d];
		if( o[d]<0 || o[d]>res ) return false;
	}
	return true;
}

// The coarse depths (few corners, every cell hits them): per-workgroup sums in
// LDS, one global atomic per corner value and workgroup.
// Estimator values per node: weight, position sum (Dim), covariance sum
// (Dim x Dim, symmetric).  Only the diagonal and upper covariance entries are
// splatted; readers mirror the lower ones (EstMirror), which keeps the sums
// exactly symmetric, as the reference's ordered sums are, with 29 % (4D) /
// 23 % (3D) fewer atomics.
template< int Dim > __host__ __device__ constexpr bool EstLower( int k ){ return k>Dim && ( k-1-Dim )/Dim > ( k-1-Dim )%Dim; }
template< int Dim > __host__ __device__ inline void EstMirror( double *v )
{
	for( int i=1 ; i<Dim ; i++ ) for( int j=0 ; j<i ; j++ ) v[ 1+Dim+i*Dim+j ] = v[ 1+Dim+j*Dim+i ];
}

// Depth t = tLo + blockIdx.y (several depths in one launch).
template< int Dim , int KR >
__global__ void __launch_bounds__( 256 ) KEstSplatLds( const double *cellP , const double *cellData , int ncell , TreeGrids< Dim > T , const double *depthScale , int tLo )
{
	const int t = tLo + (int)blockIdx.y;
	constexpr int NV = DimC< Dim >::NDNI , NT = 2*KR+2;
	extern __shared__ double acc[];
	int taps = 1 ; for( int d=0 ; d<Dim ; d++ ) taps *= NT;
	int corners = 1 ; for( int d=0 ; d<Dim ; d++ ) corners *= ( 1<<(t-1) ) + 1;
	for( int i=threadIdx.x ; i<corners*NV ; i+=blockDim.x ) acc[i] = 0;
	__syncthreads();
	const double dw = depthScale[t-1];
	const long long items = (long long)ncell*taps;
	for( long long g=(long long)blockIdx.x*blockDim.x + threadIdx.x ; g<items ; g+=(long long)gridDim.x*blockDim.x )
	{
		const int c = (int)( g / taps ) , tap = (int)( g % taps );
		int o[Dim];
		double scl;
		if( !EstTap< Dim , KR >( cellP + (long long)c*Dim , t , tap , o , scl ) ) continue;
		double *node = acc + CornerIndex< Dim >( t , o )*NV;
		for( int k=0 ; k<NV ; k++ ) if( !EstLower< Dim >( k ) ) atomicAdd( node+k , ( cellData[ (long long)c*NV+k ]*dw )*scl );
	}
	__syncthreads();
	for( int i=threadIdx.x ; i<corners*NV ; i+=blockDim.x ) if( acc[i]!=0 ) atomicAdd( T.data[t]+i , acc[i] );
}

// Depths tMin..maxDepth+1: one thread per (cell, depth, tap), FP64 atomics.
// perm: a multiplier coprime with ncell; the k-th cell processed is cell
// (k perm) mod ncell, so that neighbouring cells (which share their corners)
// are not splatted at the same time (3D depth 7: 4.54 -> 4.33 ms per build).
template< int Dim , int KR >
__global__ void KEstSplat( const double *cellP , const double *cellData , int ncell , TreeGrids< Dim > T , const double *depthScale , int tMin , unsigned int perm )
{
	constexpr int NV = DimC< Dim >::NDNI , NT = 2*KR+2;
	int taps = 1 ; for( int d=0 ; d<Dim ; d++ ) taps *= NT;
	const int nt = T.maxDepth+2-tMin;
	const long long g = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( g>=(long long)ncell*nt*taps ) return;
	const int tap = (int)( g % taps );
	const long long ct = g / taps;
	const int t = (int)( ct % nt ) + tMin;
	const int c = (int)( ( (unsigned long long)( ct / nt )*perm ) % (unsigned int)ncell );
	int o[Dim];
	double scl;
	if( !EstTap< Dim , KR >( cellP + (long long)c*Dim , t , tap , o , scl ) ) return;
	double *node = T.data[t] + CornerIndex< Dim >( t , o )*NV;
	const double dw = depthScale[t-1];
	for( int k=0 ; k<NV ; k++ ) if( !EstLower< Dim >( k ) ) atomicAdd( node+k , ( cellData[ (long long)c*NV+k ]*dw )*scl );
}

// Orders a wavefront's LDS accesses on either side (a wave's own LDS
// operations complete in order, so this only constrains the compiler).
__device__ __forceinline__ void WaveSync( void )
{
	__builtin_amdgcn_fence( __ATOMIC_RELEASE , "wavefront" );
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence( __ATOMIC_ACQUIRE , "wavefront" );
}

// Morton key of every occupied cell of depth D (the bits of its coordinates
// interleaved, coordinate 0 first at each bit, the highest bit first) and its
// index.  Sorted by key, the cells inside one cell of any coarser depth are
// consecutive: the depth-t ancestor is the key shifted right by Dim (D+1-t).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KMortonKeys( const int *cellIdx , int ncell , int D , unsigned long long *key , int *idx )
{
	const int c = blockIdx.x*blockDim.x + threadIdx.x;
	if( c>=ncell ) return;
	int F[Dim];
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) F[d] = cellIdx[ (long long)c*Dim+d ];
	unsigned long long k = 0;
	for( int b=D-1 ; b>=0 ; b-- )
#pragma unroll
		for( int d=0 ; d<Dim ; d++ ) k = ( k<<1 ) | (unsigned long long)( ( F[d]>>b ) & 1 );
	key[c] = k , idx[c] = c;
}

// KEstSplat over the cells in Morton order: the thread of the first cell of
// each depth-t cell (and tap) adds the contributions of all the cells inside
// that depth-t cell in registers and issues one atomic per value, instead of
// one per cell (the other threads return).  A cell whose mean position lies
// outside its own cell (rounding, or a position clamped into the grid) adds
// its own contribution with atomics of its own.  A run of cells also starts
// at every multiple of SplatRun Morton positions, which bounds a thread's
// serial loop for the coarse depths of large inputs.
constexpr int SplatRun = 64;
__device__ __forceinline__ bool SplatRunHead( const unsigned long long *mkey , int i , int shift )
{
	return i==0 || ( i % SplatRun )==0 || ( mkey[i-1]>>shift )!=( mkey[i]>>shift );
}
__device__ __forceinline__ bool SplatRunNext( const unsigned long long *mkey , int i , int j , int ncell , int shift )
{
	return j<ncell && ( j==i || ( j % SplatRun )!=0 ) && ( mkey[j]>>shift )==( mkey[i]>>shift );
}
// WIDE (taps a multiple of 64, so that a wavefront holds one cell and depth):
// the run's sums go through LDS and each atomic instruction covers
// consecutive values of consecutive nodes (the taps of a wavefront run along
// the last coordinate) instead of one value of 64 scattered nodes.
template< int Dim , int KR , bool WIDE >
__global__ void __launch_bounds__( 256 ) KEstSplatSeg( const double *cellP , const double *cellData , const unsigned long long *mkey , const int *morder , int ncell , TreeGrids< Dim > T , const double *depthScale , int tMin )
{
	constexpr int NV = DimC< Dim >::NDNI , NT = 2*KR+2;
	int taps = 1 ; for( int d=0 ; d<Dim ; d++ ) taps *= NT;
	const int D = T.maxDepth , nt = D+2-tMin;
	const long long g = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( g>=(long long)ncell*nt*taps ) return;
	const int tap = (int)( g % taps );
	const long long ct = g / taps;
	const int t = (int)( ct % nt ) + tMin;
	const int i = (int)( ct / nt );
	const int shift = Dim*( D+1-t );
	if( !SplatRunHead( mkey , i , shift ) ) return;   // uniform over a wavefront when WIDE
	const double dw = depthScale[t-1];
	double acc[NV];
	int oh[Dim];
	bool have = false;
	if constexpr( WIDE )
	{
#pragma unroll
		for( int k=0 ; k<NV ; k++ ) acc[k] = 0;
	}
	for( int j=i ; SplatRunNext( mkey , i , j , ncell , shift ) ; j++ )
	{
		const int c = morder[j];
		int o[Dim];
		double scl;
		if( !EstTap< Dim , KR >( cellP + (long long)c*Dim , t , tap , o , scl ) ) continue;
		if( !have )
		{
			have = true;
#pragma unroll
			for( int d=0 ; d<Dim ; d++ ) oh[d] = o[d];
#pragma unroll
			for( int k=0 ; k<NV ; k++ ) acc[k] = 0;
		}
		bool same = true;
#pragma unroll
		for( int d=0 ; d<Dim ; d++ ) same = same && o[d]==oh[d];
		if( same )
		{
#pragma unroll
			for( int k=0 ; k<NV ; k++ ) if( !EstLower< Dim >( k ) ) acc[k] += ( cellData[ (long long)c*NV+k ]*dw )*scl;
		}
		else
		{
			double *node = T.data[t] + CornerIndex< Dim >( t , o )*NV;
			for( int k=0 ; k<NV ; k++ ) if( !EstLower< Dim >( k ) ) atomicAdd( node+k , ( cellData[ (long long)c*NV+k ]*dw )*scl );
		}
	}
	if constexpr( WIDE )
	{
		__shared__ double sv[4][64*NV];
		__shared__ long long sn[4][64];
		const int lane = threadIdx.x & 63 , wv = threadIdx.x>>6;
#pragma unroll
		for( int k=0 ; k<NV ; k++ ) sv[wv][ lane*NV+k ] = acc[k];
		sn[wv][lane] = have ? CornerIndex< Dim >( t , oh )*NV : -1;
		WaveSync();
		for( int q=lane ; q<64*NV ; q+=64 )
		{
			const int src = q / NV , k = q - src*NV;
			const long long n0 = sn[wv][src];
			const double v = sv[wv][q];
			if( n0>=0 && !EstLower< Dim >( k ) && v!=0 ) atomicAdd( T.data[t] + n0 + k , v );
		}
	}
	else
	{
		if( !have ) return;
		double *node = T.data[t] + CornerIndex< Dim >( t , oh )*NV;
#pragma unroll
		for( int k=0 ; k<NV ; k++ ) if( !EstLower< Dim >( k ) ) atomicAdd( node+k , acc[k] );
	}
}

// Linear interpolation of a tree depth at p (GridSamples.inl:774-809).
template< int Dim , int NV >
__device__ __forceinline__ void TreeEval( const TreeGrids< Dim > &T , int t , const double *p , double *out , int nv )
{
	const int res = 1<<(t-1);
	int e[Dim];
	double v[Dim][2];
	for( int d=0 ; d<Dim ; d++ ){ e[d] = (int)floor( p[d]*res ) ; BSpline< 1 >( p[d]*res - e[d] , v[d] ); }
	for( int k=0 ; k<nv ; k++ ) out[k] = 0;
	for( int f=0 ; f<(1<<Dim) ; f++ )
	{
		int o[Dim];
		bool in = true;
		double scale = 1.;
		for( int d=0 ; d<Dim ; d++ ){ const int b = Bit< Dim >( f , d ) ; o[d] = e[d]+b ; in = in && o[d]<=res ; scale *= v[d][b]; }
		if( !in ) continue;
		const double *node = T.data[t] + CornerIndex< Dim >( t , o )*NV;
		for( int k=0 ; k<nv ; k++ ) out[k] += node[k]*scale;
	}
}

// Symmetric eigenvalues by cyclic Jacobi, ascending.  Up to 64 sweeps, ending
// early when the off-diagonal sum of squares is zero or when a sweep leaves
// every entry bit for bit unchanged: the sweep is then a fixed point, so the
// remaining sweeps would change nothing and the result equals the full 64
// sweeps'.  Tiny off-diagonals whose theta^2 overflows give t = 0 (identity
// rotations) while their squares stay nonzero, so without this test most
// estimator covariances ran all 64 sweeps instead of 6-20.
template< int Dim >
__device__ void SymEig( double A[Dim][Dim] , double *ev )
{
	auto same = []( double a , double b ){ return __double_as_longlong( a )==__double_as_longlong( b ); };
	for( int sweep=0 ; sweep<64 ; sweep++ )
	{
		double off = 0;
		for( int i=0 ; i<Dim ; i++ ) for( int j=i+1 ; j<Dim ; j++ ) off += A[i][j]*A[i][j];
		if( off==0 ) break;
		bool changed = false;
		for( int p=0 ; p<Dim ; p++ ) for( int q=p+1 ; q<Dim ; q++ )
		{
			if( A[p][q]==0 ) continue;
			const double theta = ( A[q][q]-A[p][p] ) / ( 2*A[p][q] );
			const double t = ( theta>=0 ? 1. : -1. ) / ( fabs( theta ) + sqrt( theta*theta+1 ) );
			const double c = 1./sqrt( t*t+1 ) , s = t*c;
			for( int k=0 ; k<Dim ; k++ ){ const double akp = A[k][p] , akq = A[k][q] ; A[k][p] = c*akp - s*akq ; A[k][q] = s*akp + c*akq ; changed |= !same( A[k][p] , akp ) || !same( A[k][q] , akq ); }
			for( int k=0 ; k<Dim ; k++ ){ const double apk = A[p][k] , aqk = A[q][k] ; A[p][k] = c*apk - s*aqk ; A[q][k] = s*apk + c*aqk ; changed |= !same( A[p][k] , apk ) || !same( A[q][k] , aqk ); }
		}
		if( !changed ) break;
	}
	for( int i=0 ; i<Dim ; i++ ) ev[i] = A[i][i];
	for( int i=1 ; i<Dim ; i++ ) for( int j=i ; j>0 && ev[j-1]>ev[j] ; j-- ){ const double t = ev[j] ; ev[j] = ev[j-1] ; ev[j-1] = t; }
}

// Per sample: fractional depth, measure and noise of the tree estimator
// (GridSamples.inl:811-896, 82-105), the R weight exp(-n^2/(2 sigma^2))
// (ExteriorPoissonRecon.cpp:222) and the scaled frame that is splatted
// (ExteriorPoisson.h:181-213).
template< int Dim >
__global__ void KEstQuery( const double *pos , const double *skewPacked , int n , TreeGrids< Dim > T , double nSigma ,
                           double *measure , double *noise , double *depth , double *rweight , double *skewScaled )
{
	constexpr int NV = DimC< Dim >::NDNI , CD = DimC< Dim >::CD , CoDim = 2;
	const int s = blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=n ) return;
	const double *p = pos + (long long)s*Dim;
	const double target = (double)( 1<<(Dim-CoDim) );
	const int maxT = T.maxDepth+1;
	auto spc = [&]( int t ){ double w ; TreeEval< Dim , NV >( T , t , p , &w , 1 ) ; return w / (double)( 1<<( (t-1)*(Dim-CoDim) ) ); };
	double dep;
	{
		// walk down from the finest depth while the count stays below the
		// target; the last two counts evaluated are the bracketing ones
		// (cur at d1, prev at d2 = d1+1), so none is evaluated twice
		int d1 = maxT , d2 = maxT+1;
		double cur = 0 , prev = 0;
		while( d2>1 )
		{
			const double v = spc( d1 );
			if( !( v<target ) ){ cur = v ; break; }   // also stops on NaN
			prev = v , d1-- , d2--;
		}
		if( d2==1 ) dep = 1 + log( prev/target ) / log( 2. );
		else if( d2==maxT+1 ) dep = maxT + log( cur/target ) / log( 2. );
		else dep = !prev ? (double)d1 : d1 + log( target/cur ) / log( prev/cur );
	}
	double info[NV];
	if( !( dep<maxT ) ) TreeEval< Dim , NV >( T , maxT , p , info , NV ) , EstMirror< Dim >( info );      // also catches NaN
	else if( !( dep>=1 ) ) TreeEval< Dim , NV >( T , 1 , p , info , NV ) , EstMirror< Dim >( info );
	else
	{
		const int d1 = min( max( (int)floor( dep ) , 1 ) , maxT-1 ) , d2 = d1+1;
		const double dd = dep-d1;
		double a[NV] , b[NV];
		TreeEval< Dim , NV >( T , d1 , p , a , NV ) , EstMirror< Dim >( a );
		TreeEval< Dim , NV >( T , d2 , p , b , NV ) , EstMirror< Dim >( b );
		for( int k=0 ; k<NV ; k++ ) info[k] = a[k]*(1.-dd) + b[k]*dd;
	}
	const double w = info[0];
	double A[Dim][Dim] , ev[Dim];
	for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) A[i][j] = info[1+Dim+i*Dim+j]/w - info[1+i]/w * info[1+j]/w;
	SymEig< Dim >( A , ev );
	double sum = 0 , low = 0;
	for( int d=0 ; d<Dim ; d++ ){ sum += ev[d] ; if( d<CoDim ) low += ev[d]; }
	const double nz = ( low*Dim ) / ( sum*CoDim );
	const double m = 1./w;
	measure[s] = m , noise[s] = nz , depth[s] = dep - 1.;
	rweight[s] = exp( -nz*nz/( 2.*nSigma*nSigma ) );
	// frame: S / |S|_F * measure * 2^(depth*CoDim)
	double sq = 0;
	{
		const double *q = skewPacked + (long long)s*CD;
		double M[Dim][Dim];
		for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) M[i][j] = 0;
		for( int i=0 , idx=0 ; i<Dim ; i++ ) for( int j=0 ; j<i ; j++ , idx++ ) M[i][j] = q[idx] , M[j][i] = -q[idx];
		for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) sq += M[i][j]*M[i][j];
		const double inv = 1./sqrt( sq );
		const double lv = (double)( 1<<( T.maxDepth*CoDim ) );
		for( int c=0 ; c<CD ; c++ ) skewScaled[ (long long)s*CD + c ] = ( ( q[c]*inv )*m )*lv;
	}
}

// The tree depths of a sample's skew splat (GridSamples.inl:393-491): t[0]
// (and t[1]) with the weights fw (-1: the whole value at one depth).
template< int Dim >
__device__ __forceinline__ int SkewDepths( double depth , int D , int ts[2] , double fw[2] )
{
	const double dep = depth + 1.;
	if( !( dep<D+1 ) ){ ts[0] = D+1 , fw[0] = -1 ; return 1; }    // also catches NaN
	if( dep<=1 ){ ts[0] = 1 , fw[0] = -1 ; return 1; }
	const int d1 = min( max( (int)floor( dep ) , 1 ) , D );
	const double dd = dep-d1;
	ts[0] = d1 , fw[0] = 1.-dd , ts[1] = d1+1 , fw[1] = dd;
	return 2;
}
// tap ii of p at depth t: its corner o and B-spline weight; false outside the grid
template< int Dim , int KR >
__device__ __forceinline__ bool SkewTap( const double *p , int t , const int *ii , int *o , double &scl )
{
	constexpr int Deg = 2*KR+1 , NT = Deg+1;
	const int res = 1<<(t-1);
	scl = 1.;
	bool in = true;
	for( int d=0 ; d<Dim ; d++ )
	{
		const int E = (int)floor( p[d]*res );
		double v[NT];
		BSpline< Deg >( p[d]*res - E , v );
		scl = scl * v[ ii[d] ];
		o[d] = E - KR + ii[d];
		in = in && o[d]>=0 && o[d]<=res;
	}
	return in;
}
// value c of sample s's contribution at depth t with weight fw, before the tap's weight
template< int Dim >
__device__ __forceinline__ double SkewValue( const double *skewScaled , int s , int c , int t , double fw )
{
	constexpr int CD = DimC< Dim >::CD;
	const double v = skewScaled[ (long long)s*CD + c ] , dw = (double)( 1ull<<( (t-1)*Dim ) );
	return fw<0 ? v*dw : ( v*fw )*dw;
}

// Skew splat at the fractional sample depth (GridSamples.inl:393-491): one
// thread per (sample, tap), up to two tree depths, FP64 atomics.  perm: the
// k-th sample processed is sample (k perm) mod n (perm coprime with n).
template< int Dim , int KR >
__global__ void KSkewSplat( const double *pos , const double *skewScaled , const double *depth , int n , TreeGrids< Dim > T , unsigned int perm )
{
	constexpr int CD = DimC< Dim >::CD , NT = 2*KR+2;
	int taps = 1 ; for( int d=0 ; d<Dim ; d++ ) taps *= NT;
	const long long g = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( g>=(long long)n*taps ) return;
	const int s = (int)( ( (unsigned long long)( g / taps )*perm ) % (unsigned int)n ) , tap = (int)( g % taps );
	const double *p = pos + (long long)s*Dim;
	int ts[2] ; double fw[2];
	const int nts = SkewDepths< Dim >( depth[s] , T.maxDepth , ts , fw );
	int ii[Dim];
	{
		int tt = tap;
		for( int d=Dim-1 ; d>=0 ; d-- ){ ii[d] = tt%NT ; tt /= NT; }
	}
	for( int k=0 ; k<nts ; k++ )
	{
		const int t = ts[k];
		int o[Dim];
		double scl;
		if( !SkewTap< Dim , KR >( p , t , ii , o , scl ) ) continue;
		double *node = T.data[t] + CornerIndex< Dim >( t , o )*CD;
		for( int c=0 ; c<CD ; c++ ) atomicAdd( node+c , SkewValue< Dim >( skewScaled , s , c , t , fw[k] )*scl );
	}
}

// The skew splat's (sample, depth) entries keyed by (depth t, the Morton key
// of the depth-t cell holding the sample): one thread per sample, whose cell
// of depth D is its clamped floor at resolution 2^D, as the host's binning
// takes it; entry 2s+k is the k-th depth of sample s, value 2s+k.  A sample
// with one depth gets the key SkewNoEntry( Dim , D ) (every bit of the sorted
// range set, above every real key) for k = 1.
__host__ __device__ inline unsigned long long SkewNoEntry( int Dim , int D ){ return ( 1ull<<( Dim*D+4 ) ) - 1; }
template< int Dim >
__global__ void __launch_bounds__( 256 ) KSkewEntries( const double *pos , const double *depth , int n , int D , unsigned long long *key , int *val )
{
	const int s = blockIdx.x*blockDim.x + threadIdx.x;
	if( s>=n ) return;
	const int res = 1<<D;
	int F[Dim];
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) F[d] = max( 0 , min( (int)floor( pos[ (long long)s*Dim+d ]*res ) , res-1 ) );
	unsigned long long m = 0;
	for( int b=D-1 ; b>=0 ; b-- )
#pragma unroll
		for( int d=0 ; d<Dim ; d++ ) m = ( m<<1 ) | (unsigned long long)( ( F[d]>>b ) & 1 );
	int ts[2] ; double fw[2];
	const int nts = SkewDepths< Dim >( depth[s] , D , ts , fw );
#pragma unroll
	for( int k=0 ; k<2 ; k++ )
	{
		key[ 2*(long long)s+k ] = k<nts ? ( (unsigned long long)ts[k]<<( Dim*D ) ) | ( m>>( Dim*( D+1-ts[k] ) ) ) : SkewNoEntry( Dim , D );
		val[ 2*(long long)s+k ] = 2*s+k;
	}
}

// KSkewSplat over the sorted entries: one thread per (entry, tap); the thread
// of the first entry of a run of equal keys (a run also starts at every
// multiple of SplatRun entries) adds the run's contributions to the depth-t
// cell in registers and issues one atomic per value (a sample whose position
// lies outside its own cell adds its own share by atomics), the others return.
// WIDE: taps a multiple of 64, the run's sums flushed through LDS as in
// KEstSplatSeg.
template< int Dim , int KR , bool WIDE >
__global__ void __launch_bounds__( 256 ) KSkewSplatRuns( const double *pos , const double *skewScaled , const double *depth , const unsigned long long *key , const int *val , long long nent , TreeGrids< Dim > T )
{
	constexpr int CD = DimC< Dim >::CD , NT = 2*KR+2;
	int taps = 1 ; for( int d=0 ; d<Dim ; d++ ) taps *= NT;
	const int D = T.maxDepth;
	const long long g = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( g>=nent*taps ) return;
	const int tap = (int)( g % taps );
	const long long e = g / taps;
	const unsigned long long ke = key[e];
	if( ke==SkewNoEntry( Dim , D ) ) return;
	if( e>0 && ( e % SplatRun )!=0 && key[e-1]==ke ) return;
	int ii[Dim];
	{
		int tt = tap;
		for( int d=Dim-1 ; d>=0 ; d-- ){ ii[d] = tt%NT ; tt /= NT; }
	}
	const int t = (int)( ke>>( Dim*D ) );
	double acc[CD];
	int oh[Dim];
	bool have = false;
	if constexpr( WIDE )
	{
#pragma unroll
		for( int q=0 ; q<CD ; q++ ) acc[q] = 0;
	}
	for( long long f=e ; f<nent && ( f==e || ( f % SplatRun )!=0 ) && key[f]==ke ; f++ )
	{
		const int s = val[f]>>1 , k = val[f] & 1;
		int ts[2] ; double fw[2];
		SkewDepths< Dim >( depth[s] , D , ts , fw );
		const double w = fw[k];
		int o[Dim];
		double scl;
		if( !SkewTap< Dim , KR >( pos + (long long)s*Dim , t , ii , o , scl ) ) continue;
		if( !have )
		{
			have = true;
#pragma unroll
			for( int d=0 ; d<Dim ; d++ ) oh[d] = o[d];
#pragma unroll
			for( int q=0 ; q<CD ; q++ ) acc[q] = 0;
		}
		bool same = true;
#pragma unroll
		for( int d=0 ; d<Dim ; d++ ) same = same && o[d]==oh[d];
		if( same )
		{
#pragma unroll
			for( int q=0 ; q<CD ; q++ ) acc[q] += SkewValue< Dim >( skewScaled , s , q , t , w )*scl;
		}
		else
		{
			double *node = T.data[t] + CornerIndex< Dim >( t , o )*CD;
			for( int q=0 ; q<CD ; q++ ) atomicAdd( node+q , SkewValue< Dim >( skewScaled , s , q , t , w )*scl );
		}
	}
	if constexpr( WIDE )
	{
		__shared__ double sv[4][64*CD];
		__shared__ long long sn[4][64];
		const int lane = threadIdx.x & 63 , wv = threadIdx.x>>6;
#pragma unroll
		for( int q=0 ; q<CD ; q++ ) sv[wv][ lane*CD+q ] = acc[q];
		sn[wv][lane] = have ? CornerIndex< Dim >( t , oh )*CD : -1;
		WaveSync();
		for( int r=lane ; r<64*CD ; r+=64 )
		{
			const int src = r / CD , q = r - src*CD;
			const long long n0 = sn[wv][src];
			const double v = sv[wv][r];
			if( n0>=0 && v!=0 ) atomicAdd( T.data[t] + n0 + q , v );
		}
	}
	else
	{
		if( !have ) return;
		double *node = T.data[t] + CornerIndex< Dim >( t , oh )*CD;
#pragma unroll
		for( int q=0 ; q<CD ; q++ ) atomicAdd( node+q , acc[q] );
	}
}

// cmap[e] = c for every occupied cell c (element index e with coordinate 0 fastest)
template< int Dim >
__global__ void KCellMap( const int *cellIdx , int ncell , int res , int *cmap )
{
	const int c = blockIdx.x*blockDim.x + threadIdx.x;
	if( c>=ncell ) return;
	long long e = 0;
	for( int d=Dim-1 ; d>=0 ; d-- ) e = e*res + cellIdx[ (size_t)c*Dim+d ];
	cmap[e] = c;
}

__global__ void KScaleArray( double *a , long long n , double s )
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( i<n ) a[i] *= s;
}

// a[i] = 0 where m[i] is 0
__global__ void KMaskZero( double *a , const unsigned char *m , long long n )
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( i<n && !m[i] ) a[i] = 0;
}

// Target field at every finest element centre: sum over tree depths up to the
// leaf containing the centre of the multilinear interpolation of the splatted
// frames (ExteriorPoisson.h:294-307, GridSamples.inl:600-651).  Also writes
// |S|^2 (full-matrix Frobenius) per element.
template< int Dim >
__global__ void KField( TreeGrids< Dim > T , double *S , double *S2 )
{
	constexpr int CD = DimC< Dim >::CD;
	const int D = T.maxDepth , res = 1<<D;
	long long ne = 1 ; for( int d=0 ; d<Dim ; d++ ) ne *= res;
	const long long e = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( e>=ne ) return;
	int I[Dim];
	{
		long long t = e;
		for( int d=0 ; d<Dim ; d++ ){ I[d] = (int)( t%res ) ; t /= res; }
	}
	double p[Dim];
	for( int d=0 ; d<Dim ; d++ ) p[d] = ( I[d] + 0.5 ) / res;
	int L = 1;
	while( L<=D )
	{
		int o[Dim];
		for( int d=0 ; d<Dim ; d++ ) o[d] = I[d]>>( D+1-L );
		if( !T.kids[L][ TreeNodeIndex< Dim >( L , o ) ] ) break;
		L++;
	}
	double s[CD];
	for( int c=0 ; c<CD ; c++ ) s[c] = 0;
	for( int t=1 ; t<=L ; t++ )
	{
		const int r = 1<<(t-1);
		int eo[Dim];
		double v[Dim][2];
		for( int d=0 ; d<Dim ; d++ ){ eo[d] = (int)floor( p[d]*r ) ; BSpline< 1 >( p[d]*r - eo[d] , v[d] ); }
		for( int f=0 ; f<(1<<Dim) ; f++ )
		{
			int o[Dim];
			double scl = 1.;
			bool in = true;
			for( int d=0 ; d<Dim ; d++ ){ const int b = Bit< Dim >( f , d ) ; scl = scl * v[d][b] ; o[d] = eo[d]+b ; in = in && o[d]<=r; }
			if( !in ) continue;
			const double *node = T.data[t] + CornerIndex< Dim >( t , o )*CD;
			for( int c=0 ; c<CD ; c++ ) s[c] = s[c] + node[c]*scl;
		}
	}
	double sq = 0;
	{
		double M[Dim][Dim];
		for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) M[i][j] = 0;
		for( int i=0 , idx=0 ; i<Dim ; i++ ) for( int j=0 ; j<i ; j++ , idx++ ) M[i][j] = s[idx] , M[j][i] = -s[idx];
		for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<Dim ; j++ ) sq += M[i][j]*M[i][j];
	}
	for( int c=0 ; c<CD ; c++ ) S[ e*CD + c ] = s[c];
	S2[e] = sq;
}

// B(f2,f1) = sum over common elements E of <S(E), FS[i1][i2]> (Hat.scalars.inl:1637-1686).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KAssembleB( const double *S , const double *FS , int res , double *B )
{
	constexpr int W = DimC< Dim >::W , C = DimC< Dim >::C , CD = DimC< Dim >::CD;
	long long nf = 1 ; for( int d=0 ; d<Dim ; d++ ) nf *= res+1;
	const long long g = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( g>=nf*W ) return;
	const long long f1 = g / W;
	const int w = (int)( g % W );
	int F1[Dim] , F2[Dim];
	Unflatten< Dim >( f1 , res , F1 );
	{
		int t = w;
		for( int d=Dim-1 ; d>=0 ; d-- ){ F2[d] = F1[d] + t%3 - 1 ; t /= 3; }
	}
	int lo[Dim] , hi[Dim] , E[Dim];
	for( int d=0 ; d<Dim ; d++ )
	{
		if( F2[d]<0 || F2[d]>res ){ B[g] = 0 ; return; }
		lo[d] = max( 0 , max( F1[d] , F2[d] )-1 );
		hi[d] = min( res , min( F1[d] , F2[d] )+1 );
		if( lo[d]>=hi[d] ){ B[g] = 0 ; return; }
		E[d] = lo[d];
	}
	double dot = 0;
	while( true )
	{
		long long e = 0;
		int i1 = 0 , i2 = 0;
		for( int d=Dim-1 ; d>=0 ; d-- ) e = e*res + E[d];
		for( int d=0 ; d<Dim ; d++ ) i1 = i1*2 + ( F1[d]-E[d] ) , i2 = i2*2 + ( F2[d]-E[d] );
		double M[Dim][Dim];
		for( int a=0 ; a<Dim ; a++ ) for( int b=0 ; b<Dim ; b++ ) M[a][b] = 0;
		const double *s = S + e*CD;
		for( int a=0 , idx=0 ; a<Dim ; a++ ) for( int b=0 ; b<a ; b++ , idx++ ) M[a][b] = s[idx] , M[b][a] = -s[idx];
		const double *st = FS + (long long)( i1*C+i2 )*Dim*Dim;
		double dd = 0;
		for( int a=0 ; a<Dim ; a++ ) for( int b=0 ; b<Dim ; b++ ) dd += st[a*Dim+b]*M[a][b];
		dot += dd;
		int d = Dim-1;
		for( ; d>=0 ; d-- ){ if( ++E[d]<hi[d] ) break ; E[d] = lo[d]; }
		if( d<0 ) break;
	}
	B[g] = dot;
}

// R = (E + E^T) * sWeight * mean(measure), E the sample-weighted delta mass of
// Hat.scalars.inl:1006-1062: for the pair (f,g) summed over the occupied cells
// adjacent to both in cell order, within a cell over its samples in order.
// Samples in cell order (k-th entry of cellSamples): positions and weights
// stored contiguously, so that the per-cell sums of REntry read sequentially.
template< int Dim >
__global__ void __launch_bounds__( 256 ) KCellOrder( const double *pos , const double *rweight , const int *cellSamples , int n , double *posS , double *rwS )
{
	const int k = blockIdx.x*blockDim.x + threadIdx.x;
	if( k>=n ) return;
	const int s = cellSamples[k];
	for( int d=0 ; d<Dim ; d++ ) posS[ (long long)k*Dim+d ] = pos[ (long long)s*Dim+d ];
	rwS[k] = rweight[s];
}
// Entry M(G, F) of R for window offset w of node F (0 outside the grid or
// without samples): the common elements E of the two supports in
// lexicographic order, each the sum over its samples (in the cell's order;
// posS, rwS as KCellOrder writes them) of the two hat values times the sample
// weight.
template< int Dim >
__device__ __forceinline__ double REntry( const double *posS , const int *cellMap , const int *cellStart , const double *rwS , int res , const int *F , int w , bool &any )
{
	constexpr int W = DimC< Dim >::W;
	int G[Dim];
	{
		int t = w;
		for( int d=Dim-1 ; d>=0 ; d-- ){ G[d] = F[d] + t%3 - 1 ; t /= 3; }
	}
	int lo[Dim] , hi[Dim] , E[Dim];
	any = false;
	for( int d=0 ; d<Dim ; d++ )
	{
		if( G[d]<0 || G[d]>res ) return 0;
		lo[d] = max( 0 , max( F[d] , G[d] )-1 );
		hi[d] = min( res , min( F[d] , G[d] )+1 );
		E[d] = lo[d];
	}
	// lexicographic order of the pair: a = min(F,G), b = max(F,G)
	bool FleG = true;
	for( int d=0 ; d<Dim ; d++ ){ if( F[d]<G[d] ){ FleG = true ; break; } if( F[d]>G[d] ){ FleG = false ; break; } }
	const int *A = FleG ? F : G , *Bn = FleG ? G : F;
	const bool diag = ( w==W/2 );
	double acc = 0;
	while( true )
	{
		long long e = 0;
		for( int d=Dim-1 ; d>=0 ; d-- ) e = e*res + E[d];
		const int c = cellMap[e];
		if( c>=0 )
		{
			double value = 0;
			for( int k=cellStart[c] ; k<cellStart[c+1] ; k++ )
			{
				double va = 1. , vb = 1.;
				for( int d=0 ; d<Dim ; d++ )
				{
					double lp = posS[ (long long)k*Dim+d ] * res;
					lp -= E[d];
					va *= ( A[d]-E[d] ) ? lp : (1.-lp);
					vb *= ( Bn[d]-E[d] ) ? lp : (1.-lp);
				}
				value += va * vb * rwS[k];
			}
			if( diag ) value /= 2.;
			if( any ) acc = acc + value ; else acc = value , any = true;
		}
		int d = Dim-1;
		for( ; d>=0 ; d-- ){ if( ++E[d]<hi[d] ) break ; E[d] = lo[d]; }
		if( d<0 ) break;
	}
	return any ? ( diag ? acc+acc : acc ) : 0;
}
// The nodes with an occupied incident element, once each (list order is
// arbitrary): thread t marks corner t mod 2^Dim of occupied cell t / 2^Dim.
template< int Dim >
__global__ void __launch_bounds__( 256 ) KOccupiedNodes( const int *cellIdx , int ncell , int res , int *flag , unsigned int *list , unsigned int *count )
{
	const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( t>=( (long long)ncell<<Dim ) ) return;
	const int c = (int)( t>>Dim ) , k = (int)( t & ( (1<<Dim)-1 ) );
	long long f = 0;
	for( int d=0 ; d<Dim ; d++ ) f = f*( res+1 ) + cellIdx[ (size_t)c*Dim+d ] + ( ( k>>d )&1 );
	if( atomicExch( flag+f , 1 )==0 ) list[ atomicAdd( count , 1u ) ] = (unsigned int)f;
}
// One thread per window entry of a listed node (R is zero elsewhere: every
// entry of a node without an occupied incident element vanishes).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KAssembleR( const double *posS , const int *cellMap , const int *cellStart , const double *rwS ,
                                                      int res , const double *scale , double *R , ColLayout< Dim > lay ,
                                                      const unsigned int *list , const unsigned int *count , float *fR )
{
	constexpr int W = DimC< Dim >::W;
	const long long g = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( g>=(long long)count[0]*W ) return;
	const int w = (int)( g % W );
	int F[Dim];
	Unflatten< Dim >( list[ g/W ] , res , F );
	bool any;
	const double v = REntry< Dim >( posS , cellMap , cellStart , rwS , res , F , w , any );
	const double r = any ? v * scale[0] * scale[1] : 0;
	const long long o = lay.col( F )*W + w;
	R[o] = r;
	if( fR ) fR[o] = (float)r;   // the FP32 sweeps' copy (optional)
}

// ---- B assembly, one thread per finest node f1 and all 3^Dim outputs of its
// column (Hat.scalars.inl:1637-1686).  The 2^Dim incident elements' fields are
// loaded once; every output (f2,f1) sums its common elements in lexicographic
// order, each element's <S(E), FS[i1][i2]> in (a,b) order without contraction,
// with the unit-resolution table FS scaled by res^(2-Dim) (an exact power of
// two), so the values equal the reference's.
template< int Dim > struct FSTab;
template<> struct FSTab< 3 >{ static constexpr double v( int i1 , int i2 , int a , int b ){ return wedge3::FS[i1][i2][a][b]; } };
template<> struct FSTab< 4 >{ static constexpr double v( int i1 , int i2 , int a , int b ){ return wedge4::FS[i1][i2][a][b]; } };

template< int Dim > constexpr int WOff( int w , int d ){ for( int k=Dim-1 ; k>d ; k-- ) w /= 3 ; return w%3 - 1; }
template< int Dim > constexpr int PackedIndex( int a , int b ){ int idx = 0 ; for( int i=0 ; i<Dim ; i++ ) for( int j=0 ; j<i ; j++ , idx++ ) if( i==a && j==b ) return idx ; return -1; }

// dot of FS[i1][i2] (elements incident to f1 indexed by e-bits, i1/i2 relative corners) with S(E)
template< int Dim , int i1 , int i2 , int ab >
__device__ __forceinline__ void FSDotTerm( const double *s , double &dd )
{
	constexpr int a = ab / Dim , b = ab % Dim;
	if constexpr( a!=b && FSTab< Dim >::v( i1 , i2 , a , b )!=0.0 )
	{
		constexpr double st = FSTab< Dim >::v( i1 , i2 , a , b );
		// M(a,b) = p for a>b, -p for a<b
		const double m = a>b ? s[ PackedIndex< Dim >( a , b ) ] : -s[ PackedIndex< Dim >( b , a ) ];
		dd = __dadd_rn( dd , __dmul_rn( st , m ) );
	}
}
template< int Dim , int i1 , int i2 , int ... abs >
__device__ __forceinline__ double FSDot( const double *s , std::integer_sequence< int , abs... > )
{
	double dd = 0;
	( FSDotTerm< Dim , i1 , i2 , abs >( s , dd ) , ... );
	return dd;
}
// output w of node f1: elements b (bits, coordinate 0 most significant) with E = F1-1+b common to f1 and f2 = f1+off(w)
template< int Dim , int w , int b >
__device__ __forceinline__ void BElement( const double (&S)[1<<Dim][DimC< Dim >::CD] , double &dot )
{
	constexpr bool common = [](){ for( int d=0 ; d<Dim ; d++ ){ const int e = -1 + Bit< Dim >( b , d ) , o = WOff< Dim >( w , d ) ; if( o<e || o>e+1 ) return false; } return true; }();
	if constexpr( common )
	{
		constexpr int i1 = [](){ int i = 0 ; for( int d=0 ; d<Dim ; d++ ) i = i*2 + ( 1 - Bit< Dim >( b , d ) ) ; return i; }();
		constexpr int i2 = [](){ int i = 0 ; for( int d=0 ; d<Dim ; d++ ) i = i*2 + ( WOff< Dim >( w , d ) + 1 - Bit< Dim >( b , d ) ) ; return i; }();
		dot = __dadd_rn( dot , FSDot< Dim , i1 , i2 >( S[b] , std::make_integer_sequence< int , Dim*Dim >{} ) );
	}
}
template< int Dim , int w , int ... bs >
__device__ __forceinline__ double BOutput( const double (&S)[1<<Dim][DimC< Dim >::CD] , std::integer_sequence< int , bs... > )
{
	double dot = 0;
	( BElement< Dim , w , bs >( S , dot ) , ... );
	return dot;
}
template< int Dim , int ... ws >
__device__ __forceinline__ void BOutputs( const double (&S)[1<<Dim][DimC< Dim >::CD] , const int *F , int res , double scale , double *col , float *fcol , std::integer_sequence< int , ws... > )
{
	( [&]{
		bool in = true;
		for( int d=0 ; d<Dim ; d++ ){ const int g = F[d] + WOff< Dim >( ws , d ) ; in = in && g>=0 && g<=res; }
		const double v = in ? __dmul_rn( BOutput< Dim , ws >( S , std::make_integer_sequence< int , 1<<Dim >{} ) , scale ) : 0.;
		col[ws] = v;
		if( fcol ) fcol[ws] = (float)v;
	}() , ... );
}

// outputs W0 + ws of a node, to out[ws] (stride 1)
template< int Dim , int W0 , int ... ws >
__device__ __forceinline__ void BOutputsChunk( const double (&S)[1<<Dim][DimC< Dim >::CD] , const int *F , int res , double scale , double *out , std::integer_sequence< int , ws... > )
{
	( [&]{
		constexpr int w = W0 + ws;
		bool in = true;
		for( int d=0 ; d<Dim ; d++ ){ const int g = F[d] + WOff< Dim >( w , d ) ; in = in && g>=0 && g<=res; }
		out[ws] = in ? __dmul_rn( BOutput< Dim , w >( S , std::make_integer_sequence< int , 1<<Dim >{} ) , scale ) : 0.;
	}() , ... );
}
template< int Dim , int Q , int CH >
__device__ __forceinline__ void BColumnsChunk( const double (&Sv)[1<<Dim][DimC< Dim >::CD] , const int *F , int res , double scale , bool live , double *stage , double *B , float *fB , long long k0 , long long N )
{
	constexpr int W = DimC< Dim >::W;
	const int lane = threadIdx.x & 63;
	if( live ) BOutputsChunk< Dim , Q*CH >( Sv , F , res , scale , stage + lane*CH , std::make_integer_sequence< int , CH >{} );
	__syncthreads();
	for( int i=lane ; i<64*CH ; i+=64 )
	{
		const int cl = i / CH , wl = i - cl*CH;
		if( k0+cl<N )
		{
			const double v = stage[i];
			B[ ( k0+cl )*W + Q*CH + wl ] = v;
			if( fB ) fB[ ( k0+cl )*W + Q*CH + wl ] = (float)v;
		}
	}
	__syncthreads();
}
template< int Dim , int CH , int ... qs >
__device__ __forceinline__ void BColumns( const double (&Sv)[1<<Dim][DimC< Dim >::CD] , const int *F , int res , double scale , bool live , double *stage , double *B , float *fB , long long k0 , long long N , std::integer_sequence< int , qs... > )
{
	( BColumnsChunk< Dim , qs , CH >( Sv , F , res , scale , live , stage , B , fB , k0 , N ) , ... );
}

// KAssembleBNode for dense levels, one thread per column of B (the nodes in
// the column order, colour by colour): each wavefront stages 27 outputs of
// its 64 columns in LDS at a time and writes them as consecutive runs of
// 27 doubles, instead of one strided double per lane and output.  The values
// are KAssembleBNode's, bit for bit.  fB (optional): the same rounded to float
// (the FP32 sweeps' copy, written here instead of by a narrowing pass).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KAssembleBCols( const double *S , int res , double scale , double *B , ColLayout< Dim > lay , float *fB )
{
	constexpr int W = DimC< Dim >::W , C = DimC< Dim >::C , CD = DimC< Dim >::CD , CH = 27;
	__shared__ double stage[4][64*CH];
	const long long N = lay.cstart[C];
	const long long k = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	const long long k0 = k - ( threadIdx.x & 63 );
	const bool live = k<N;
	int F[Dim];
	double Sv[C][CD];
	{
		int c = 0;
#pragma unroll
		for( int q=1 ; q<C ; q++ ) c += k>=lay.cstart[q] ? 1 : 0;
		ColourSet< Dim > cs ; cs.set( res , c );
		cs.node( live ? k - lay.cstart[c] : 0 , F );
	}
#pragma unroll
	for( int b=0 ; b<C ; b++ )
	{
		bool in = live;
		long long e = 0;
		for( int d=Dim-1 ; d>=0 ; d-- ){ const int E = F[d]-1+Bit< Dim >( b , d ) ; in = in && E>=0 && E<res ; e = e*res + E; }
#pragma unroll
		for( int c=0 ; c<CD ; c++ ) Sv[b][c] = in ? S[ e*CD + c ] : 0.;
	}
	BColumns< Dim , CH >( Sv , F , res , scale , live , stage[ threadIdx.x>>6 ] , B , fB , k0 , N , std::make_integer_sequence< int , W/CH >{} );
}

// fB (optional, dense levels): the values rounded to float as well (the FP32
// sweeps' copy).
template< int Dim >
__global__ void __launch_bounds__( 256 ) KAssembleBNode( const double *S , int res , double scale , double *B , ColLayout< Dim > lay , const unsigned char *emask , const unsigned char *fmask , float *fB )
{
	constexpr int W = DimC< Dim >::W , C = DimC< Dim >::C , CD = DimC< Dim >::CD;
	long long nf = 1 ; for( int d=0 ; d<Dim ; d++ ) nf *= res+1;
	const long long f1 = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if( f1>=nf ) return;
	int F[Dim];
	Unflatten< Dim >( f1 , res , F );
	double Sv[C][CD];
#pragma unroll
	for( int b=0 ; b<C ; b++ )
	{
		bool in = true;
		long long e = 0;
		for( int d=Dim-1 ; d>=0 ; d-- ){ const int E = F[d]-1+Bit< Dim >( b , d ) ; in = in && E>=0 && E<res ; e = e*res + E; }
		if( emask ) in = in && emask[e];   // adapted levels: only indexed elements are integrated (Hat.scalars.inl:1640-1650)
#pragma unroll
		for( int c=0 ; c<CD ; c++ ) Sv[b][c] = in ? S[ e*CD + c ] : 0.;
	}
	double *col = B + lay.col( F )*W;
	BOutputs< Dim >( Sv , F , res , scale , col , fB && !fmask ? fB + lay.col( F )*W : nullptr , std::make_integer_sequence< int , W >{} );
	if( fmask )   // adapted levels: entries only between indexed functions
	{
		for( int w=0 ; w<W ; w++ )
		{
			bool in = fmask[f1]!=0;
			long long g = 0;
			for( int d=0 , t=w , p=W/3 ; d<Dim ; d++ , p/=3 ){ const int G = F[d] + ( t/p )%3 - 1 ; in = in && G>=0 && G<=res ; g = g*(res+1) + G; }
			if( !in || !fmask[g] ) col[w] = 0;
		}
	}
}


// ---- T = P^T M_f by fine node, one launch per parity class of the fine node.
// For a fine node K of parity class CLS (bit d = K_d & 1) every entry m = K + delta
// of its column has its coarse parents i at compile-time window offsets
// wo_d = i_d - (K_d>>1) + 1:
//   K_d even: delta -1 -> {0,1} (1/2 each), 0 -> {1} (1), +1 -> {1,2} (1/2 each)
//   K_d odd : delta -1 -> {1} (1), 0 -> {1,2} (1/2 each), +1 -> {2} (1)
// so each thread scatters its column into W register accumulators; the terms
// of every output arrive in increasing m (the Eigen product order of
// KGalerkinT's gather) and the values are identical.
template< int Dim > struct GalerkinShape
{
	static constexpr int W = DimC< Dim >::W;
	static constexpr int Pow3( int e ){ return e==0 ? 1 : 3*Pow3( e-1 ); }
	static constexpr int Digit( int w , int d ){ return ( w / Pow3( Dim-1-d ) ) % 3; }
	// number of parents of the entry in dimension d
	static constexpr int Opts( int cls , int wm , int d ){ const int p = ( cls>>d )&1 , dl = Digit( wm , d ) - 1 ; return p==0 ? ( dl==0 ? 1 : 2 ) : ( dl==0 ? 2 : 1 ); }
	// window offset of option `opt` in dimension d
	static constexpr int Off( int cls , int wm , int d , int opt )
	{
		const int p = ( cls>>d )&1 , dl = Digit( wm , d ) - 1;
		if( p==0 ) return dl==-1 ? opt : ( dl==0 ? 1 : 1+opt );
		return dl==-1 ? 1 : ( dl==0 ? 1+opt : 2 );
	}
	static constexpr int NumCombos( int cls , int wm ){ int n = 1 ; for( int d=0 ; d<Dim ; d++ ) n *= Opts( cls , wm , d ) ; return n; }
	// option of dimension d in combination `combo` (mixed radix, first dimension most significant)
	static constexpr int OptOf( int cls , int wm , int combo , int d ){ int r = 1 ; for( int e=Dim-1 ; e>d ; e-- ) r *= Opts( cls , wm , e ) ; return ( combo / r ) % Opts( cls , wm , d ); }
	static constexpr int Wo( int cls , int wm , int combo ){ int w = 0 ; for( int d=0 ; d<Dim ; d++ ) w = w*3 + Off( cls , wm , d , OptOf( cls , wm , combo , d ) ) ; return w; }
	// the weight of an entry with two parents in a dimension is 1/2 there; an
	// entry with one parent has weight 1 (m even, m = 2i)
	static constexpr double Weight( int cls , int wm ){ double w = 1 ; for( int d=0 ; d<Dim ; d++ ) if( Opts( cls , wm , d )==2 ) w *= 0.5 ; return w; }
};

// Scatter of one column entry into the accumulators of the outputs whose
// first window digit is WO0 (acc[wo % 27]).
template< int Dim , int CLS , int WM , int WO0 , int ... Cs >
__device__ __forceinline__ void GalerkinScatter( double mv , const int *K , int cres , double *acc , std::integer_sequence< int , Cs... > )
{
	typedef GalerkinShape< Dim > G;
	constexpr double pmi = G::Weight( CLS , WM );
	( [&]{
		constexpr int wo = G::Wo( CLS , WM , Cs );
		if constexpr( wo/27==WO0 )
		{
			bool in = true;
#pragma unroll
			for( int d=0 ; d<Dim ; d++ )
			{
				const int off = G::Off( CLS , WM , d , G::OptOf( CLS , WM , Cs , d ) );
				const int i = ( K[d]>>1 ) + off - 1;
				in = in && i>=0 && i<=cres;
			}
			if( in ) acc[wo%27] = __dadd_rn( acc[wo%27] , __dmul_rn( pmi , mv ) );
		}
	}() , ... );
}

// The column entries WM = 27 CH + i (i in Is) of one fine node, from LDS,
// scattered into the accumulators of window digit WO0 (entries outside the
// grid are skipped; a zero entry adds zero, which leaves every accumulator
// as it is: they start at +0 and never become -0).
template< int Dim , int CLS , int CH , int WO0 , int ... Is >
__device__ __forceinline__ void GalerkinChunk( const double *col , const int *K , int fres , int cres , double *acc , std::integer_sequence< int , Is... > )
{
	typedef GalerkinShape< Dim > G;
	( [&]{
		constexpr int WM = 27*CH + Is;
		bool in = true;
#pragma unroll
		for( int d=0 ; d<Dim ; d++ ){ const int m = K[d] + G::Digit( WM , d ) - 1 ; in = in && m>=0 && m<=fres; }
		if( in ) GalerkinScatter< Dim , CLS , WM , WO0 >( col[Is] , K , cres , acc , std::make_integer_sequence< int , G::NumCombos( CLS , WM ) >{} );
	}() , ... );
}
template< int Dim , int CLS , int CH , int ... WOs >
__device__ __forceinline__ void GalerkinChunkWave( int wave , const double *col , const int *K , int fres , int cres , double *acc , std::integer_sequence< int , WOs... > )
{
	( [&]{ if( wave==WOs ) GalerkinChunk< Dim , CLS , CH , WOs >( col , K , fres , cres , acc , std::make_integer_sequence< int , 27 >{} ); }() , ... );
}

// Chunk by chunk (27 column entries each): stage the block's columns in LDS,
// then each wavefront scatters its lanes' columns into its outputs.
template< int Dim , int CLS , int ... CHs >
__device__ __forceinline__ void GalerkinTChunks( const double *Mf , float *fMf , long long col0 , int lim , bool have , int wave , const int *K , int fres , int cres , double *buf , double *acc , std::integer_sequence< int , CHs... > )
{
	constexpr int W = DimC< Dim >::W , NW = W/27;
	const int lane = threadIdx.x & 63;
	( [&]{
		if( CHs ) __syncthreads();
		for( int e=threadIdx.x ; e<64*27 ; e+=64*NW )
		{
			const long long o = ( col0 + e/27 )*W + 27*CHs + e%27;
			const double v = e<lim ? Mf[o] : 0.;
			buf[e] = v;
			if( fMf && e<lim ) fMf[o] = (float)v;   // every entry of Mf is loaded once: its float copy
		}
		__syncthreads();
		if( have ) GalerkinChunkWave< Dim , CLS , CHs >( wave , buf + lane*27 , K , fres , cres , acc , std::make_integer_sequence< int , NW >{} );
	}() , ... );
}

// T is offset-major over the fine columns in colour-grouped order:
// T[wo*Nf + col] (wo the coarse window offset), so a wavefront's stores here
// and its loads in KGalerkinC are contiguous.  64 fine nodes of parity class
// CLS per workgroup (block b of the class), one wavefront per first window
// digit of the outputs (3D: one, 4D: three, 27 accumulators each): the
// columns pass through LDS 27 entries at a time (contiguous loads), are
// scattered from LDS into the accumulators, and leave one offset plane at a
// time (contiguous stores).
template< int Dim , int CLS >
__device__ __forceinline__ void GalerkinTBlock( const double *Mf , float *fMf , double *T , long long Nf , int cres , const ColourSet< Dim > &cs , long long colBase , unsigned b , double *buf )
{
	constexpr int W = DimC< Dim >::W , NCH = W/27;
	const int lane = threadIdx.x & 63 , wave = threadIdx.x >> 6 , fres = 2*cres;
	const long long g = (long long)b*64 + lane;
	const bool have = g<cs.total;
	int K[Dim];
	if( have ) cs.node( g , K );
	else for( int d=0 ; d<Dim ; d++ ) K[d] = 0;
	const long long col0 = colBase + (long long)b*64;
	const int lim = (int)min( (long long)64 , cs.total - (long long)b*64 )*27;   // LDS entries that belong to the class
	double acc[27];
#pragma unroll
	for( int w=0 ; w<27 ; w++ ) acc[w] = 0;
	GalerkinTChunks< Dim , CLS >( Mf , fMf , col0 , lim , have , wave , K , fres , cres , buf , acc , std::make_integer_sequence< int , NCH >{} );
	if( !have ) return;
#pragma unroll
	for( int w=0 ; w<27 ; w++ ) T[ ( 27*wave + w )*Nf + colBase + g ] = acc[w];
}
template< int Dim , int ... Cs >
__device__ __forceinline__ void GalerkinTDispatch( int cls , const double *Mf , float *fMf , double *T , long long Nf , int cres , long long colBase , unsigned b , double *buf , std::integer_sequence< int , Cs... > )
{
	( [&]{
		if( cls!=Cs ) return;
		ColourSet< Dim > cs ; cs.set( 2*cres , Cs );
		if( (long long)b*64<cs.total ) GalerkinTBlock< Dim , Cs >( Mf , fMf , T , Nf , cres , cs , colBase , b , buf );   // block-uniform
	}() , ... );
}
// T = P^T M of a fine level (Nf functions): blockIdx.y is the parity class of
// the fine nodes; 64 x 3^(Dim-3) threads.  fMf (optional): M rounded to float
// (the FP32 sweeps' copy, written as M passes through)
template< int Dim >
__global__ void __launch_bounds__( Dim==3 ? 64 : 192 ) KGalerkinT( const double *Mf , double *T , long long Nf , int cres , ColLayout< Dim > lay , float *fMf )
{
	__shared__ double buf[64*27];
	const int cls = blockIdx.y;
	GalerkinTDispatch< Dim >( cls , Mf , fMf , T , Nf , cres , lay.cstart[cls] , blockIdx.x , buf , std::make_integer_sequence< int , DimC< Dim >::C >{} );
}

// M_c = T P: entry (J, I = J + o) sums T[K][I] P[K][J] over the fine nodes
// K = 2J + delta (|delta_d| <= 1 inside the grid; delta_d in {-1,0,1} where
// o_d = 0, {0,1} where o_d = 1, {-1,0} where o_d = -1) in increasing K, the
// first nonzero term taken as it is (the Eigen product order).  The deltas
// are unrolled 27 at a time (delta_0 by delta_0 in 4D): a chunk's loads are
// all issued before its terms are added in order.
template< int Dim , int ... Ds >
__device__ __forceinline__ void GalerkinCChunk( const double *T , int Nf , const long long *cstart , const int *J , const int *o , int cres , int D0 , double &acc , bool &any , std::integer_sequence< int , Ds... > )
{
	// Ds: the deltas of the last three dimensions (3D: all of them) in base 3, 0 -> -1
	constexpr int R = Dim==3 ? 3 : 4;
	const int fres = 2*cres;
	double v[27];
	( [&]{
		int dl[R];
		if constexpr( Dim==4 ) dl[0] = D0;
		for( int q=0 ; q<3 ; q++ ) dl[R-3+q] = ( Ds / ( q==0 ? 9 : q==1 ? 3 : 1 ) ) % 3 - 1;
		bool ok = true;
		int cls = 0 , ci = 0 , wo = 0;
#pragma unroll
		for( int d=0 ; d<Dim ; d++ )
		{
			const int Kd = 2*J[d] + dl[d] , p = Kd&1;
			ok = ok && ( o[d]!=1 || dl[d]>=0 ) && ( o[d]!=-1 || dl[d]<=0 ) && Kd>=0 && Kd<=fres;
			cls |= p<<d;
			ci = ci*( p ? cres : cres+1 ) + ( Kd>>1 );
			wo = wo*3 + ( o[d] - ( dl[d]<0 ? -1 : 0 ) + 1 );
		}
		v[Ds] = ok ? T[ (long long)wo*Nf + cstart[cls] + ci ] : 0.;
	}() , ... );
	( [&]{
		int nz = 0;
		if constexpr( Dim==4 ) nz += D0!=0;
		nz += ( Ds/9 )%3!=1 , nz += ( Ds/3 )%3!=1 , nz += Ds%3!=1;
		const double pkj = nz==0 ? 1. : nz==1 ? 0.5 : nz==2 ? 0.25 : nz==3 ? 0.125 : 0.0625;
		const double t = v[Ds];
		if( t!=0 ){ const double pr = __dmul_rn( t , pkj ) ; acc = any ? __dadd_rn( acc , pr ) : pr ; any = true; }
	}() , ... );
}

template< int Dim >
__device__ __forceinline__ double GalerkinEntry( const double *T , int Nf , const long long *cstart , const int *J , int cres , int wi , const unsigned char *fmaskC , int j )
{
	int o[Dim];
	bool in = true;
	int ii = 0;
	{
		int t = wi;
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ o[d] = t%3 - 1 ; t /= 3 ; const int Id = J[d] + o[d] ; in = in && Id>=0 && Id<=cres; }
#pragma unroll
		for( int d=0 ; d<Dim ; d++ ) ii = ii*( cres+1 ) + J[d] + o[d];
	}
	// adapted levels: only indexed coarse functions have rows and columns
	if( in && fmaskC ) in = fmaskC[j]!=0 && fmaskC[ii]!=0;
	if( !in ) return 0.;
	double acc = 0;
	bool any = false;
	if constexpr( Dim==3 ) GalerkinCChunk< 3 >( T , Nf , cstart , J , o , cres , 0 , acc , any , std::make_integer_sequence< int , 27 >{} );
	else
	{
		// one chunk at a time (a chunk's 27 loads in flight, not 81), chunks
		// outside the delta range of o_0 or the grid skipped
#pragma unroll 1
		for( int d0=-1 ; d0<=1 ; d0++ )
		{
			if( ( o[0]==1 && d0<0 ) || ( o[0]==-1 && d0>0 ) || 2*J[0]+d0<0 || 2*J[0]+d0>2*cres ) continue;
			GalerkinCChunk< 4 >( T , Nf , cstart , J , o , cres , d0 , acc , any , std::make_integer_sequence< int , 27 >{} );
		}
	}
	return acc;
}

// Two thread maps (nc and Nf below 2^31).  perEntry: thread g is entry
// (g / 3^Dim, g % 3^Dim), for small levels (latency).  Else a workgroup is
// 64 consecutive coarse nodes (natural order, one per lane) by 4 window
// offsets at a time, so for a fixed delta a wavefront reads consecutive
// columns of one parity class of one offset plane of T.
template< int Dim >
__global__ void __launch_bounds__( 256 ) KGalerkinC( const double *T , double *Mc , int cres , ColLayout< Dim > lay , ColLayout< Dim > layF , int Nf , const unsigned char *fmaskC , int perEntry )
{
	constexpr int W = DimC< Dim >::W;
	int nc = 1 ; for( int d=0 ; d<Dim ; d++ ) nc *= cres+1;
	int j , wi0 , wstep;
	if( perEntry )
	{
		const long long g = (long long)blockIdx.x*blockDim.x + threadIdx.x;
		if( g>=(long long)nc*W ) return;
		j = (int)( g / W ) , wi0 = (int)( g % W ) , wstep = W;
	}
	else
	{
		j = (int)blockIdx.x*64 + ( threadIdx.x & 63 ) , wi0 = threadIdx.x >> 6 , wstep = 4;
		if( j>=nc ) return;
	}
	int J[Dim];
	{
		int t = j;
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- ){ const int q = t/( cres+1 ) ; J[d] = t - q*( cres+1 ) ; t = q; }
	}
	const long long outBase = lay.col( J )*W;
	for( int wi=wi0 ; wi<W ; wi+=wstep ) Mc[ outBase + wi ] = GalerkinEntry< Dim >( T , Nf , layF.cstart , J , cres , wi , fmaskC , j );
}

// All iterations of a small level in one workgroup with the coefficients (and,
// when they fit, the R and B columns) resident in LDS; barrier between colours.
template< int Dim >
__device__ __forceinline__ void GsSmallLds( const LevelView< Dim > &L , int iters , int rbInLds )
{
	constexpr int C = DimC< Dim >::C , W = DimC< Dim >::W;
	extern __shared__ double sm[];
	double *sx = sm , *sy = sm + L.N;
	for( long long i=threadIdx.x ; i<L.N ; i+=blockDim.x ) sx[i] = L.x[i] , sy[i] = L.y[i];
	LevelView< Dim > Ls = L;
	Ls.x = sx , Ls.y = sy;
	{
		double *sK = sm + 2*L.N;
		for( int i=threadIdx.x ; i<W*W ; i+=blockDim.x ) sK[i] = L.K[i];
		Ls.K = sK;
	}
	if( rbInLds )
	{
		double *sR = sm + 2*L.N + W*W , *sB = sR + L.N*W;
		for( long long i=threadIdx.x ; i<L.N*W ; i+=blockDim.x ) sR[i] = L.R[i] , sB[i] = L.B[i];
		Ls.R = sR , Ls.B = sB;
	}
	__syncthreads();
	// colour c's first column, lay.cstart[c], kept as a running sum (no dynamic index into the view)
	for( int it=0 ; it<iters ; it++ ) for( int c=0 , colStart=0 ; c<C ; c++ )
	{
		ColourSet< Dim > cs ; cs.set( L.res , c );
		for( long long g=threadIdx.x ; g<cs.total*C ; g+=blockDim.x )
		{
			const long long k = g / C;
			const int lane = (int)( g % C );
			int F[Dim];
			cs.node( k , F );
			const long long i = FlatIndex< Dim >( F , L.stride );
			double q[5];
			NodeGroup< Dim , MODE_GS >( Ls , F , i , colStart+k , lane , q );
			if( lane==0 && ( !L.fmask || L.fmask[i]==1 ) )
			{
				double dx , dy;
				if( LevelStep< Dim >( L , q , dx , dy ) ) sx[i] += dx , sy[i] += dy;
			}
		}
		colStart += (int)cs.total;
		__syncthreads();
	}
	for( long long i=threadIdx.x ; i<L.N ; i+=blockDim.x ) L.x[i] = sx[i] , L.y[i] = sy[i];
}
template< int Dim >
__global__ void __launch_bounds__( Dim==3 ? 1024 : 512 ) KGsSmallLds( LevelView< Dim > L , int iters , int rbInLds ){ GsSmallLds< Dim >( L , iters , rbInLds ); }

// ---- Tiny levels (a few elements per colour phase), one workgroup for all
// iterations.  Every element has exactly one corner of each colour, so a colour
// phase is: element stage (each element's wedge terms for the node at its
// corner of that colour, one wavefront per element) -> barrier -> node stage
// (a 16-lane row per node: its elements' contributions + its 3^Dim window
// terms + the step) -> barrier.  Coefficients, the Dirichlet rows (and R/B when
// they fit) and a colour-sorted node table live in LDS for the whole launch;
// all reductions are DPP / permlane-swap butterflies (no LDS round trips).
//
// Element stage, per element: lanes 0..2C-1 fetch the corner values relabelled
// so that the node is corner 0; lane (j, ch) contracts row j of the cubic table
// with pair chunk ch (its table slice is held in registers) over the pair
// wedges w_kl = x_k y_l - x_l y_k; the quadratic table rows and the five sums
// over j finish on lanes ch = 0, 1, 2.
template< int Dim > struct TinyShape;
template<> struct TinyShape< 3 >{ static constexpr int NCH = 8 , QC = 4; };    // 7 rows x 8 chunks of 4 pairs (28 pairs padded to 32)
template<> struct TinyShape< 4 >{ static constexpr int NCH = 4 , QC = 30; };   // 15 rows x 4 chunks of 30 pairs
template< int Dim > constexpr int TinyMaxThreads( void ){ return Dim==3 ? 1024 : 512; }   // 4D: 256 VGPRs for the 60 of the table slice

// v + v[lane'] for the DPP lane pattern CTRL (the whole wavefront must be active)
template< int CTRL > __device__ __forceinline__ double DppAdd( double v ){ return v + __builtin_amdgcn_mov_dpp( v , CTRL , 0xF , 0xF , true ); }
enum { DPP_XOR1 = 0xB1 , DPP_XOR2 = 0x4E , DPP_HALF_MIRROR = 0x141 , DPP_MIRROR = 0x140 , DPP_ROR4 = 0x124 , DPP_ROR8 = 0x128 };
// sum over the 16 lanes of each row
__device__ __forceinline__ double RowSum( double v ){ return DppAdd< DPP_MIRROR >( DppAdd< DPP_HALF_MIRROR >( DppAdd< DPP_XOR2 >( DppAdd< DPP_XOR1 >( v ) ) ) ); }
// v[l] + v[l^16] (ROWS=16) or v[l] + v[l^32] (ROWS=32) with the gfx950 permlane swaps
template< int ROWS >
__device__ __forceinline__ double SwapAdd( double v )
{
	const unsigned long long b = (unsigned long long)__double_as_longlong( v );
	const unsigned lo = (unsigned)b , hi = (unsigned)( b>>32 );
	unsigned l0 , l1 , h0 , h1;
	if constexpr( ROWS==16 )
	{
		const auto l = __builtin_amdgcn_permlane16_swap( lo , lo , false , false ) , h = __builtin_amdgcn_permlane16_swap( hi , hi , false , false );
		l0 = l[0] , l1 = l[1] , h0 = h[0] , h1 = h[1];
	}
	else
	{
		const auto l = __builtin_amdgcn_permlane32_swap( lo , lo , false , false ) , h = __builtin_amdgcn_permlane32_swap( hi , hi , false , false );
		l0 = l[0] , l1 = l[1] , h0 = h[0] , h1 = h[1];
	}
	return __longlong_as_double( (long long)( ( (unsigned long long)h0<<32 ) | l0 ) ) + __longlong_as_double( (long long)( ( (unsigned long long)h1<<32 ) | l1 ) );
}

// The step of NodeStep with the reciprocal of det from v_rcp_f64 refined by
// two Newton iterations r <- r + r (1 - det r) (instead of the IEEE division
// sequence; used on the latency-bound tiny levels).
__device__ __forceinline__ bool NodeStepFast( double q20 , double q02 , double q11 , double q10 , double q01 , double &dx , double &dy )
{
	const double H00 = 2*q20 , H11 = 2*q02 , H01 = q11;
	const double det = H00*H11 - H01*H01;
	double r = __builtin_amdgcn_rcp( det );
	r = fma( r , fma( -det , r , 1. ) , r );
	r = fma( r , fma( -det , r , 1. ) , r );
	dx = -0.5*( ( H11*q10 - H01*q01 )*r );
	dy = -0.5*( ( H00*q01 - H01*q10 )*r );
	return fabs( det )>1e-20;
}

template< int Dim > struct TinyLds
{
	static constexpr int C = DimC< Dim >::C , W = DimC< Dim >::W , P = ( C*(C-1) )/2;
	static constexpr bool Direct = Dim==3;                                     // 3D lanes form their pair wedges from the corners
	static constexpr int WB = Direct ? 0 : TinyShape< Dim >::NCH * TinyShape< Dim >::QC;
	static constexpr int PerWave = 2*C + WB;
	static constexpr int PairSlots = ( P + 63 ) / 64;
	// doubles before the node table: x, y, quadratic table, contributions,
	// per-wave scratch, then (R + dWeight K, B) per node when RB, else the
	// Dirichlet rows; rounded up to an even count
	__host__ __device__ static size_t Doubles( long long N , long long ne , int waves , bool rb )
	{
		const size_t d = 2*N + (C-1)*(C-1) + 5*ne + (size_t)waves*PerWave + ( rb ? 2*N*W : W*W );
		return ( d+1 ) & ~(size_t)1;
	}
	// bytes of dynamic LDS: the doubles, the node table (4 ints per node) and the colour starts
	static size_t Bytes( long long N , long long ne , int waves , bool rb ){ return Doubles( N , ne , waves , rb )*sizeof(double) + ( 4*N + C + 1 )*sizeof(int); }
};

// Requires: res a power of two, at most 4 * (blockDim/64) nodes per colour.
template< int Dim , bool RB >
__device__ __forceinline__ void GsTiny( const LevelView< Dim > &L , int iters )
{
	constexpr int C = DimC< Dim >::C , W = DimC< Dim >::W , J = C-1 , P = ( C*(C-1) )/2;
	constexpr int NCH = TinyShape< Dim >::NCH , QC = TinyShape< Dim >::QC;
	constexpr int PS = TinyLds< Dim >::PairSlots , WM = ( W + 15 )/16;
	constexpr int MC = (W/2)/16 , RC = (W/2)%16;   // slot and row lane of the centre window term
	constexpr bool DIRECT = TinyLds< Dim >::Direct;
	const int lane = threadIdx.x & 63 , wave = __builtin_amdgcn_readfirstlane( threadIdx.x >> 6 ) , nWaves = blockDim.x >> 6;
	const int row = lane >> 4 , rl = lane & 15;
	const int res = L.res , lres = __builtin_ctz( (unsigned)res );
	const int N = (int)L.N , ne = 1 << ( lres*Dim );
	int strideI[Dim];
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) strideI[d] = (int)L.stride[d];

	extern __shared__ double sm[];
	double *sx = sm , *sy = sx + N , *sA = sy + N , *contrib = sA + J*J , *scr = contrib + 5*ne , *tail = scr + nWaves*TinyLds< Dim >::PerWave;
	double *sRK = tail , *sB = tail + N*W , *sK = tail;
	int *nodeInfo = (int *)( sm + TinyLds< Dim >::Doubles( N , ne , nWaves , RB ) ) , *colStart = nodeInfo + 4*N;
	double *cb = scr + wave*TinyLds< Dim >::PerWave , *wb = cb + 2*C;
	for( int i=threadIdx.x ; i<N ; i+=blockDim.x ) sx[i] = L.x[i] , sy[i] = L.y[i];
	for( int i=threadIdx.x ; i<J*J ; i+=blockDim.x ) sA[i] = WA< Dim >( i/J , i%J );
	if constexpr( !RB ) for( int i=threadIdx.x ; i<W*W ; i+=blockDim.x ) sK[i] = L.K[i];
	// node table, sorted by colour: (flat index, boundary case, first incident
	// element, low/high boundary bits), bit Dim-1-d for coordinate d
	{
		auto start = [&]( int c ){ int s = 0 ; for( int c2=0 ; c2<c ; c2++ ){ ColourSet< Dim > cs ; cs.set( res , c2 ) ; s += (int)cs.total; } return s; };
		if( threadIdx.x<=C ) colStart[ threadIdx.x ] = start( threadIdx.x );
		for( int n=threadIdx.x ; n<N ; n+=blockDim.x )
		{
			int F[Dim] , t = n;
			for( int d=Dim-1 ; d>=0 ; d-- ){ F[d] = t % (res+1) ; t /= res+1; }
			int colour = 0 , cse = 0 , e0 = 0 , low = 0 , high = 0;
			for( int d=0 ; d<Dim ; d++ )
			{
				colour |= ( F[d]&1 )<<d;
				cse = cse*3 + ( F[d]==0 ? 0 : ( F[d]==res ? 2 : 1 ) );
				e0 += ( F[d]-1 ) * ( 1<<( lres*d ) );
				if( F[d]==0 ) low |= 1<<( Dim-1-d );
				if( F[d]==res ) high |= 1<<( Dim-1-d );
			}
			ColourSet< Dim > cs ; cs.set( res , colour );
			int k = 0;
			for( int d=0 ; d<Dim ; d++ ) k = k*cs.cnt[d] + ( F[d]>>1 );
			int *o = nodeInfo + 4*( start( colour ) + k );
			o[0] = n , o[1] = cse , o[2] = e0 , o[3] = low | ( high<<8 );
		}
	}
	// R + dWeight K and B by table position (the columns of R and B are in the same colour order)
	if constexpr( RB )
	{
		__syncthreads();
		for( int k=threadIdx.x ; k<N*W ; k+=blockDim.x ){ const int cse = nodeInfo[ 4*(k/W)+1 ] ; sRK[k] = L.R[k] + L.K[ cse*W + k%W ]*L.dW , sB[k] = L.B[k]; }
	}
	if constexpr( !DIRECT ) for( int p=P+lane ; p<TinyLds< Dim >::WB ; p+=64 ) wb[p] = 0;   // padding pairs stay zero

	// element stage: lane (j, ch) holds row j of the cubic table on pair chunk ch
	const int j = lane / NCH , ch = lane % NCH;
	const bool rowLane = j<J;
	double Treg[QC];
#pragma unroll
	for( int q=0 ; q<QC ; q++ ){ const int pp = ch*QC+q ; Treg[q] = ( rowLane && pp<P ) ? WT< Dim >( j , pp ) : 0.; }
	// ch 0 forms u_j = sum_k A(j,k) y_k (from the y corners), ch 1 v_j = sum_k A(j,k) x_k;
	// the quadratic table rows stay in registers in 3D and are read from LDS in 4D
	constexpr int JR = Dim==3 ? J : 1;
	double Areg[JR];
#pragma unroll
	for( int k=0 ; k<JR ; k++ ) Areg[k] = rowLane ? WA< Dim >( j , k ) : 0.;
	const double *aRow = sA + ( rowLane ? j : 0 )*J;
	const int dotBase = ( ch==0 ? C : 0 ) + 1;
	// contribution slots: lanes 0..2 store their first sum (a20, a02, b01), lanes 0..1 their second (a11, b10)
	const double f0 = ( lane==2 ? -2. : 1. )*L.sW , f1 = ( lane==0 ? -2. : 2. )*L.sW;
	// pairs: DIRECT, the lane's own chunk; otherwise the pairs lane + 64 s it stages
	constexpr int NPK = DIRECT ? QC : PS;
	int pk[NPK] , pl[NPK];
#pragma unroll
	for( int s=0 ; s<NPK ; s++ )
	{
		const int p = DIRECT ? ch*QC+s : lane + 64*s;
		pk[s] = p<P ? PK< Dim >( p ) : 0 , pl[s] = p<P ? PL< Dim >( p ) : 0;
	}
	// node stage: row lane rl takes incident element rl (rl<C) and window terms rl + 16 m
	int eoff = 0;
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) eoff += Bit< Dim >( rl , d ) * ( 1<<( lres*d ) );
	int woff[WM] , wneg[WM] , wpos[WM];
#pragma unroll
	for( int m=0 ; m<WM ; m++ )
	{
		int t = rl + 16*m;
		woff[m] = 0 , wneg[m] = 0 , wpos[m] = 0;
#pragma unroll
		for( int d=Dim-1 ; d>=0 ; d-- )
		{
			const int o = t%3 - 1 ; t /= 3;
			woff[m] += o*strideI[d];
			if( o<0 ) wneg[m] |= 1<<( Dim-1-d );
			if( o>0 ) wpos[m] |= 1<<( Dim-1-d );
		}
	}
	__syncthreads();

	for( int it=0 ; it<iters ; it++ ) for( int c=0 ; c<C ; c++ )
	{
		// ---- this wave's nodes (one per 16-lane row): table entry and window
		// values are read first, their latency hides behind the element stage
		const int c1 = colStart[c+1] , n = colStart[c] + 4*wave + row;
		const bool valid = n<c1;
		const int nn = valid ? n : 0;
		const int i = nodeInfo[4*nn] , cse = nodeInfo[4*nn+1] , e0 = nodeInfo[4*nn+2] , lh = nodeInfo[4*nn+3];
		const int low = valid ? lh & 0xFF : 0xFF , high = valid ? lh>>8 : 0xFF;   // an invalid row sees no neighbours and no elements
		double wx[WM] , wy[WM] , wR[WM] , wB[WM];
#pragma unroll
		for( int m=0 ; m<WM ; m++ )
		{
			const int w = rl + 16*m;
			const bool ok = valid && w<W && !( wneg[m] & low ) && !( wpos[m] & high );
			const int r = ok ? i + woff[m] : 0 , k = ok ? n*W + w : 0;
			const double xr = sx[r] , yr = sy[r];
			double Rv , Bv;
			if constexpr( RB ) Rv = sRK[k] , Bv = sB[k];
			else Rv = L.R[k] + sK[ ok ? cse*W+w : 0 ]*L.dW , Bv = L.B[k];
			wx[m] = ok ? xr : 0. , wy[m] = ok ? yr : 0. , wR[m] = ok ? Rv : 0. , wB[m] = ok ? Bv : 0.;
		}
		// ---- element stage
		for( int e=wave ; e<ne ; e+=nWaves )
		{
			int E[Dim] , bb = 0;
#pragma unroll
			for( int d=0 ; d<Dim ; d++ ){ E[d] = ( e>>( lres*d ) ) & ( res-1 ) ; bb |= ( ( ( c>>d ) ^ E[d] ) & 1 )<<( Dim-1-d ); }
			{
				const int a = ( lane & (C-1) ) ^ bb;
				int o = 0;
#pragma unroll
				for( int d=0 ; d<Dim ; d++ ) o += ( E[d] + Bit< Dim >( a , d ) ) * strideI[d];
				const double v = ( lane & C ) ? sy[o] : sx[o];
				if( lane<2*C ) cb[lane] = v;
			}
			WaveSync();
			double z0 = 0 , z1 = 0;
			if constexpr( DIRECT )
			{
#pragma unroll
				for( int q=0 ; q<QC ; q+=2 )
				{
					z0 = fma( Treg[q  ] , cb[ pk[q  ] ]*cb[ C+pl[q  ] ] - cb[ pl[q  ] ]*cb[ C+pk[q  ] ] , z0 );
					z1 = fma( Treg[q+1] , cb[ pk[q+1] ]*cb[ C+pl[q+1] ] - cb[ pl[q+1] ]*cb[ C+pk[q+1] ] , z1 );
				}
			}
			else
			{
#pragma unroll
				for( int s=0 ; s<PS ; s++ ) if( lane+64*s<P ) wb[ lane+64*s ] = cb[ pk[s] ]*cb[ C+pl[s] ] - cb[ pl[s] ]*cb[ C+pk[s] ];
				WaveSync();
#pragma unroll
				for( int q=0 ; q<QC ; q+=2 ){ z0 = fma( Treg[q] , wb[ ch*QC+q ] , z0 ) ; z1 = fma( Treg[q+1] , wb[ ch*QC+q+1 ] , z1 ); }
			}
			// sum over the NCH chunk lanes of row j
			double z = DppAdd< DPP_XOR2 >( DppAdd< DPP_XOR1 >( z0 + z1 ) );
			if constexpr( NCH==8 ) z = DppAdd< DPP_HALF_MIRROR >( z );
			double d0 = 0 , d1 = 0;
#pragma unroll
			for( int k=0 ; k<J ; k+=2 )
			{
				d0 = fma( Dim==3 ? Areg[ k<JR ? k : 0 ] : aRow[k] , cb[ dotBase+k ] , d0 );
				if( k+1<J ) d1 = fma( Dim==3 ? Areg[ k+1<JR ? k+1 : 0 ] : aRow[k+1] , cb[ dotBase+k+1 ] , d1 );
			}
			const double dot = d0 + d1;
			const double xj = rowLane ? cb[ j+1 ] : 0. , yj = rowLane ? cb[ C+j+1 ] : 0.;
			// ch 0: (y.u , x.u) = (a20 , a11); ch 1: (x.v , y.z) = (a02 , b10); ch 2: (x.z , 0) = (b01 , 0)
			double s0 = ch==0 ? yj*dot : ( ch==1 ? xj*dot : ( ch==2 ? xj*z : 0. ) );
			double s1 = ch==0 ? xj*dot : ( ch==1 ? yj*z : 0. );
			// sum over j: inside a 16-lane row the rows j sit NCH lanes apart, then across the four rows
			if constexpr( NCH==8 ) s0 = DppAdd< DPP_ROR8 >( s0 ) , s1 = DppAdd< DPP_ROR8 >( s1 );
			else s0 = DppAdd< DPP_ROR8 >( DppAdd< DPP_ROR4 >( s0 ) ) , s1 = DppAdd< DPP_ROR8 >( DppAdd< DPP_ROR4 >( s1 ) );
			s0 = SwapAdd< 32 >( SwapAdd< 16 >( s0 ) ) , s1 = SwapAdd< 32 >( SwapAdd< 16 >( s1 ) );
			// contrib[5e + 0..4] = (q20, q02, q01, q11, q10) of the element's colour-c node
			if( lane<3 ) contrib[ 5*e+lane ] = s0*f0;
			if( lane<2 ) contrib[ 5*e+3+lane ] = s1*f1;
			WaveSync();
		}
		// ---- window terms of the wave's nodes
		double q20 = 0 , q02 = 0 , q11 = 0 , q10 = 0 , q01 = 0;
#pragma unroll
		for( int m=0 ; m<WM ; m++ )
		{
			const double Rv = wR[m] , Bv = wB[m];
			if( m==MC && rl==RC ){ q20 += Rv ; q02 += Rv; }
			q10 += 2.*Rv*wx[m];
			q01 += 2.*Rv*wy[m];
			if( !( m==MC && rl==RC ) ){ q10 -= 2.*Bv*wy[m] ; q01 += 2.*Bv*wx[m]; }
		}
		__syncthreads();
		// ---- the contributions of the nodes' elements, the step
		{
			const bool okE = rl<C && !( ( ( ~rl & low ) | ( rl & high ) ) & (C-1) );
			const double *o = contrib + 5*( okE ? e0 + eoff : 0 );
			const double o0 = o[0] , o1 = o[1] , o2 = o[2] , o3 = o[3] , o4 = o[4];
			if( okE ) q20 += o0 , q02 += o1 , q01 += o2 , q11 += o3 , q10 += o4;
		}
		q20 = RowSum( q20 ) , q02 = RowSum( q02 ) , q11 = RowSum( q11 ) , q10 = RowSum( q10 ) , q01 = RowSum( q01 );
		double dx , dy;
		bool step;
		if( L.separate ){ NodeStepSeparate( q20 , q02 , q11 , q10 , q01 , dx , dy ) ; step = true; }
		else step = NodeStepFast( q20 , q02 , q11 , q10 , q01 , dx , dy );
		if( valid && rl==RC && step && ( !L.fmask || L.fmask[i]==1 ) ) sx[i] = wx[MC] + dx , sy[i] = wy[MC] + dy;
		__syncthreads();
	}
	for( int i=threadIdx.x ; i<N ; i+=blockDim.x ) L.x[i] = sx[i] , L.y[i] = sy[i];
}
template< int Dim , bool RB >
__global__ void __launch_bounds__( TinyMaxThreads< Dim >() ) KGsTiny( LevelView< Dim > L , int iters ){ GsTiny< Dim , RB >( L , iters ); }

// mean of the per-sample measures (deterministic): out[1] = sum / n, out[0] =
// sWeight.  The partials are staged in LDS by the whole block and summed in
// order by one thread.  nb <= 1024.
__global__ void __launch_bounds__( 256 ) KMeanMeasure( const double *partials , int nb , long long n , double sWeight , double *out )
{
	__shared__ double sh[1024];
	for( int b=threadIdx.x ; b<nb ; b+=blockDim.x ) sh[b] = partials[b];
	__syncthreads();
	if( threadIdx.x ) return;
	double s = 0;
	for( int b=0 ; b<nb ; b++ ) s += sh[b];
	out[0] = sWeight , out[1] = s / (double)n;
}
} // namespace epr

namespace epr
{
// ---- Single-wavefront levels (3D, resolution 1 or 2: one or eight elements,
// 8 or 27 nodes).  One wavefront, no barriers.  Lane (g, s) = (lane/8,
// lane%8): group g is element g (at resolution 2 its coordinates are the bits
// of g, bit 2-d for coordinate d, which is also its parity pattern e).  Every
// lane of group g keeps in registers the element's 8 corner pairs by parity
// pattern -- slot k holds the corner whose node has parity pattern k, corner
// k ^ e -- and the 28 slot wedges x_a y_b - x_b y_a (a < b); lane s also keeps
// slot s's pair.  Colour c's node in every element is the corner of parity
// pattern B = CellCorner(c), so with that node relabelled to corner 0, corner
// k of the relabelled element is slot k ^ B: the element stage of each colour
// addresses registers at compile-time positions.  A colour phase:
//   * row lane s < 7 contracts row s of the cubic table against the relabelled
//     wedges and the quadratic table against the relabelled corners, with its
//     own corner (s+1) ^ B fetched from lane (s+1) ^ B of the group;
//   * lane s adds window terms s + 8m of the node (neighbour pairs and its
//     R + dW K and B entries, from LDS) in the group that owns the node: the
//     groups sharing colour c's node differ in the pattern bits of B, the
//     owner has them 0, the others carry zero weights;
//   * the five coefficients are summed over the 8 lanes (DPP), then over the
//     groups sharing the node (lane xor 8, 16, 32 for pattern bits 0, 1, 2);
//     every butterfly step adds two equal partial sums, so all lanes of those
//     groups hold the same bits and take the same step;
//   * slot B and its 7 wedges are refreshed, and the owner's lane s = B
//     writes the node's pair to LDS for the window loads that follow.
// The cubic and quadratic table rows and the window weights are held
// pre-multiplied by sW and 2, powers of two for these levels.
template< int Dim > __host__ __device__ constexpr int CellCorner( int colour ){ int b = 0 ; for( int d=0 ; d<Dim ; d++ ) if( ( colour>>d ) & 1 ) b |= 1<<( Dim-1-d ) ; return b; }

__device__ __forceinline__ double BPermute( int addr , double v )
{
	const unsigned long long u = (unsigned long long)__double_as_longlong( v );
	const int lo = __builtin_amdgcn_ds_bpermute( addr , (int)(unsigned)u ) , hi = __builtin_amdgcn_ds_bpermute( addr , (int)(unsigned)( u>>32 ) );
	return __longlong_as_double( (long long)( ( (unsigned long long)(unsigned)hi<<32 ) | (unsigned)lo ) );
}
// sum over the 8 lanes of each half row (the whole wavefront must be active)
__device__ __forceinline__ double OctSum( double v ){ return DppAdd< DPP_HALF_MIRROR >( DppAdd< DPP_XOR2 >( DppAdd< DPP_XOR1 >( v ) ) ); }
// v + v[lane ^ 8 M'] over the bits M' of M (8, 16 and 32 for bits 0, 1, 2)
template< int M > __device__ __forceinline__ double GroupSum8( double v )
{
	if constexpr( ( M & 1 )!=0 ) v = DppAdd< DPP_ROR8 >( v );
	if constexpr( ( M & 2 )!=0 ) v = SwapAdd< 16 >( v );
	if constexpr( ( M & 4 )!=0 ) v = SwapAdd< 32 >( v );
	return v;
}
// position of slot pair (a, b), a < b, in the 28 stored wedges
__host__ __device__ constexpr int WedgeSlot( int a , int b ){ return a*8 + b - ( ( a+1 )*( a+2 ) )/2; }

template< int RES >
struct WaveState
{
	static constexpr int C = 8 , J = 7 , P = 28 , W = 27 , WM = 4 , G = RES*RES*RES , NN = ( RES+1 )*( RES+1 )*( RES+1 );
	double X[C] , Y[C];    // the group's element, by parity pattern
	double Wd[P];          // slot wedges
	double T[P] , A[J];    // row s of the cubic and quadratic tables (times sW)
	double xl , yl;        // slot s's pair
	int nbr[C][WM];        // single wavefront: window neighbours of colour c's node (LDS index; NN: the zero pad)
	unsigned nodeP[2] , csrcP[2];   // block, 8 bits per colour: colour c's node, and the element holding it as corner s
	unsigned offP;                  // block, 8 bits per window slot m: flat offset of window term s + 8m, plus 64
	int lane , s , g , own , act , myNode;
};

// z0, z1 += row terms of the relabelled wedges (even / odd pairs)
template< int B , int ... ps >
__device__ __forceinline__ void WaveContract( const double *Wd , const double *T , double &z0 , double &z1 , std::integer_sequence< int , ps... > )
{
	( [&]{
		constexpr int k = WTab< 3 >::K( ps ) ^ B , l = WTab< 3 >::L( ps ) ^ B;
		const double w = k<l ? Wd[ WedgeSlot( k , l ) ] : -Wd[ WedgeSlot( l , k ) ];
		if constexpr( ( ps & 1 )!=0 ) z1 = fma( T[ps] , w , z1 ); else z0 = fma( T[ps] , w , z0 );
	}() , ... );
}
// the 7 wedges of slot B
template< int B , int ... ks >
__device__ __forceinline__ void WaveRefresh( const double *X , const double *Y , double *Wd , std::integer_sequence< int , ks... > )
{
	( [&]{
		if constexpr( ks!=B )
		{
			constexpr int a = ks<B ? ks : B , b = ks<B ? B : ks;
			Wd[ WedgeSlot( a , b ) ] = X[a]*Y[b] - X[b]*Y[a];
		}
	}() , ... );
}

// The element stage of the colour of pattern B: row lane s's terms of the five
// coefficients (q20, q02, q11, q10, q01) of the element's colour node.
template< int B , typename St >
__device__ __forceinline__ void WaveElement( const St &S , double *q )
{
	constexpr int J = 7 , P = 28;
	// row lane s: its corner (s+1) ^ B of the relabelled element, from lane (s+1) ^ B of the group
	const int src = ( S.lane & ~7 ) | ( S.s<J ? ( ( S.s+1 ) ^ B ) : S.s );
	const double xq = BPermute( src*4 , S.xl ) , yq = BPermute( src*4 , S.yl );
	double z0 = 0 , z1 = 0;
	WaveContract< B >( S.Wd , S.T , z0 , z1 , std::make_integer_sequence< int , P >{} );
	const double z = z0 + z1;
	double u = 0 , v = 0;
#pragma unroll
	for( int k=0 ; k<J ; k++ ){ u = fma( S.A[k] , S.Y[ ( k+1 )^B ] , u ) ; v = fma( S.A[k] , S.X[ ( k+1 )^B ] , v ); }
	const double xj = S.s<J ? xq : 0. , yj = S.s<J ? yq : 0.;
	q[0] = yj*u , q[1] = xj*v , q[2] = -2.*( xj*u ) , q[3] = 2.*( yj*z ) , q[4] = -2.*( xj*z );
}
// slot B moved by (dx, dy): the copies and the slot's 7 wedges
template< int B , typename St >
__device__ __forceinline__ void WaveUpdate( St &S , double dx , double dy )
{
	S.X[B] += dx , S.Y[B] += dy;
	if( S.s==B ) S.xl += dx , S.yl += dy;
	WaveRefresh< B >( S.X , S.Y , S.Wd , std::make_integer_sequence< int , 8 >{} );
}

template< int RES , int COL >
__device__ __forceinline__ void WavePhase( const LevelView< 3 > &L , WaveState< RES > &S , double2 *sxy , const double2 *swin )
{
	using St = WaveState< RES >;
	constexpr int J = St::J , P = St::P , WM = St::WM;
	constexpr int B = CellCorner< 3 >( COL );
	// window terms: weights and neighbour pairs (written by earlier phases)
	double2 wv[WM] , nb[WM];
#pragma unroll
	for( int m=0 ; m<WM ; m++ ) wv[m] = swin[ ( COL*WM + m )*64 + S.lane ] , nb[m] = sxy[ S.nbr[COL][m] ];
	double qe[5];
	WaveElement< B >( S , qe );
	double q20 = qe[0] , q02 = qe[1] , q11 = qe[2] , q10 = qe[3] , q01 = qe[4];
#pragma unroll
	for( int m=0 ; m<WM ; m++ )
	{
		q10 = fma( wv[m].x , nb[m].x , q10 ) , q10 = fma( -wv[m].y , nb[m].y , q10 );
		q01 = fma( wv[m].x , nb[m].y , q01 ) , q01 = fma( wv[m].y , nb[m].x , q01 );
	}
	if( S.s==5 ){ q20 += 0.5*wv[1].x ; q02 += 0.5*wv[1].x; }   // the centre term, window slot 13 = 5 + 8
	q20 = OctSum( q20 ) , q02 = OctSum( q02 ) , q11 = OctSum( q11 ) , q10 = OctSum( q10 ) , q01 = OctSum( q01 );
	if constexpr( RES==2 )
		q20 = GroupSum8< B >( q20 ) , q02 = GroupSum8< B >( q02 ) , q11 = GroupSum8< B >( q11 ) , q10 = GroupSum8< B >( q10 ) , q01 = GroupSum8< B >( q01 );
	double dx , dy;
	bool step;
	if( L.separate ){ NodeStepSeparate( q20 , q02 , q11 , q10 , q01 , dx , dy ) ; step = true; }
	else step = NodeStepFast( q20 , q02 , q11 , q10 , q01 , dx , dy );
	if( step && ( ( S.act>>COL ) & 1 ) )
	{
		WaveUpdate< B >( S , dx , dy );
		if( S.s==B && ( ( S.own>>COL ) & 1 ) ) sxy[ S.myNode ] = make_double2( S.xl , S.yl );
	}
	WaveSync();
}
template< int RES , int ... cs >
__device__ __forceinline__ void WaveIteration( const LevelView< 3 > &L , WaveState< RES > &S , double2 *sxy , const double2 *swin , std::integer_sequence< int , cs... > )
{
	( WavePhase< RES , cs >( L , S , sxy , swin ) , ... );
}

// Requires a dense 3D level of resolution RES (1 or 2); one workgroup of one wavefront.
template< int RES >
__device__ __forceinline__ void GsWave( const LevelView< 3 > &L , int iters )
{
	using St = WaveState< RES >;
	constexpr int Dim = 3 , C = St::C , J = St::J , P = St::P , W = St::W , WM = St::WM , G = St::G , NN = St::NN;
	__shared__ double2 sxy[ NN+1 ];
	__shared__ double2 swin[ C*WM*64 ];
	St S;
	S.lane = threadIdx.x , S.s = S.lane & 7 , S.g = S.lane>>3;
	const int g = S.g , e = g<G ? g : 0;
	const bool live = g<G;
	int E[Dim];
#pragma unroll
	for( int d=0 ; d<Dim ; d++ ) E[d] = Bit< Dim >( e , d );
	auto flat = [&]( const int *F ){ int i = 0 ; for( int d=0 ; d<Dim ; d++ ) i += F[d]*(int)L.stride[d] ; return i; };
	auto slotNode = [&]( int k , int *F ){ for( int d=0 ; d<Dim ; d++ ) F[d] = E[d] + Bit< Dim >( k ^ e , d ) ; return flat( F ); };
	for( int i=S.lane ; i<=NN ; i+=64 ) sxy[i] = i<NN ? make_double2( L.x[i] , L.y[i] ) : make_double2( 0. , 0. );
#pragma unroll
	for( int k=0 ; k<C ; k++ ){ int F[Dim] ; const int i = slotNode( k , F ) ; S.X[k] = live ? L.x[i] : 0. , S.Y[k] = live ? L.y[i] : 0.; }
	{
		int F[Dim];
		S.myNode = slotNode( S.s , F );
		S.xl = live ? L.x[ S.myNode ] : 0. , S.yl = live ? L.y[ S.myNode ] : 0.;
	}
#pragma unroll
	for( int a=0 ; a<C ; a++ )
#pragma unroll
		for( int b=a+1 ; b<C ; b++ ) S.Wd[ WedgeSlot( a , b ) ] = S.X[a]*S.Y[b] - S.X[b]*S.Y[a];
	const int j = S.s<J ? S.s : 0;
	const double sW = L.sW;
#pragma unroll
	for( int p=0 ; p<P ; p++ ) S.T[p] = ( live && S.s<J ) ? WT< Dim >( j , p )*sW : 0.;
#pragma unroll
	for( int k=0 ; k<J ; k++ ) S.A[k] = ( live && S.s<J ) ? WA< Dim >( j , k )*sW : 0.;
	// per colour: ownership, activity, window weights and neighbours of the group's node
	S.own = 0 , S.act = 0;
#pragma unroll
	for( int c=0 ; c<C ; c++ )
	{
		const int Bc = CellCorner< Dim >( c );
		int F[Dim];
		const int n = slotNode( Bc , F );
		const bool owner = live && ( e & Bc )==0;
		if( owner ) S.own |= 1<<c;
		if( live && ( !L.fmask || L.fmask[n]==1 ) ) S.act |= 1<<c;
		int cse = 0;
		for( int d=0 ; d<Dim ; d++ ) cse = cse*3 + ( F[d]==0 ? 0 : ( F[d]==RES ? 2 : 1 ) );
		const long long col = L.lay.col( F );
#pragma unroll
		for( int m=0 ; m<WM ; m++ )
		{
			const int w = S.s + 8*m;
			int Fn[Dim] , t = w;
			bool ok = owner && w<W;
			for( int d=Dim-1 ; d>=0 ; d-- ){ Fn[d] = F[d] + t%3 - 1 ; t /= 3 ; ok = ok && Fn[d]>=0 && Fn[d]<=RES; }
			S.nbr[c][m] = ok ? flat( Fn ) : NN;
			double r = 0 , b = 0;
			if( ok )
			{
				r = 2.*( L.R[ col*W + w ] + L.K[ cse*W + w ]*L.dW );
				b = w==W/2 ? 0. : 2.*L.B[ col*W + w ];
			}
			swin[ ( c*WM + m )*64 + S.lane ] = make_double2( r , b );
		}
	}
	__syncthreads();
	for( int it=0 ; it<iters ; it++ ) WaveIteration< RES >( L , S , sxy , swin , std::make_integer_sequence< int , C >{} );
	__syncthreads();
	for( int i=S.lane ; i<NN ; i+=64 ){ const double2 q = sxy[i] ; L.x[i] = q.x , L.y[i] = q.y; }
}
template< int RES >
__global__ void __launch_bounds__( 64 ) KGsWave( LevelView< 3 > L , int iters ){ GsWave< RES >( L , iters ); }
// ---- The 3D resolution-4 level (64 elements, 125 nodes) in one workgroup of
// eight wavefronts: the element stage of the single-wavefront form (group g =
// element g, coordinates the base-4 digits of g), then one exchange through
// LDS per colour phase.  Lane 0 of each group writes the element's five
// coefficients (double-buffered by phase parity); after the barrier every
// group sharing colour c's node gathers them -- lane s the element holding
// the node as corner s -- adds the node's window terms (weights per node and
// window slot in LDS, neighbour pairs from LDS, the node's own pair from its
// registers), sums over its 8 lanes and takes the step itself, so all of them
// hold the same bits without a second exchange.  Only colour c's nodes change
// in phase c and no window of a colour-c node holds another colour-c node, so
// the owner's LDS write of the new pair races with no read.
template< int COL >
__device__ __forceinline__ void BlockPhase( const LevelView< 3 > &L , WaveState< 4 > &S , double2 *sxy , const double2 *swt , double *contrib )
{
	constexpr int B = CellCorner< 3 >( COL ) , WM = 4 , NC = 64 , NN = 125;
	constexpr int par = COL & 1;
	double q[5];
	WaveElement< B >( S , q );
#pragma unroll
	for( int k=0 ; k<5 ; k++ ) q[k] = OctSum( q[k] );
	if( S.s==0 )
	{
		double *cw = contrib + ( par*( NC+1 ) + S.g )*6;
		*(double2 *)cw = make_double2( q[0] , q[1] ) , *(double2 *)( cw+2 ) = make_double2( q[2] , q[3] ) , cw[4] = q[4];
	}
	__syncthreads();
	const int n = ( S.nodeP[COL>>2]>>( 8*( COL&3 ) ) ) & 0xFF , src = ( S.csrcP[COL>>2]>>( 8*( COL&3 ) ) ) & 0xFF;
	const double *cr = contrib + ( par*( NC+1 ) + src )*6;
	const double2 c01 = *(const double2 *)cr , c23 = *(const double2 *)( cr+2 );
	q[0] = c01.x , q[1] = c01.y , q[2] = c23.x , q[3] = c23.y , q[4] = cr[4];
	double2 wv[WM] , nb[WM];
#pragma unroll
	for( int m=0 ; m<WM ; m++ ) wv[m] = swt[ n*32 + S.s + 8*m ] , nb[m] = sxy[ min( max( n + (int)( ( S.offP>>( 8*m ) ) & 0xFF ) - 64 , 0 ) , NN-1 ) ];
	if( S.s==5 ) nb[1] = make_double2( S.X[B] , S.Y[B] );   // the node itself: window slot 13
#pragma unroll
	for( int m=0 ; m<WM ; m++ )
	{
		q[3] = fma( wv[m].x , nb[m].x , q[3] ) , q[3] = fma( -wv[m].y , nb[m].y , q[3] );
		q[4] = fma( wv[m].x , nb[m].y , q[4] ) , q[4] = fma( wv[m].y , nb[m].x , q[4] );
	}
	if( S.s==5 ){ q[0] += 0.5*wv[1].x ; q[1] += 0.5*wv[1].x; }
#pragma unroll
	for( int k=0 ; k<5 ; k++ ) q[k] = OctSum( q[k] );
	double dx , dy;
	bool step;
	if( L.separate ){ NodeStepSeparate( q[0] , q[1] , q[2] , q[3] , q[4] , dx , dy ) ; step = true; }
	else step = NodeStepFast( q[0] , q[1] , q[2] , q[3] , q[4] , dx , dy );
	if( step && ( ( S.act>>COL ) & 1 ) )
	{
		WaveUpdate< B >( S , dx , dy );
		if( S.s==B && ( ( S.own>>COL ) & 1 ) ) sxy[n] = make_double2( S.xl , S.yl );
	}
}
template< int ... cs >
__device__ __forceinline__ void BlockIteration( const LevelView< 3 > &L , WaveState< 4 > &S , double2 *sxy , const double2 *swt , double *contrib , std::integer_sequence< int , cs... > )
{
	( BlockPhase< cs >( L , S , sxy , swt , contrib ) , ... );
}
// dynamic LDS of GsBlock: the pairs (+ pad), window weights (32 slots per
// node), contributions (2 parities x 65 rows x 6 doubles)
constexpr int BlockLdsBytes = ( 126 + 125*32 )*16 + 2*65*6*8;

// Requires a dense 3D level of resolution 4; one workgroup of 512 threads and BlockLdsBytes of dynamic LDS.
__device__ __forceinline__ void GsBlock( const LevelView< 3 > &L , int iters )
{
	using St = WaveState< 4 >;
	constexpr int Dim = 3 , RES = 4 , C = 8 , J = 7 , P = 28 , W = 27 , WM = 4 , NN = 125 , NC = 64;
	extern __shared__ double2 smb[];
	double2 *sxy = smb , *swt = smb + 126;
	double *contrib = (double *)( swt + NN*32 );
	St S;
	S.lane = threadIdx.x & 63 , S.s = threadIdx.x & 7 , S.g = threadIdx.x>>3;
	const int g = S.g;
	int E[Dim];
	E[0] = g>>4 , E[1] = ( g>>2 ) & 3 , E[2] = g & 3;
	int e = 0;
	for( int d=0 ; d<Dim ; d++ ) e |= ( E[d] & 1 )<<( Dim-1-d );
	auto flat = [&]( const int *F ){ int i = 0 ; for( int d=0 ; d<Dim ; d++ ) i += F[d]*(int)L.stride[d] ; return i; };
	auto slotNode = [&]( int k , int *F ){ for( int d=0 ; d<Dim ; d++ ) F[d] = E[d] + Bit< Dim >( k ^ e , d ) ; return flat( F ); };
	// LDS: pairs, zeroed contributions, window weights by node
	for( int i=threadIdx.x ; i<=NN ; i+=blockDim.x ) sxy[i] = i<NN ? make_double2( L.x[i] , L.y[i] ) : make_double2( 0. , 0. );
	for( int i=threadIdx.x ; i<2*( NC+1 )*6 ; i+=blockDim.x ) contrib[i] = 0.;
	for( int i=threadIdx.x ; i<NN*32 ; i+=blockDim.x )
	{
		const int n = i>>5 , w = i & 31;
		int F[Dim] , Fn[Dim] , t = n , u = w;
		for( int d=Dim-1 ; d>=0 ; d-- ){ F[d] = t % ( RES+1 ) ; t /= RES+1; }
		bool ok = w<W;
		for( int d=Dim-1 ; d>=0 ; d-- ){ Fn[d] = F[d] + u%3 - 1 ; u /= 3 ; ok = ok && Fn[d]>=0 && Fn[d]<=RES; }
		double r = 0 , b = 0;
		if( ok )
		{
			int cse = 0;
			for( int d=0 ; d<Dim ; d++ ) cse = cse*3 + ( F[d]==0 ? 0 : ( F[d]==RES ? 2 : 1 ) );
			const long long col = L.lay.col( F );
			r = 2.*( L.R[ col*W + w ] + L.K[ cse*W + w ]*L.dW );
			b = w==W/2 ? 0. : 2.*L.B[ col*W + w ];
		}
		swt[i] = make_double2( r , b );
	}
	// registers: the element's corners by parity pattern, wedges, table rows
#pragma unroll
	for( int k=0 ; k<C ; k++ ){ int F[Dim] ; const int i = slotNode( k , F ) ; S.X[k] = L.x[i] , S.Y[k] = L.y[i]; }
	{
		int F[Dim];
		S.myNode = slotNode( S.s , F );
		S.xl = L.x[ S.myNode ] , S.yl = L.y[ S.myNode ];
	}
#pragma unroll
	for( int a=0 ; a<C ; a++ )
#pragma unroll
		for( int b=a+1 ; b<C ; b++ ) S.Wd[ WedgeSlot( a , b ) ] = S.X[a]*S.Y[b] - S.X[b]*S.Y[a];
	const int j = S.s<J ? S.s : 0;
	const double sW = L.sW;
#pragma unroll
	for( int p=0 ; p<P ; p++ ) S.T[p] = S.s<J ? WT< Dim >( j , p )*sW : 0.;
#pragma unroll
	for( int k=0 ; k<J ; k++ ) S.A[k] = S.s<J ? WA< Dim >( j , k )*sW : 0.;
	S.offP = 0;
#pragma unroll
	for( int m=0 ; m<WM ; m++ )
	{
		int t = S.s + 8*m , o = 0;
		for( int d=Dim-1 ; d>=0 ; d-- ){ o += ( t%3 - 1 )*(int)L.stride[d] ; t /= 3; }
		S.offP |= (unsigned)( o + 64 )<<( 8*m );   // |o| <= 31 for resolution 4
	}
	// per colour: the group's node, ownership, activity, and the element holding the node as corner s
	S.own = 0 , S.act = 0 , S.nodeP[0] = S.nodeP[1] = S.csrcP[0] = S.csrcP[1] = 0;
#pragma unroll
	for( int c=0 ; c<C ; c++ )
	{
		const int Bc = CellCorner< Dim >( c );
		int F[Dim];
		const int n = slotNode( Bc , F );
		S.nodeP[c>>2] |= (unsigned)n<<( 8*( c&3 ) );
		bool owner = true , in = true;
		int Es[Dim];
		for( int d=0 ; d<Dim ; d++ )
		{
			owner = owner && E[d]==min( F[d] , RES-1 );
			Es[d] = F[d] - Bit< Dim >( S.s , d );
			in = in && Es[d]>=0 && Es[d]<RES;
		}
		S.csrcP[c>>2] |= (unsigned)( in ? Es[0]*16 + Es[1]*4 + Es[2] : NC )<<( 8*( c&3 ) );
		if( owner ) S.own |= 1<<c;
		if( !L.fmask || L.fmask[n]==1 ) S.act |= 1<<c;
	}
	__syncthreads();
	for( int it=0 ; it<iters ; it++ ) BlockIteration( L , S , sxy , swt , contrib , std::make_integer_sequence< int , C >{} );
	__syncthreads();
	for( int i=threadIdx.x ; i<NN ; i+=blockDim.x ){ const double2 q = sxy[i] ; L.x[i] = q.x , L.y[i] = q.y; }
}
__global__ void __launch_bounds__( 512 ) KGsBlock( LevelView< 3 > L , int iters ){ GsBlock( L , iters ); }
} // namespace epr
