/*
 * CPU oracle for the ExteriorPoissonRecon solve path (dense grid, --aRadius -1,
 * and adapted hierarchies given by the reference's own tree).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch FP64 restatement of the
 * reference algorithm, written to reproduce the reference's arithmetic in the
 * reference's order wherever that order is observable.  Only tests/,
 * __graft_entry__.smoke() and the CPU-baseline leg of bench.py may load it, and
 * only as the checker; the product library never links or calls it.
 *
 * Every function in epr_oracle.cpp cites the reference file:line it follows.
 * Parity status of each piece is pinned by tests/test_oracle_golden.py against
 * vectors written by the reference itself (oracle/ref_dump.cpp).  The one
 * piece restated with a different method is the Dim x Dim symmetric
 * eigen-solve in the noise estimate (reference: Eigen 3.3.7
 * SelfAdjointEigenSolver, Include/GridSamples.inl:94); the oracle uses cyclic
 * Jacobi, equal to rounding.
 */
#ifndef EPR_ORACLE_H
#define EPR_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
	int    depth;          /* --depth            (ExteriorPoissonRecon.cpp:65)  */
	int    seed;           /* --seed             (:66)                          */
	int    tRadius;        /* --tRadius          (:67)                          */
	int    dRadius;        /* --dRadius          (:68)                          */
	int    minSolveDepth;  /* --minSolveDepth    (:70)                          */
	int    maxSolveDepth;  /* --maxSolveDepth    (:72), <0 means "= depth"      */
	int    gsIters;        /* --gsIters          (:71)                          */
	int    singleLevel;    /* --singleLevel      (:76)                          */
	double nSigma;         /* --nSigma           (:57)                          */
	double scale;          /* --scale            (:58)                          */
	double sWeight;        /* --sWeight          (:59), scales R                */
	double dWeight;        /* --dWeight          (:60), scales the stiffness    */
	double iMult;          /* --iMult            (:61)                          */
	int    separate;       /* --separate: SeparateGaussSeidelSolve (ExteriorPoissonSolvers.h:198-250) */
	double jitter;         /* --jitter           (:62, applied at :174-180)     */
} orc_params;

void   orc_default_params( orc_params *p );

/* Builds everything the solve needs (ExteriorPoissonRecon.cpp:163-236).
 * pos: n*dim world positions; skew: n*dim*dim full frame matrices m[c][r]
 * (PLY properties skew_c_r).  Returns NULL on error (message via orc_error). */
void  *orc_create( int dim , const orc_params *p , size_t n , const double *pos , const double *skew );
/* The same for an adapted hierarchy (ExteriorPoissonRecon --aRadius K,
 * Include/Hat.adaptive.inl:349-729), given as it appears in a reference .tree:
 * per level l = 0..depth concatenated, fmask holds (2^l+1)^dim function flags
 * (0 not indexed, 1 active, 2 padding; coordinate 0 most significant), emask
 * (2^l)^dim element flags (coordinate 0 least significant), and forder the
 * nforder[l] dense indices of the indexed functions in the hierarchy's
 * numbering.  Coefficients stay in the dense layout, zero outside the
 * hierarchy; only active functions are relaxed; B integrates indexed elements
 * only; c sums indexed elements only; P and the Galerkin products keep indexed
 * rows and columns and accumulate in the hierarchy's numbering. */
void  *orc_create_adapted( int dim , const orc_params *p , size_t n , const double *pos , const double *skew ,
                           const unsigned char *fmask , const unsigned char *emask , const long long *forder , const size_t *nforder );
void   orc_destroy( void *h );
const char *orc_error( void );

int    orc_dim( void *h );
size_t orc_num_functions( void *h , int level );
size_t orc_num_elements( void *h , int level );

/* Named arrays: unit_pos, skew_packed, unit_to_world, cells, cell_counts,
 * measure, noise, sample_depth, S_elem, c.  Returns the count (writes at most
 * cap values when out!=NULL). */
size_t orc_get( void *h , const char *name , double *out , size_t cap );

/* y = M x for M in {'R','B'} at a level (N_l x N_l), or 'P' (prolongation
 * from level into level+1). */
void   orc_matvec( void *h , int level , char which , const double *x , double *y );

/* (E_fit, E_reg) at a level, SystemEnergy.h:486-497 */
void   orc_energies( void *h , int level , const double *x , const double *y , double out[2] );
/* wedge square norm ||dx^dy||^2, Hat.products.inl:189-231 */
double orc_wedge( void *h , int level , const double *x , const double *y );
/* per-node quadratic fit (q20,q02,q11,q10,q01), SystemEnergy.h:542-643 */
void   orc_qfit( void *h , int level , const double *x , const double *y , double *q );
/* paired multicolour Gauss-Seidel, ExteriorPoissonSolvers.h:156-196 */
void   orc_gs( void *h , int level , double *x , double *y , int iters );
/* the random initial guess at minSolveDepth, ExteriorPoisson.h:333-341 */
void   orc_init( void *h , double *x , double *y );
/* the whole cascadic solve, ExteriorPoisson.h:316-409.  hx/hy receive the
 * hierarchical coefficients of every level concatenated (level 0 first);
 * fx/fy (may be NULL) the finest-level non-hierarchical solution; E the final
 * (E_fit, E_reg). */
void   orc_solve( void *h , double *hx , double *hy , double *fx , double *fy , double E[2] );
/* writes <hdr>.tree in the reference's byte layout (ExteriorPoissonRecon.cpp:259-270) */
int    orc_write_tree( void *h , const char *path , const double *hx , const double *hy , unsigned long long indexerHash );

/* ---- simplicial extraction (epr_levelset.cpp) ---- */
/* the canonical cube's Dim! simplices, (dim+1) vertex bit-indices each
 * (Misha/HyperCubeSimplices.h:47-170); returns the simplex count */
int    orc_cube_simplices( int dim , int *out );
/* MarchingSimplices::RegularSubGridTriangulation (orient=true/false,
 * addCenters=false), Misha/MarchingSimplices.h:622-706: V (nV x dim) and
 * S (nS x (dim+1)) are malloc'ed, release with orc_free */
int    orc_subgrid_triangulation( int dim , size_t nCells , const unsigned *cells , int orient ,
                                  size_t *nV , double **V , size_t *nS , unsigned **S );
/* MarchingSimplices::LevelSet (codim 1 or 2), Misha/MarchingSimplices.h:410-576:
 * values are nV x codim, vertex data nV x vdim; LV (nLV x vdim) and LS
 * (nLS x (dim-codim+1)) are malloc'ed */
int    orc_level_set( int dim , int codim , int vdim , size_t nV , const double *vdata , size_t nS , const unsigned *S ,
                      const double *values , const double *iso , size_t *nLV , double **LV , size_t *nLS , unsigned **LS );
void   orc_free( void *p );

/* ---- Extract pipeline (epr_extract.cpp) ---- */
/* Extract.cpp:109-259 on a .tree file: the level-set nodes (nCells x dim),
 * the finest cumulative values ((2^D+1)^dim pairs), and the oriented joint
 * zero set with 3*dim doubles per vertex (world position, frames n_0, n_1)
 * and dim-1 vertices per simplex.  all as in IsoTree::levelSetNodes.  Arrays
 * are malloc'ed (orc_free).  Nonzero on error (orc_extract_error). */
int    orc_extract( const char *treePath , int all , size_t *nCells , unsigned **cells , double **values ,
                    size_t *nV , double **V , size_t *nS , unsigned **S );
/* orc_extract with Extract.cpp:153-159, 217-238: with a density file and trim > 0,
 * the face-adjacent components whose largest samples-per-cell estimate exceeds
 * trim, in component order; compMax receives every component's maximum. */
int    orc_extract_trimmed( const char *treePath , int all , const char *densityPath , double trim ,
                            size_t *nCells , unsigned **cells , double **values ,
                            size_t *nV , double **V , size_t *nS , unsigned **S , size_t *nComp , double **compMax );
/* Extract.cpp:296-310: the codimension-1 zero set of one component, world positions */
int    orc_extract_level_set( const char *treePath , int component , size_t *nV , double **V , size_t *nS , unsigned **S );
const char *orc_extract_error( void );
unsigned long long orc_indexer_hash( int dim , int adapted );
/* MarchingSimplices::Orient of k-vertex simplices (k = 2, 3) in place */
int    orc_orient( int k , size_t nS , unsigned *simplices );

#ifdef __cplusplus
}
#endif
#endif
