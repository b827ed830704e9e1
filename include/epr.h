/*
 * epr.h -- C ABI of the MI355X (gfx950) ExteriorPoissonRecon solve path.
 *
 * One context is one reconstruction on the current HIP device.  Contexts are
 * not thread-safe; several may coexist.  All buffers in these signatures are
 * host pointers; device memory stays inside the context.  Every call returns
 * EPR_OK (0) or a negative status, with a message in epr_last_error().
 *
 * The entry points replace the reference's C++ call boundaries of the
 * reconstruction (dense grid and adapted hierarchy), the multigrid solver and
 * the Extract program; each declaration names the reference interface it
 * replaces (paths relative to the reference tree).  The few without a
 * reference counterpart (many reconstructions at once, the density grid,
 * the timing hooks) say so.
 */
#ifndef EPR_H
#define EPR_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPR_OK             0
#define EPR_ERR_ARG       -1
#define EPR_ERR_STATE     -2
#define EPR_ERR_HIP       -3
#define EPR_ERR_INPUT     -4

typedef struct epr_ctx epr_ctx;

/* Command-line parameters of ExteriorPoissonRecon/ExteriorPoissonRecon.cpp:52-78
 * (same names, same defaults). */
typedef struct
{
	int    depth;          /* --depth 5 */
	int    seed;           /* --seed 0 */
	int    tRadius;        /* --tRadius 1 */
	int    dRadius;        /* --dRadius 1 */
	int    minSolveDepth;  /* --minSolveDepth 0 */
	int    maxSolveDepth;  /* --maxSolveDepth (<0: = depth) */
	int    gsIters;        /* --gsIters 10 */
	int    singleLevel;    /* --singleLevel */
	int    verbose;        /* --verbose */
	double nSigma;         /* --nSigma 0.25 */
	double scale;          /* --scale 1.1 */
	double sWeight;        /* --sWeight 100 (screening, scales R) */
	double dWeight;        /* --dWeight 0.00625 (Dirichlet, scales the stiffness) */
	double iMult;          /* --iMult 2 */
	int    aRadius;        /* --aRadius 3: -1 selects the dense grid (HierarchicalRegularIndexer),
	                          0..5 the adapted hierarchy of that radius (HierarchicalAdaptedIndexer,
	                          Include/Hat.adaptive.h:208-331) */
	int    separate;       /* --separate: SeparateGaussSeidelSolve (ExteriorPoissonSolvers.h:198-250)
	                          instead of PairedGaussSeidelSolve */
	double jitter;         /* --jitter 0: a random translation of the unit-cube samples
	                          (ExteriorPoissonRecon.cpp:174-180) */
	int    fp32Sweeps;     /* --fp32 (no reference counterpart, default 0): on dense levels of at
	                          least 32768 functions swept by colour launches, every iteration of a
	                          level but its last runs in FP32 (coefficients and R, B, K rounded to
	                          float); the last runs in FP64 from the widened iterate; smaller levels
	                          and the energies stay FP64 */
} epr_params;

/* Fills the reference defaults (ExteriorPoissonRecon.cpp:56-72). */
void epr_default_params( epr_params *p );

/* Replaces the construction of ExteriorPoisson::Reconstructor<Dim,Energy>
 * (ExteriorPoisson/ExteriorPoisson.h:56-72, 126-313); dim is 3 or 4. */
int  epr_create( int dim , const epr_params *p , epr_ctx **out );
void epr_destroy( epr_ctx *ctx );
const char *epr_last_error( const epr_ctx *ctx );   /* NULL: the calling thread's last context-free (mesh) error */

/* Hands over the samples read by ReadSamples<Dim>
 * (ExteriorPoissonRecon/ExteriorPoissonRecon.cpp:282-303): pos is n x dim
 * world positions, skew is n x dim x dim full frame matrices m[c][r] (PLY
 * properties skew_c_r).  The unit-cube map of ExteriorPoisson.h:100-123 is
 * applied here. */
int  epr_set_samples( epr_ctx *ctx , size_t n , const double *pos , const double *skew );

/* Everything the reference does between reading the samples and solving:
 * OrderedSampler (Include/OrderedSamples.inl:32-85), TreeEstimator
 * (Include/GridSamples.inl:740-896), the skew splat and target field
 * (ExteriorPoisson.h:175-264, 290-309), R = deltaMass (Hat.scalars.inl:976-1063),
 * B = stiffness(S) (Hat.scalars.inl:1584-1690) and the Galerkin chain
 * (Include/SystemEnergy.h:464-477).  With aRadius >= 0 it first builds the
 * adapted hierarchy (HierarchicalAdaptedIndexer's constructor,
 * Include/Hat.adaptive.inl:349-376, 587-729): levels keep their dense node
 * layout, functions outside the hierarchy hold 0, and only the active
 * functions are updated. */
int  epr_build( epr_ctx *ctx );

/* Sets the coefficients of one level; x0/y0 NULL selects the reference's
 * random initial guess at minSolveDepth (ExteriorPoisson.h:333-341: glibc's
 * rand() sequence after srand(seed), drawn from a private copy of that
 * generator, so the process-wide rand() state is neither read nor changed);
 * the random start also zeroes the levels above `level`. */
int  epr_init( epr_ctx *ctx , int level , const double *x0 , const double *y0 );

/* iters paired multicolour Gauss-Seidel iterations at one level
 * (ExteriorPoisson/ExteriorPoissonSolvers.h:156-196). */
int  epr_gs_sweeps( epr_ctx *ctx , int level , int iters );

/* One colour of one of those iterations: every active node of colour
 * `colour` (0 .. 2^dim-1, the parity colour of ExteriorPoisson.h:161-169)
 * takes its step -- the body of the colour loop of PairedGaussSeidelSolve
 * (ExteriorPoissonSolvers.h:163-185).  Levels swept by one-workgroup kernels
 * (the few coarsest) are an error. */
int  epr_gs_colour( epr_ctx *ctx , int level , int colour );

/* x[level+1] += P x[level], y likewise (ExteriorPoisson.h:386). */
int  epr_prolong_add( epr_ctx *ctx , int level );

/* (E_fit, E_reg) of the current coefficients of a level
 * (Include/SystemEnergy.h:486-497). */
int  epr_energies( epr_ctx *ctx , int level , double out[2] );

/* epr_energies without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a static 4 MiB device array, one per device, shared by
 * every call on that device and kept until the process ends: calls that
 * overlap on one device race.  Same kernels and summation order as
 * epr_energies (same result bits).  Levels whose partial sums do not fit --
 * more than 131071 * 256 / 2^dim functions (3D up to depth 7, 4D up to
 * depth 5 on the dense grid) -- are an error (EPR_ERR_STATE, with the limit
 * in epr_last_error). */
int  epr_energies_noalloc( epr_ctx *ctx , int level , double out[2] );

/* The whole cascadic schedule with the random initial guess
 * (ExteriorPoisson.h:316-409); afterwards the coefficients are hierarchical
 * as in the reference's .tree output. out_E (may be NULL) receives the final
 * (E_fit, E_reg) at maxSolveDepth. */
int  epr_solve( epr_ctx *ctx , double out_E[2] );

/* Many independent reconstructions.  No counterpart in the reference, which
 * runs one reconstruction per process (ExteriorPoissonRecon.cpp:201-279);
 * these are for callers with many small inputs, whose coarse levels alone
 * would leave most of the GPU idle.  Context-free errors (epr_last_error(NULL)).
 *
 * epr_build_batch: epr_build on each of n contexts from up to `threads` host
 * threads (0: 4, one per hardware queue of the HIP runtime's default), each
 * context's kernels on its own stream.
 *
 * epr_solve_batch: epr_solve on each of n built contexts of one dimension,
 * with every level's sweeps of all contexts in shared launches (one
 * workgroup per context on the coarse levels, one grid over all contexts'
 * nodes of a colour on the others; contexts with fp32Sweeps take their float
 * iterations in shared float grids).  Each context ends with exactly the
 * coefficients and energies its own epr_solve gives (with the same tuned
 * plans); energies (may be NULL) receives n x 2 values (E_fit, E_reg).
 * --verbose prints nothing here.  Contexts of equal depth are solved
 * together, the groups side by side on their own streams; the whole device
 * work is recorded into one HIP graph on the first call and replayed by later
 * calls on the same contexts while their buffers, counts, parameters and
 * plans are unchanged (only the target-field constants are uploaded per
 * call); at most 32767 contexts. */
int  epr_build_batch( size_t n , epr_ctx *const *ctxs , int threads );
int  epr_solve_batch( size_t n , epr_ctx *const *ctxs , double *energies );

/* Autotuning of the sweep plans (no reference counterpart: the reference has
 * one CPU loop per level, ExteriorPoissonSolvers.h:156-196).  A plan is, per
 * dimension and level resolution, the sweep kernel of the level (one of the
 * one-workgroup kernels, or one launch per colour) and the workgroup size of
 * its colour launches (64, 128 or 256 threads); one table per device, kept
 * until the process ends, consulted whenever a context plans a level.
 *
 * epr_autotune: times every candidate plan on each level of this built
 * context's solve (on its own data; x and y are restored) and keeps the
 * fastest for single-context solves; flags bit 0 also tunes the FP32 colour
 * launches.  report (may be NULL) receives one line per trial, "level res
 * nodes precision kind threads us_per_iteration", kept ones ending in " *",
 * truncated to cap bytes including the terminating NUL.
 *
 * epr_autotune_batch: the same for epr_solve_batch on these contexts: per
 * level, the sweeps of every context with that level under each candidate
 * ("level res nodes contexts precision kind threads us" lines).
 *
 * epr_tune_get: the current device's table as text, one line per entry
 * "dim res batched fp32 kind threads us" (kind -1: the built-in rule); *len
 * (may be NULL) receives the full length, buf (may be NULL) the text
 * truncated to cap bytes including the NUL.  epr_tune_set: replaces the
 * current device's table with such a text (NULL or "": the built-in rules). */
int  epr_autotune( epr_ctx *ctx , int flags , char *report , size_t cap );
int  epr_autotune_batch( size_t n , epr_ctx *const *ctxs , int flags , char *report , size_t cap );
int  epr_tune_get( char *buf , size_t cap , size_t *len );
int  epr_tune_set( const char *text );

/* Number of hat functions (2^level+1)^dim of a level (the dense layout in
 * which coefficients are exchanged, also for adapted reconstructions). */
size_t epr_num_functions( const epr_ctx *ctx , int level );

/* Copies the coefficients of one level to/from the host (Reconstructor::x/y,
 * ExteriorPoisson.h:83-86). */
int  epr_get_coeffs( epr_ctx *ctx , int level , double *x , double *y );
int  epr_set_coeffs( epr_ctx *ctx , int level , const double *x , const double *y );

/* Per-node quadratic fits (q20,q02,q11,q10,q01) of a level at its current
 * coefficients (Include/SystemEnergy.h:542-643); q holds 5*N doubles. */
int  epr_quadratic_fits( epr_ctx *ctx , int level , double *q );

/* y = M x for M = 'R' or 'B' of a level, or 'P' (level -> level+1)
 * (SystemEnergy.h:160-176 operators). */
int  epr_apply( epr_ctx *ctx , int level , char which , const double *x , double *y );

/* Named build products: unit_pos, cells, cell_counts, measure, noise,
 * sample_depth, S_elem, c, unit_to_world; for adapted reconstructions also
 * adapted_fmask_<d> (0 not indexed, 1 active, 2 padding, per dense node),
 * adapted_emask_<d>, adapted_forder_<d> and adapted_eorder_<d> (dense index of
 * each function / element in the hierarchy's numbering) and adapted_nactive_<d>.
 * Writes min(count,cap) values when out is not NULL and returns the count in
 * *count. */
int  epr_get_array( epr_ctx *ctx , const char *name , double *out , size_t cap , size_t *count );

/* Writes <path> in the reference's .tree layout
 * (ExteriorPoissonRecon.cpp:259-270) from the hierarchical coefficients left
 * by epr_solve; indexer_hash is the reference build's
 * typeid(HierarchicalRegularIndexer<Dim>).hash_code() or, for adapted
 * reconstructions, typeid(HierarchicalAdaptedIndexer<Dim>).hash_code().
 * Adapted trees carry the coefficients in the hierarchy's numbering followed
 * by HierarchicalAdaptedIndexer::write (Include/Hat.adaptive.inl:408-414). */
int  epr_write_tree( epr_ctx *ctx , const char *path , unsigned long long indexer_hash );

/* Context-free: builds the adapted hierarchy of n world positions (the
 * HierarchicalAdaptedIndexer constructor, Include/Hat.adaptive.inl:349-376,
 * 587-729, after the unit-cube map with the given bounding-box scale) and
 * writes its serialisation (Hat.adaptive.inl:408-414) to path.  Host only. */
int  epr_adapted_indexer_write( int dim , int depth , int aRadius , double scale , size_t n , const double *pos , const char *path );

/* Writes the finest-level (phi1,phi2) values as a .grid file
 * (Misha/RegularGrid.inl:440-471 layout, 2 DOUBLE values per node). */
int  epr_write_grid( epr_ctx *ctx , const char *path );

/* Writes the density estimator as <out>.density.tree
 * (ExteriorPoissonRecon.cpp:272-277: the voxel-to-world XForm, then
 * TreeEstimator::write, Include/GridSamples.inl:734-738): the estimator tree
 * with the reference's node set and depth-first order, each node's
 * DensityAndNoiseInfo statistics, and the target samples per cell.  Extract
 * reads it with --density (epr_iso_read_density). */
int  epr_write_density( epr_ctx *ctx , const char *path );

/* Measurement hook (no reference counterpart): runs iters Gauss-Seidel
 * iterations at one level bracketed by HIP events on the context's stream and
 * returns the elapsed device milliseconds and the number of kernel launches. */
int  epr_time_gs_sweeps( epr_ctx *ctx , int level , int iters , double *ms , int *launches );

/* Measurement hook (no reference counterpart): the same iterations computed
 * in FP32 -- the level's coefficients and operators rounded to float, the
 * colour launches instantiated for float, timed like epr_time_gs_sweeps --
 * with the float coefficients written back to the level (so epr_energies
 * gives the FP32 iterate's FP64 energy).  Dense levels swept by colour
 * launches only. */
int  epr_time_gs_sweeps_fp32( epr_ctx *ctx , int level , int iters , double *ms , int *launches );

/* ---- Simplicial extraction (the reference's Extract, Extract/Extract.cpp:161-208) ----
 * A mesh handle holds vertices (nVertices x vertexDim doubles) and simplices
 * (nSimplices x simplexVertices indices) in host memory. */
typedef struct epr_mesh epr_mesh;
int  epr_mesh_info( const epr_mesh *m , size_t *nVertices , int *vertexDim , size_t *nSimplices , int *simplexVertices );
int  epr_mesh_copy( const epr_mesh *m , double *vertices , unsigned *simplices );
void epr_mesh_destroy( epr_mesh *m );

/* Replaces MarchingSimplices::RegularSubGridTriangulation< dim >( nCells , F ,
 * orient , false ) (Misha/MarchingSimplices.h:622-706): the dim! simplices of
 * every listed cell (cells: nCells x dim unsigned offsets), vertices numbered
 * in order of first appearance, positions as doubles. */
int  epr_subgrid_triangulation( int dim , size_t nCells , const unsigned *cells , int orient , epr_mesh **out );

/* Replaces MarchingSimplices::LevelSet< double , codim , dim >( mesh , PF , iso )
 * (Misha/MarchingSimplices.h:410-576), codim 1 or 2: the (dim-codim)-simplices
 * of the zero set of codim values per vertex (values: nV x codim; iso NULL =
 * 0) over a mesh of dim-simplices (simplices: nS x (dim+1)); output vertices
 * carry the interpolation of the input vertex data (nV x vdim), numbered and
 * ordered as the reference numbers and orders them.  Codim 2 is the joint zero
 * set of Extract.cpp:205; codim 1 the per-function sets of its --levelSets
 * output (:303-309). */
int  epr_level_set( int dim , int codim , int vdim , size_t nV , const double *vdata , size_t nS , const unsigned *simplices ,
                    const double *values , const double *iso , epr_mesh **out );

/* ---- Multigrid V-cycle (the reference's Include/MG.h) ----
 * Sparse matrices are CSC with sorted row indices (Eigen's storage). */
typedef struct epr_mg epr_mg;

/* Replaces MGSolver( M , Ps , pNum ) (Include/MG.h:83-108): M is n x n;
 * prolongation i (i = 0..pNum-1) is pRows[i] x pCols[i] and maps level i to
 * level i+1, level pNum being M's.  The Galerkin matrices
 * M_i = P_i^T M_{i+1} P_i are formed on the device; the coarsest (at most
 * 4096 unknowns, symmetric positive definite) is inverted there. */
int  epr_mg_create( size_t n , const unsigned *colptr , const unsigned *rowidx , const double *values ,
                    int pNum , const size_t *pRows , const size_t *pCols ,
                    const unsigned *const *pColptr , const unsigned *const *pRowidx , const double *const *pValues , epr_mg **out );
int  epr_mg_num_levels( const epr_mg *mg );
size_t epr_mg_level_size( const epr_mg *mg , int level );

/* The colour lists MCI(level) of MGSolver::vCycle's multicolour form
 * (MG.h:146-186) for a level 1..pNum: colour c holds
 * indices[starts[c] .. starts[c+1]). */
int  epr_mg_set_colours( epr_mg *mg , int level , int nColours , const unsigned *starts , const unsigned *indices );

/* MGSolver::vCycle( b , x , iters , MCI ) (MG.h:146-186); x is updated in place. */
int  epr_mg_vcycle( epr_mg *mg , const double *b , double *x , unsigned iters );

/* iters multicolour GSRelaxer sweeps (MG.h:62-79) of a level's matrix. */
int  epr_mg_relax( epr_mg *mg , int level , const double *b , double *x , unsigned iters );
void epr_mg_destroy( epr_mg *mg );

/* Measurement hook (no reference counterpart): a plain device-to-device copy
 * kernel over two buffers of `bytes` bytes on the current device, run reps
 * times after one warm-up launch; returns the mean milliseconds per copy.  It
 * calibrates the HBM byte counters (FETCH_SIZE, WRITE_SIZE) against a kernel
 * of known traffic and gives the achievable copy bandwidth. */
int  epr_time_copy( size_t bytes , int reps , double *ms );

/* The reference's synthetic input generator, Sample/Sample.cpp over
 * Include/Samples.h (host only, no device needed): `type` as Sample's --type
 * (curves in 3D: torus_knot:<p>:<q>, spiral:<r>, borromean_rings,
 * spherical:1:<n>:<l>; surfaces in 4D: clifford_torus, hopf_torus:<n>:<l>),
 * res samples of a curve or res^2 of a surface, at the parameters Samples.h's
 * GetSamples draws (random from std::default_random_engine unless regular),
 * mapped to the unit cube (Sample.cpp:87-118), with positional (pNoise),
 * angular (aNoise, degrees) and spot (sNoise) noise (Sample.cpp:120-201),
 * mapped back.  *dim and *count receive the dimension and the sample count;
 * when pos and skew are given and cap >= *count, pos receives count x dim
 * positions and skew count x dim x dim frame matrices (entry (c, r) at
 * [i][c][r], the reference PLY's skew_c_r), bit for bit the values the
 * reference writes.  Other types are an error ("Could not parse sample
 * type"). */
int  epr_sample_generate( const char *type , unsigned int res , int regular , double aNoise , double pNoise , double sNoise ,
                          int *dim , size_t *count , double *pos , double *skew , size_t cap );

/* Measurement hooks (no reference counterpart), for the HBM roofline of the
 * transfer and energy kernels and the batched sweep.
 *
 * epr_time_op: reps launches (after one warm-up) of one kernel of a built
 * context at `level`, on scratch outputs (the context's data is unchanged):
 * "prolong" (KProlong2, level -> level+1, x and y: out += P in), "galerkin_t"
 * (KGalerkinT, T = P^T R of the level), "galerkin_c" (KGalerkinC, the coarser
 * level's R from T), "energy" (KEnergy + its fixed-order row sums).  *ms
 * receives milliseconds per launch, *bytes the launch's algorithmic HBM
 * bytes (every input and output value once).
 *
 * epr_time_batch_sweeps: iters Gauss-Seidel iterations of `level` as
 * epr_solve_batch runs them on colour grids (one grid per colour over every
 * dense context of the batch with that level), FP32 on the float copies when
 * fp32 is set (the FP64 iterate is left as it is), else FP64 with x and y
 * restored afterwards; *ms all launches, *launches their number, *nodes the
 * node updates of one iteration summed over the contexts. */
int  epr_time_op( epr_ctx *ctx , const char *op , int level , int reps , double *ms , double *bytes );
int  epr_time_batch_sweeps( size_t n , epr_ctx *const *ctxs , int level , int fp32 , int iters , double *ms , int *launches , long long *nodes );

/* Wall-clock seconds of the last epr_build / epr_solve (device-synchronised). */
double epr_last_seconds( const epr_ctx *ctx , int which /* 0 build, 1 solve */ );

/* ---- Surface extraction from a .tree (the reference's Extract/Extract.cpp:96-330) ----
 * An iso handle holds a solved hierarchy read from a .tree file (dense or
 * adapted): the finest regular values of both functions in device memory,
 * recomposed as the reference's IsoTree does (Include/IsoTree.inl:45-73,
 * 314-338), and the indexed finest elements of an adapted hierarchy. */
typedef struct epr_iso epr_iso;

/* Replaces Extract.cpp:109-129 (read dim, indexer hash, voxel-to-world, the
 * hierarchical x and y, the indexer; IsoTree construction).  Unrecognised
 * hashes and dimensions other than 3 and 4 are errors, as Extract.cpp:352-363
 * makes them.  Errors: epr_iso_last_error( NULL ). */
int  epr_iso_open( const char *treePath , epr_iso **out );
const char *epr_iso_last_error( const epr_iso *iso );
/* depth = the finest depth D; xform (optional) receives the (dim+1)^2 voxel-to-world coordinates as stored. */
int  epr_iso_info( const epr_iso *iso , int *dim , int *depth , int *adapted , double xform[25] );
/* The finest regular values, (2^D+1)^dim pairs (phi1, phi2), last coordinate fastest. */
int  epr_iso_values( epr_iso *iso , double *out );

/* Replaces IsoTree::refineZeroCrossing + levelSetNodes( all ) (IsoTree.inl:165-234):
 * the finest elements of the level set, in tree order, as dim offsets each.
 * cells may be NULL (count only); at most cap cells are copied. */
int  epr_iso_level_set_nodes( epr_iso *iso , int all , size_t cap , unsigned *cells , size_t *count );

/* Replaces Extract.cpp:161-259 up to the PLY write: triangulation of the
 * level-set nodes, vertex values, the joint zero set, MarchingSimplices::Orient,
 * per-vertex gradient frames and the voxel-to-world map.  The mesh has
 * 3*dim doubles per vertex (world position, then the frame n_c_d for c = 0, 1,
 * d = 0..dim-1) and dim-1 vertices per simplex.  all = 1 for the mesh output
 * alone, 0 when the --levelSets output is also requested (Extract.cpp:129, 161). */
int  epr_iso_extract( epr_iso *iso , int all , epr_mesh **out );

/* Replaces Extract.cpp:153-159 (TreeEstimator::read): a density file as the
 * reference's ExteriorPoissonRecon writes it (<out>.density.tree). */
int  epr_iso_read_density( epr_iso *iso , const char *densityPath );

/* epr_iso_extract followed by Extract.cpp:217-238 when a density file has been
 * read and trimDensity > 0: the face-adjacent components of the oriented set
 * (MarchingSimplices::FaceAdjacentConnectedComponents) whose largest
 * samples-per-cell estimate over their vertices exceeds trimDensity are kept,
 * component by component in the reference's traversal order. */
int  epr_iso_extract_trimmed( epr_iso *iso , int all , double trimDensity , epr_mesh **out );
/* The largest samples-per-cell estimate of every component of the last trimmed
 * extraction, in component order (the values Extract --verbose prints). */
int  epr_iso_component_densities( const epr_iso *iso , size_t cap , double *out , size_t *count );

/* Context-free: the <out>.density.grid output (the reference's README lists
 * .grid / .density.grid outputs of an older version; its current code writes
 * only the .density.tree this reads).  The density estimator's samples per
 * cell (TreeEstimator::samplesPerCell, Include/GridSamples.inl:845-850, the
 * quantity Extract --trimDensity thresholds) at every node of the
 * (2^depth+1)^dim grid, in the function index order, written as a .grid with
 * one DOUBLE per node (Misha/RegularGrid.inl:440-471 layout, the density
 * file's voxel-to-world transform in the header). */
int  epr_density_grid( const char *densityPath , int dim , int depth , const char *gridPath );

/* Replaces Extract.cpp:296-310: the codimension-1 zero set of one component
 * (0 or 1) over the all = 0 triangulation, dim world coordinates per vertex,
 * dim vertices per simplex. */
int  epr_iso_level_set( epr_iso *iso , int component , epr_mesh **out );

/* Replaces Extract.cpp:132-144: header.0.grid and header.1.grid of the finest
 * regular values (RegularGrid<dim,double>::write, Misha/RegularGrid.inl:441-471). */
int  epr_iso_write_grid( epr_iso *iso , const char *header );

/* Seconds of the last open (0), level-set nodes and triangulation (1), level
 * set (2), orientation and trimming (3), frames (4); no reference counterpart. */
double epr_iso_seconds( const epr_iso *iso , int which );
void epr_iso_destroy( epr_iso *iso );

/* Replaces MarchingSimplices::Orient( simplexIndices ) (Misha/MarchingSimplices.h:179-246)
 * for k-vertex simplices, k = 2 (segments) or 3 (triangles), in place: the
 * faces' incidences are paired (a face shared by three simplices is an error)
 * and each face-connected component is oriented like its first simplex, with
 * the reference's walk order deciding on non-orientable components. */
int  epr_orient( int k , size_t nS , unsigned *simplices );

/* ---- Learned initializer ----
 * No counterpart in the reference.  A convolutional U-Net on the hat-function
 * node grid of one coarse level L, (2^L+1)^dim nodes (four grids of 2^L+1,
 * 2^(L-1)+1, 2^(L-2)+1, 2^(L-3)+1 nodes per side, 16/32/64/128 channels,
 * 3^dim kernels, BF16 storage and FP32 accumulation on the device's matrix
 * cores), maps the target field averaged onto that level to the level's
 * coefficients (x, y); they replace the random initial guess
 * (ExteriorPoisson.h:333-341) and the reference's cascade refines them from
 * level L on.  Errors: epr_unet_last_error( net ) (NULL: the calling
 * thread's last error of epr_unet_create). */
typedef struct epr_unet epr_unet;
int  epr_unet_create( int dim , int level , int inChannels , epr_unet **out );   /* inChannels: dim*(dim-1)/2 for epr_learned_features */
void epr_unet_destroy( epr_unet *net );
const char *epr_unet_last_error( const epr_unet *net );
size_t epr_unet_num_params( const epr_unet *net );
size_t epr_unet_num_nodes( const epr_unet *net );   /* (2^level+1)^dim */
/* Parameters in the canonical order: per convolution (encoder 2 per grid,
 * finest first; decoder 2 per grid, coarsest first; the output) the weights
 * W[tap][out][in] then the biases[out]; tap = base-3 offset index, first
 * coordinate most significant.  set resets the optimizer. */
int  epr_unet_set_params( epr_unet *net , const float *params );
int  epr_unet_get_params( epr_unet *net , float *params );
/* in: batch x nodes x inChannels, out: batch x nodes x 2 (host). */
int  epr_unet_forward( epr_unet *net , int batch , const float *in , float *out );
/* The training loss -- per sample the mean squared distance to the target
 * after the best rotation of the (x, y) plane, averaged over the batch -- and
 * its gradient with respect to every parameter (canonical order; grad may be
 * NULL).  target: batch x nodes x 2. */
int  epr_unet_loss_gradients( epr_unet *net , int batch , const float *in , const float *target , double *loss , float *grad );
/* Training on the device: the set is uploaded once; each step takes `batch`
 * samples of a seeded shuffle, runs the forward and backward passes and an
 * Adam update (learning rate lr); losses (may be NULL) receives every step's
 * loss. */
int  epr_unet_set_dataset( epr_unet *net , size_t count , const float *in , const float *target );
int  epr_unet_train( epr_unet *net , int steps , int batch , double lr , unsigned seed , double *losses );

/* The network input of a built reconstruction at a level: the finest target
 * field averaged over each element of the level, then over each node's
 * incident elements, divided by its root mean square; out: nodes x C(dim,2). */
int  epr_learned_features( epr_ctx *ctx , int level , float *out );
/* Sets the coefficients of a level from the network (the reference's random
 * start replaced): its (x, y), y negated when that lowers the energy, scaled by
 * the factor that minimises the fitting energy E_fit = |grad x ^ grad y - S|^2;
 * finer levels are zeroed.  info (may be NULL): the factor, the negation
 * (0/1), the level's energy afterwards.  A prediction with no
 * fitting-energy-lowering scale is an error (use epr_init's random start). */
int  epr_init_learned( epr_ctx *ctx , epr_unet *net , int level , double info[3] );
/* epr_solve with the cascade started at the network's level from its
 * prediction instead of the random start at minSolveDepth. */
int  epr_solve_learned( epr_ctx *ctx , epr_unet *net , double out_E[2] );
/* epr_solve (net NULL) or epr_solve_learned (net given) with the iteration
 * count of every level d of the cascade at iters[d] (depth+1 entries; NULL:
 * the reference's ceil( iMult^(depth-d) gsIters )).  No reference
 * counterpart: the learned start needs far fewer coarse iterations than the
 * reference's schedule gives its random start. */
int  epr_solve_schedule( epr_ctx *ctx , epr_unet *net , const int *iters , double out_E[2] );

/* typeid( MishaK::Hat::Hierarchical{Regular,Adapted}Indexer< dim > ).hash_code()
 * of a libstdc++ build of the reference: the value a .tree carries after its
 * dimension (ExteriorPoissonRecon.cpp:263). */
unsigned long long epr_indexer_hash( int dim , int adapted );

#ifdef __cplusplus
}
#endif
#endif
