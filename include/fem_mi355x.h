/*
 * fem_mi355x.h -- C ABI of the MI355X (gfx950) P1 FEM assembly + AMG-PCG library.
 *
 * Plain C: opaque handle, POD parameter struct, pointers + sizes, integer
 * status codes.  No function in this library calls exit(); every failure is a
 * non-zero status with a message in fem_last_error().
 *
 * Each entry point names the reference interface it replaces
 * (SCIInstitute/SCI-Solver_FEM, paths relative to the repository root).
 *
 * Threading: calls on one handle must be serialised by the caller; distinct
 * handles are independent.  All device work of a handle is queued on one HIP
 * stream (fem_set_stream), and every call returns only after that work is done
 * (except fem_apply_preconditioner_noalloc, which only enqueues it).
 */
#ifndef FEM_MI355X_H
#define FEM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum fem_status {
  FEM_OK = 0,
  FEM_ERR_ARG = 1,          /* invalid argument / inconsistent sizes        */
  FEM_ERR_IO = 2,           /* file missing or malformed                    */
  FEM_ERR_HIP = 3,          /* HIP runtime error                            */
  FEM_ERR_EMPTY = 4,        /* no matrix (reference: std::invalid_argument) */
  FEM_ERR_NOT_CONVERGED = 5,/* solve ended at maxIters above tolerance      */
  FEM_ERR_NUMERIC = 6,      /* non-positive diagonal, NaN, coarse not SPD   */
  FEM_ERR_STATE = 7,        /* call out of order (e.g. grad before solve)   */
  FEM_ERR_CAPACITY = 8      /* a kernel capacity limit was exceeded         */
};

typedef struct fem_handle fem_handle;

/* ---- solver parameters -----------------------------------------------------
 * Field names and defaults of the reference's public FEMSolver members,
 * src/FEMSolver.h:40-61, defaults src/FEMSolver.cu:11-33.  Fields marked
 * "accepted, no effect" are read by nothing in the reference either
 * (src/core/cuda/amg_level.cu:51,106; src/core/cuda/cgcycle.cu:7;
 * src/FEMSolver.cu:99-129).  Fields after `blockSize` are additions. */
typedef struct fem_params {
  int verbose;            /* verbose_                                        */
  int maxLevels;          /* maxLevels_          100                          */
  int maxIters;           /* maxIters_           100                          */
  int preInnerIters;      /* preInnerIters_      5  pre-smoothing sweeps      */
  int postInnerIters;     /* postInnerIters_     5  post-smoothing sweeps     */
  int postRelaxes;        /* postRelaxes_        1  post sweeps x this count  */
  int cycleIters;         /* cycleIters_         1  accepted, no effect       */
  int dsType;             /* dsType_             0  accepted, no effect       */
  int topSize;            /* topSize_            256 coarsest level < this    */
  int randMisParameters;  /* randMisParameters_  90102 digits: min aggregate
                             size (x/10000%10), coarse depth, fine MIS depth */
  int partitionMaxSize;   /* partitionMaxSize_   512 accepted, no effect      */
  int aggregatorType;     /* aggregatorType_     0  randomized MIS-2; 3 AggMIS:
                             MIS, aggregate-to-nearest, merge-split to sizes
                             [p/1000%1000, p%1000] of randMisParameters p
                             (MIS depth p/10^7%10, 0 = 2); 4 / 5: host
                             flood-fill MIS of depth p/10^7%10, nearest-root
                             aggregation, merge-split for 4 only; 1, 2 (METIS):
                             not built (FEM_ERR_ARG)                          */
  int convergeType;       /* convergeType_       0  accepted, no effect       */
  double tolerance;       /* tolerance_          1e-6 relative residual       */
  int cycleType;          /* cycleType_          0  V; 1 W (coarse correction:
                             two W-cycles of the next level); 2 F (an F-cycle
                             then a V-cycle of the next level; the PCG then
                             takes the flexible beta); 3 K: a V-cycle (the
                             reference runs K as its outer CG)                */
  int solverType;         /* solverType_         0: one V-cycle, 1: PCG       */
  double smootherWeight;  /* smootherWeight_     1.0                          */
  double proOmega;        /* proOmega_           0.67 prolongator smoothing   */
  int device;             /* device_             0                            */
  int blockSize;          /* blockSize_          256 accepted, no effect      */
  /* --- additions --- */
  int precision;          /* 0: FP64 hierarchy, 1: FP32 hierarchy + FP64 PCG,
                             2: as 1, and for one right-hand side the level-0
                             smoother runs on D^-1/2 A D^-1/2 with FP16
                             off-diagonals (row sums kept exact through the
                             FP32 diagonal)                                   */
  int smootherType;       /* 0: Jacobi, weight smootherWeight*4/(3*rho(D^-1A))
                             1: Chebyshev of degree pre/postInnerIters        */
  int aggregationRule;    /* 0: first aggregated neighbour in index order (the
                             reference kernel's behaviour), 1: most-adjacent
                             aggregate, ties to the lowest aggregate id       */
  int coarseSolve;        /* 0: dense FP64 inverse of the coarsest level
                             (Cholesky, computed on the host at setup),
                             applied on the device in FP64 in every precision */
  unsigned long long seed;/* seed of the randomized MIS (reference: time())  */
  double strengthThreshold;/* 0: aggregate the full pattern at level 0 and the
                             induced aggregate graph below (the reference);
                             > 0: aggregate the strength graph of every level,
                             |a_ij| >= theta sqrt(a_ii a_jj)                   */
  int launchMode;         /* 1: each PCG iteration (V-cycle, dots, updates) is
                             captured once as a HIP graph and replayed;
                             0: kernels launched one by one (same results)    */
} fem_params;

/* Fill *p with the reference defaults (src/FEMSolver.cu:11-33). */
void fem_params_default(fem_params* p);

/* ---- construction ----------------------------------------------------------*/
/* FEMSolver::FEMSolver(fname, isTetMesh, verbose), src/FEMSolver.cu:9-44:
 * reads <base>.node/<base>.ele (is_tet != 0, TetMesh::read,
 * src/core/cuda/tetmesh.cu:251-376) or an ASCII PLY/OFF triangle mesh
 * (TriMesh::read, src/core/aggmis/cuda/TriMesh_io.cu:146-257), builds the P1
 * pattern and assembles stiffness + mass on the GPU
 * (FEMSolver::getMatrixFromMesh, src/FEMSolver.cu:141-171).  The element
 * coefficient is the reference's material-label table
 * (src/core/cuda/perform_element_loop_3D.cuh:585-610), mass coefficient 1. */
int fem_create_from_files(const char* base, int is_tet, int verbose, fem_handle** out);

/* Same pipeline for a mesh given as arrays (the generated workloads).
 * dim = 3: conn is ne x 4 tets; dim = 2: conn is ne x 3 triangles.
 * xyz is nv x 3 (z ignored in 2D).  kappa: per-element conductivity or NULL
 * (all ones).  Operator K = sum_e kappa_e S_e + mass_coeff * M.
 * flags bit 0: 2D mass with the signed Jacobian
 * (src/core/cuda/perform_element_loop_2D.cuh:92).  The host arrays are
 * copied by four library threads (one HIP stream each, kept per device). */
int fem_create_from_arrays(int dim, int64_t nv, const double* xyz, int64_t ne,
                           const int32_t* conn, const double* kappa,
                           double mass_coeff, int flags, int verbose,
                           fem_handle** out);

/* FEMSolver::~FEMSolver, src/FEMSolver.cu:46-51. */
void fem_destroy(fem_handle* h);

/* No counterpart in the reference.  Frees the library's reserve of device
 * blocks: temporaries that mesh builds and AMG setups release are kept (at
 * most 8 GiB per device) for later builds and setups to reuse, instead of a
 * hipFree each.  The reserve is also freed whenever a library allocation
 * fails.  For callers that need the memory for their own allocations; must
 * not overlap with a fem_create_* or fem_setup call. */
void fem_release_cached_memory(void);

/* Message of the last failure on h (or of the last failed create if h is NULL). */
const char* fem_last_error(const fem_handle* h);

/* Public solver fields (src/FEMSolver.h:40-61).  Changing a field that shapes
 * the hierarchy makes the next solve rebuild it. */
int fem_set_params(fem_handle* h, const fem_params* p);
int fem_get_params(const fem_handle* h, fem_params* p);

/* Queue all device work of h on `stream` (a hipStream_t; NULL = own stream). */
int fem_set_stream(fem_handle* h, void* stream);

/* ---- matrix access ---------------------------------------------------------*/
/* FEMSolver::getMatrixRows, src/FEMSolver.cu:95-97. */
int64_t fem_get_matrix_rows(const fem_handle* h);
int64_t fem_get_matrix_nnz(const fem_handle* h);
int64_t fem_get_num_elements(const fem_handle* h);

/* The assembled A_h_ (src/FEMSolver.cu:170) as CSR in user vertex numbering,
 * columns ascending.  Any output pointer may be NULL. */
int fem_get_matrix_csr(const fem_handle* h, int64_t* row_ptr, int32_t* cols, double* vals);

/* FEMSolver::readMatlabSparseMatrix, src/FEMSolver.cu:177-356: replace A by
 * the file matrix plus 1e-12 on every mesh-pattern slot.  Returns 0, or the
 * reference's 1 / -1 codes on a malformed file (message in fem_last_error). */
int fem_read_matlab_sparse(fem_handle* h, const char* filename);

/* Replace the per-element conductivity (user element order) and reassemble
 * in place on the device; the hierarchy is rebuilt by the next setup/solve. */
int fem_update_kappa(fem_handle* h, const double* kappa);

/* ---- setup and solve -------------------------------------------------------*/
/* AMG::setup, src/core/cuda/amg.cu:79-144: aggregation, smoothed prolongator,
 * Galerkin coarse operators, smoother data, coarsest-level factorisation.
 * Idempotent; fem_solve calls it when needed. */
int fem_setup(fem_handle* h);

typedef struct fem_solve_info {
  int iterations;          /* PCG iterations (max over right-hand sides)    */
  int converged;           /* 1 if every column reached tolerance           */
  double rel_residual;     /* true ||b - A x|| / ||b|| (max over columns)   */
  double setup_ms;         /* hierarchy build time of this call (0 if reused)*/
  double solve_ms;         /* device time of the iteration                  */
  int num_levels;
  double grid_complexity;
  double operator_complexity;
} fem_solve_info;

/* FEMSolver::solveFEM(x_h, b_h), src/FEMSolver.cu:58-93, with AMG::solve,
 * src/core/cuda/amg.cu:173-200 and CG_Flex_Cycle, src/core/cuda/cgcycle.cu:7-69.
 * b, x: nv x nrhs, row-major (vertex-major), user numbering; x holds the
 * initial guess on entry and the solution on return.  on_device != 0: b and x
 * are device pointers.  solverType 0 applies one V-cycle; 1 runs PCG to
 * `tolerance` on the true FP64 residual.  Returns FEM_ERR_NOT_CONVERGED (with
 * x filled) if maxIters ends the iteration first. */
int fem_solve(fem_handle* h, const double* b, double* x, int nrhs, int on_device,
              fem_solve_info* info);

/* ---- adjoint (no counterpart in the reference) -----------------------------*/
/* dLdkappa[e] = -|T_e| grad(lambda)_e . grad(u)_e with K lambda = dLdu, solved
 * on the hierarchy of the last setup (K symmetric).  lambda_out may be NULL.
 * u, dLdu, lambda_out: nv, user numbering; dLdkappa: ne, user element order. */
int fem_adjoint_grad(fem_handle* h, const double* u, const double* dLdu,
                     double* dLdkappa, double* lambda_out, int on_device,
                     fem_solve_info* info);

/* Element gradient only: dLdkappa[e] = -|T_e| grad(a)_e . grad(u)_e. */
int fem_element_grad(fem_handle* h, const double* a, const double* u,
                     double* dLdkappa, int on_device);

/* y = A x with the assembled FP64 operator (user numbering, nv x nrhs). */
int fem_apply_operator(fem_handle* h, const double* x, double* y, int nrhs, int on_device);

/* z = M r: one application of the V-cycle preconditioner (user numbering). */
int fem_apply_preconditioner(fem_handle* h, const double* r, double* z, int on_device);

/* z = M r like fem_apply_preconditioner(h, r, z, 1) (bitwise the same z),
 * allocating no device memory: no counterpart in the reference; for callers
 * that cannot allocate (code under HIP graph capture, serving loops, a
 * caller's own Krylov iteration).  r, z: device pointers, nv doubles each,
 * user numbering.  The work (r and z in internal numbering, r.z partials)
 * lives in one static device array per GPU (about 32 MiB, held until the
 * process ends), so nv must not exceed fem_noalloc_max_rows() (2^21);
 * larger operators return FEM_ERR_CAPACITY.  Needs fem_setup (hierarchy and
 * work vectors) beforehand, the handle's device current and profiling off
 * (else FEM_ERR_STATE).  Enqueues on the handle's stream and returns without
 * synchronizing.  Every call on one GPU shares the array: calls that overlap
 * on one GPU (other streams, other handles, other threads) race. */
int fem_apply_preconditioner_noalloc(fem_handle* h, const double* r, double* z);
int64_t fem_noalloc_max_rows(void);

/* ---- hierarchy inspection (tests and diagnostics) ---------------------------*/
int fem_get_num_levels(const fem_handle* h);
/* Level `lvl` operator in the library's internal numbering: rows, nnz. */
int fem_get_level_info(const fem_handle* h, int lvl, int64_t* n, int64_t* nnz);
/* Copy level data (internal numbering) to host arrays; NULL skips an item.
 * aggregates: n entries (coarse id of each row; absent on the coarsest). */
int fem_get_level_matrix(const fem_handle* h, int lvl, int32_t* row_ptr, int32_t* cols,
                         double* vals, int32_t* aggregates);
/* Prolongator of level lvl -> lvl+1 (n_lvl x n_lvl+1 CSR); NULL skips. */
int fem_get_level_prolongator(const fem_handle* h, int lvl, int64_t* nnz,
                              int32_t* row_ptr, int32_t* cols, double* vals);
/* y = M x for level lvl's operator (which = 0, n_lvl x n_lvl), prolongator
 * (1, n_lvl x n_lvl+1) or restriction (2, n_lvl+1 x n_lvl) through the
 * one-column kernels the V-cycle launches, in the precision it applies them
 * (FP32 values and vectors in precisions 1 and 2, FP64 in 0); which = 3 on the
 * coarsest level: y = A_c^-1 x, the direct solve the V-cycle applies there
 * (FP64 inverse and FP64 sums in every precision, x and y rounded to the
 * level's vector type; replaces the host LU of src/core/cuda/amg.cu:101-107,
 * amg_level.cu:25-31).  Host arrays in internal numbering.  A test and
 * diagnostic hook beside the level exports above (the reference applies these
 * matrices inside AMG_Level::cycle, src/core/cuda/amg_level.cu:22-128). */
int fem_apply_level(fem_handle* h, int lvl, int which, const double* x, double* y);
/* Smoother data of level lvl (not the coarsest): rho, the power-iteration
 * estimate of rho(D^-1 A); the Chebyshev interval [lmin, lmax] =
 * [1.1 rho / 30, 1.1 rho] and the Jacobi weight omega = smootherWeight *
 * 4 / (3 rho) derived from it.  Any output may be NULL.  Diagnostic hook; the
 * reference's smoother has a fixed weight (src/core/cuda/gauss_seidel.cu:1279-1376). */
int fem_get_level_smoother(const fem_handle* h, int lvl, double* rho, double* lmin, double* lmax,
                           double* omega);
/* One smoothing step on level lvl through the kernel the V-cycle launches, in
 * the hierarchy's precision (precision 2: the FP32 kernels): mode 0 (Jacobi)
 * y = x + c1 D^-1 (b - A x); mode 1 (Chebyshev) d = c1 d + c2 D^-1 (b - A x)
 * and y = x + d.  Host arrays of n_lvl entries in internal numbering; d is
 * read and overwritten.  Diagnostic hook for the smoother that replaces the
 * reference's partition-local Jacobi sweeps (src/core/cuda/gauss_seidel.cu:
 * 1279-1376, 3627-3736). */
int fem_apply_smoother_step(fem_handle* h, int lvl, int mode, double c1, double c2, const double* x,
                            const double* b, double* d, double* y);
/* internal -> user vertex map (nv entries). */
int fem_get_permutation(const fem_handle* h, int32_t* perm);

/* Run the device randomized distance-k MIS + aggregation on a host CSR graph
 * (no diagonal) with the given parameters; writes the aggregate id of every
 * node and returns the aggregate count in *n_agg.  Exposes the kernels of
 * RandMIS_Aggregator / misHelpers::aggregateGraph
 * (src/core/cuda/misHelpers.cu:513-601, randomizedMIS_GPU.cu:154-272). */
int fem_device_aggregate(int64_t n, const int32_t* adj_ptr, const int32_t* adj,
                         int depth, int min_size, int rule, unsigned long long seed,
                         int32_t* aggregate, int32_t* n_agg);

/* Condition an aggregation (ids 0..n_agg-1, all in use) of a host CSR graph
 * (no diagonal) to sizes [min_size, max_size] with the merge-split
 * conditioner of aggregatorType 3 (MergeSplitConditionerGPU::Condition,
 * src/core/aggmis/cuda/AggMIS_MergeSplitConditioner.cu:1625-1715, as NewMIS
 * calls it, src/core/cuda/ComputePermutationMethods.cu:403-407).  Host only;
 * rewrites aggregate and returns the new count in *n_agg_out. */
int fem_condition_aggregates(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int32_t* aggregate,
                             int64_t n_agg, int min_size, int max_size, int64_t* n_agg_out);

/* aggregatorType 4 / 5, host diagnostic: the flood-fill distance-`depth` MIS
 * (AggMIS::MIS::FloodFillMIS, src/core/aggmis/cuda/AggMIS_MIS_CPU.cu:25-140)
 * of a symmetric graph (CSR without the diagonal) from the start node the
 * setup derives from `seed` (returned in *starter_out when not NULL; the
 * reference starts at rand() % n after srand(time(NULL))), and every node
 * aggregated to the nearest root (AggMIS::Aggregation::AggregateToNearest,
 * AggMIS_Aggregation_CPU.cu:4-92).  roots_out[i] = 1 for a root, 0 else;
 * aggregate ids are dense, numbered in root index order.  Type 4 then runs
 * fem_condition_aggregates; type 5 uses the aggregation as it is. */
int fem_flood_fill_aggregate(int64_t n, const int32_t* adj_ptr, const int32_t* adj, int depth,
                             unsigned long long seed, int64_t* starter_out, int32_t* roots_out,
                             int32_t* aggregate_out, int64_t* n_agg_out);

/* ---- profiling -------------------------------------------------------------*/
/* Time each launch of the named kernel family with HIP events on the handle's
 * stream while enabled.  Families: "assembly", "element_grad", "pcg_spmv"
 * (FP64 outer operator + p.q), "smooth_l0" (level-0 smoothing step),
 * "smooth_l0_out" (fused last level-0 step), "resid_l0" (level-0 residual),
 * "smooth_l0_h" / "smooth_l0_h_out" (the FP16 level-0 smoother of precision 2).
 * Timing a launch makes the stream wait between launches (no graph replay). */
int fem_profile_enable(fem_handle* h, int enable);
int fem_profile_get(const fem_handle* h, const char* family, int64_t* launches,
                    double* total_ms, double* bytes_per_launch);

/* ---- host IO (no GPU needed) -----------------------------------------------*/
/* TetMesh::read (src/core/cuda/tetmesh.cu:251-376): float-parsed coordinates,
 * 0/1-based detection, optional label column.  Arrays are malloc'd; free
 * with fem_free.  labels is filled with zeros if the file has none. */
int fem_read_tetgen(const char* base, int64_t* nv, double** xyz, int64_t* ne,
                    int32_t** tets, int32_t** labels);
/* TriMesh::read subset: ASCII PLY and OFF vertices + triangle faces
 * (src/core/aggmis/cuda/TriMesh_io.cu:259-502,692-708). */
int fem_read_trimesh(const char* filename, int64_t* nv, double** xyz, int64_t* ne,
                     int32_t** tris);
void fem_free(void* p);

/* FEMSolver::readMatlabArray, src/FEMSolver.cu:358-449.  With out == NULL
 * only *n is set.  Returns 0 or -1 (reference codes). */
int fem_read_matlab_array(const char* filename, double* out, int64_t* n);
/* FEMSolver::writeMatlabArray, src/FEMSolver.cu:451-517 (byte-identical). */
int fem_write_matlab_array(const char* filename, const double* x, int64_t n);
/* MAT v5 sparse reader half of readMatlabSparseMatrix (src/FEMSolver.cu:177-311):
 * COO entries rounded to float as the reference does when round_to_float != 0.
 * With rows == NULL only *nnz, *nrows, *ncols are set. */
int fem_read_matlab_sparse_coo(const char* filename, int round_to_float, int64_t* nrows,
                               int64_t* ncols, int64_t* nnz, int32_t* rows, int32_t* cols,
                               double* vals);
/* FEMSolver::writeVTK, src/FEMSolver.cu:519-582 (tets when dim==3). */
int fem_write_vtk(const char* fname, int dim, int64_t nv, const double* xyz, int64_t ne,
                  const int32_t* conn, const double* values, int64_t nvalues);

/* Library version string. */
const char* fem_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FEM_MI355X_H */
