/* fhgpu -- MI355X (gfx950) solve path of the statically condensed HHO face
 * system.  C ABI: plain pointers and sizes, int status (0 = OK, < 0 = error;
 * the message of the last error of the calling thread is fh_last_error()).
 *
 * Device vectors are raw HBM pointers owned by the caller (any allocator:
 * hipMalloc, a torch tensor, ...).  Matrices are created from HOST arrays
 * (copied; the library owns their device memory).  All work of one context
 * runs on its own HIP stream; functions that return a host scalar or solver
 * statistics synchronise that stream.  Handles are not thread-safe.
 *
 * Citations are files of the reference code (pmatalon/fhhos4, src/...).
 */
#ifndef FHGPU_H
#define FHGPU_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct fh_ctx fh_ctx;
typedef struct fh_mat fh_mat;
typedef struct fh_bjac fh_bjac;
typedef struct fh_coloring fh_coloring;
typedef struct fh_solver fh_solver;

/* storage / vector precision codes */
enum { FH_F64 = 0, FH_F32 = 1, FH_BF16 = 2 };

/* ------------------------------------------------------------ context */
/* Replaces the reference's implicit host execution context
 * (src/Utils/ParallelLoop.h:34-57, BaseParallelLoop thread pool). */
int fh_ctx_create(int device, fh_ctx** out);
int fh_ctx_destroy(fh_ctx* ctx);
int fh_ctx_sync(fh_ctx* ctx);
/* Error reporting; replaces Utils::FatalError (src/Utils/Utils.h:258-263). */
const char* fh_last_error(void);

/* ------------------------------------------------------------ matrices */
/* Condensed operator / transfer matrix from the reference's storage: scalar
 * CSR (RowMajor, int32, src/Utils/Types.h:11-22), converted to dense bs x bs
 * blocks.  Replaces NonZeroCoefficients::Fill into SparseMatrix
 * (src/Utils/NonZeroCoefficients.h:94-96).  keep_host keeps FP64 values on
 * the host for set-up algorithms (aggregation, Galerkin, transpose). */
int fh_mat_from_csr(fh_ctx* ctx, int nrows, int ncols, int bs, const int* rowptr, const int* colidx,
                    const double* vals, int keep_host, fh_mat** out);
/* Same, from block arrays (nbr+1 row pointers, nnzb block columns, nnzb*bs*bs
 * row-major block values). */
int fh_mat_from_blocks(fh_ctx* ctx, int nbr, int nbc, int bs, const int* rowptr, const int* colidx,
                       const double* vals, int keep_host, fh_mat** out);
/* Copy with values stored in FH_F64 / FH_F32 / FH_BF16. */
int fh_mat_convert(fh_ctx* ctx, const fh_mat* A, int prec, fh_mat** out);
int fh_mat_destroy(fh_mat* A);
int fh_mat_info(const fh_mat* A, int* nbr, int* nbc, int* bs, int64_t* nnzb, int* prec);
/* Host copy of the block arrays of a matrix (rowptr: nbr+1, colidx: nnzb,
 * vals: nnzb*bs*bs FP64); NULL pointers are skipped.  Values come from the
 * host copy (keep_host) or, for a device-only FP64 matrix, from HBM after
 * the device's pending work. */
int fh_mat_blocks(const fh_mat* A, int* rowptr, int* colidx, double* vals);
/* Device pointers of the block structure and values (for profiling tools). */
int fh_mat_device_arrays(const fh_mat* A, const int** rowptr, const int** colidx, const void** vals);
/* Execution layout of full-matrix products.  on = 1 builds a second copy of
 * the values: a sliced copy when every row fits one wavefront (blocks per
 * row * n_f / 2 <= 64 for even n_f, blocks per row * n_f <= 64 for odd n_f;
 * a wavefront owns consecutive block rows, lane = (row, block slot, column
 * pair)), otherwise a chunked copy (a wavefront per row, rows cut into
 * chunks of 64 / (n_f / 2) or 64 / n_f block slots); *active (may be NULL)
 * tells whether A has one.  on = 0 drops it and keeps A on the block-CSR
 * kernels.  Without a call the library builds one at the first full-matrix
 * product when the matrix has at least 4096 block rows and padding adds at
 * most 25% to the values (the chunked copy: FP64, FP32, or BF16 with even
 * n_f; uneven rows of 32768 or more block rows: a sliced copy in two parts,
 * short and long rows).  No
 * reference counterpart: the reference has one CSR storage
 * (src/Utils/Types.h:11-22); the copy is a GPU execution detail of the same
 * operator. */
int fh_mat_set_sliced(fh_ctx* ctx, const fh_mat* A, int on, int* active);
/* Uses (on != 0) or stops using the element layout of an operator assembled
 * on the GPU from hexahedral cells (fh_hho_assemble_gpu keeps the condensed
 * cell matrices, upper triangles, beside the block rows and uses them for
 * full products, residuals and Chebyshev steps: about 55% of the block
 * rows' bytes per product).  *active: whether A now uses one (0 when A has
 * none, or when the environment sets FHGPU_ELEM=0).  No reference
 * counterpart: the reference assembles A from the same cell matrices
 * (Diffusion_HHO.h:633-637); this keeps them as a GPU execution detail. */
int fh_mat_set_elem(fh_ctx* ctx, const fh_mat* A, int on, int* active);
/* Execution layout of A's full-matrix products: 0 block CSR (or not decided
 * yet), 1 sliced copy, 2 sliced copy in two parts (short and long rows, each
 * with its own width: uneven rows such as a prolongation's), 3 chunked copy,
 * 4 SELL-64 copy (scalar square operators: 64 rows per slice, slot-major),
 * 5 element layout (fh_mat_set_elem).  No reference counterpart (an
 * execution detail). */
int fh_mat_layout(const fh_mat* A, int* layout);

/* -------------------------------------------- exported linear system */
/* Writes the scalar operator as Matrix Market text, byte-compatible with the
 * reference's export of A.dat / level<l>_A.dat: ExportModule::ExportMatrix
 * (src/Utils/ExportModule.h:68-73) -> Eigen::saveMarket (header
 * "%%MatrixMarket matrix coordinate  real general", "rows cols nnz", 1-based
 * "i j v" in row-major order, "%.64e" values).  Stored entries: the scalar
 * pattern the matrix was created from (keep_host) or every non-zero entry. */
int fh_mat_export_market(fh_ctx* ctx, const fh_mat* A, const char* path);
/* Reads such a file (Eigen::loadMarket semantics: duplicates summed) into a
 * block matrix with bs x bs blocks (rows and columns multiples of bs). */
int fh_mat_import_market(fh_ctx* ctx, const char* path, int bs, int keep_host, fh_mat** out);
/* Host vector <-> "%%MatrixMarket matrix array real general" text, as
 * ExportModule::ExportVector (ExportModule.h:75-80, Eigen::saveMarketVector)
 * writes b.dat and solutionFaces.dat.  Import: *n receives the length; v
 * (capacity cap) is filled when not NULL. */
int fh_vec_export_market(int64_t n, const double* v, const char* path);
int fh_vec_import_market(const char* path, int64_t cap, double* v, int64_t* n);

/* ------------------------------------------------------------ kernels */
/* y = alpha A x + beta y.  Replaces `A * x` of Eigen's SparseMatrix in the
 * solvers (src/Solver/Krylov/ConjugateGradient.h:70,
 * Multigrid/Level.h:139-159).  vprec: FH_F64 (matrix FP64) or FH_F32 (matrix
 * FP32 or BF16, FP32 accumulation). */
int fh_spmv(fh_ctx* ctx, const fh_mat* A, double alpha, const void* x, double beta, void* y, int vprec);
/* r = b - A x  (src/Solver/Multigrid/Multigrid.h:478). */
int fh_residual(fh_ctx* ctx, const fh_mat* A, const void* b, const void* x, void* r, int vprec);
/* Deterministic dot product, FP64 accumulation (Eigen dot in
 * src/Solver/Krylov/ConjugateGradient.h:73,86). */
int fh_dot(fh_ctx* ctx, int64_t n, const void* x, const void* y, int vprec, double* out);
/* y = a x + b y  (vector updates, ConjugateGradient.h:67,77,82). */
int fh_axpby(fh_ctx* ctx, int64_t n, double a, const void* x, double b, void* y, int vprec);
/* Precision conversion of a vector. */
int fh_convert(fh_ctx* ctx, int64_t n, const void* src, int sprec, void* dst, int dprec);

/* In-situ timing of the block-row kernel: every launch of operation `mode`
 * (0 spmv, 1 residual, 2 smoothing, 3 fused Chebyshev step, 4 residual of
 * the first entry of every block row, fused with the p-restriction to one
 * mode per face) on matrix A is
 * bracketed by HIP events on the context's stream until fh_profile_end,
 * which returns the number of launches and their mean duration in ms.  Replaces the reference's Timer
 * accounting of matrix-vector products (src/Utils/Timer.h:60-123). */
int fh_profile_begin(fh_ctx* ctx, const fh_mat* A, int mode);
int fh_profile_end(fh_ctx* ctx, int* launches, double* mean_ms);

/* ------------------------------------------------------------ smoothers */
/* Block Jacobi set-up: inverse of every diagonal block, stored in dprec
 * (FH_F64 / FH_F32).  Replaces BlockJacobi::Setup
 * (src/Solver/FixedPoint/BlockJacobi.h:57-74). */
int fh_bjac_create(fh_ctx* ctx, const fh_mat* A, double omega, int dprec, fh_bjac** out);
int fh_bjac_destroy(fh_bjac* bj);
/* Download of the block inverses (nbr*bs*bs doubles). */
int fh_bjac_inverses(fh_ctx* ctx, const fh_bjac* bj, double* host_out);
/* nsweeps damped block-Jacobi sweeps x <- x + w D^-1 (b - A x); work is a
 * device vector of the same length (x_is_zero: x is treated as 0).
 * Replaces BlockJacobi::ExecuteOneIteration (BlockJacobi.h:77-127). */
int fh_bjac_sweep(fh_ctx* ctx, const fh_bjac* bj, const fh_mat* A, const void* b, void* x, void* work, int nsweeps,
                  int x_is_zero, int vprec);
/* No counterpart in the reference.  fh_bjac_sweep for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the ping-pong vector of the sweeps is a
 * file-scope device array of 2^21 doubles (16 MiB) per GPU, held until the
 * process ends, instead of the caller's `work`.  Vectors longer than
 * 2^21 FP64 or 2^22 FP32 entries are rejected (-5).  The call allocates
 * nothing: it never builds A's sliced / chunked copy (a matrix whose layout
 * is not decided yet runs on the block-CSR kernels) and times no launches.
 * Calls that overlap on one GPU share the array and race. */
int fh_bjac_sweep_noalloc(fh_ctx* ctx, const fh_bjac* bj, const fh_mat* A, const void* b, void* x, int nsweeps,
                          int x_is_zero, int vprec);
/* Greedy colouring of the block graph (block-row order). */
int fh_coloring_create(fh_ctx* ctx, const fh_mat* A, fh_coloring** out);
int fh_coloring_destroy(fh_coloring* col);
int fh_coloring_get(const fh_coloring* col, int* ncolours, int* colour_of_row);
/* One multicolour block Gauss-Seidel sweep, dir 0 forward / 1 backward /
 * 2 symmetric; uses the block inverses of bj.  Replaces the lexicographic
 * BlockGaussSeidel::ExecuteOneIteration (BlockGaussSeidel.h:75-191) with a
 * colour ordering (same fixed point, different sweep order). */
int fh_mcgs_sweep(fh_ctx* ctx, const fh_mat* A, const fh_bjac* bj, const fh_coloring* col, int dir, const void* b,
                  void* x, int vprec);

/* ------------------------------------------------------------ AMG set-up */
/* One pairwise aggregation pass on block rows (host algorithm, bit-exact with
 * AlgebraicMesh::Build + PairwiseAggregation::Perform,
 * src/Solver/Multigrid/AggregAMG/AlgebraicMesh.h:54-153,
 * PairwiseAggregation.h:9-205).  agg_out: nbr ints. */
int fh_aggregate_pairwise(fh_ctx* ctx, const fh_mat* A, double theta, int* agg_out, int* nagg, int* coarsest);
/* Galerkin operator of a piecewise-constant block prolongation
 * (AggregLevel::GalerkinOperator, AggregAMG.h:156-180, extended to blocks). */
int fh_galerkin_aggregate(fh_ctx* ctx, const fh_mat* A, const int* agg, int nagg, fh_mat** out);

/* Galerkin operator P^T A P of a block prolongation (host; matrices with
 * keep_host).  Level::ComputeGalerkinOperator (src/Solver/Multigrid/Level.h:133-137). */
int fh_galerkin_rap(fh_ctx* ctx, const fh_mat* A, const fh_mat* P, fh_mat** out);

/* ------------------------------------------------------------ solvers */
typedef struct fh_solver_args {
    const char* solver;        /* "cg" | "fcg" | "mg" | "aggregamg" | "bj" | "bgs"  (SolverFactory.h:126-233) */
    const char* prec;          /* "" | "bj" | "bj23" | "bgs" | "mg" | "aggregamg" */
    char cycle;                /* 'V' | 'W' | 'K' (Multigrid.h:18-19) */
    int wloops;
    int nu1, nu2;              /* pre/post smoothing iterations */
    const char* pre_smoother;  /* SmootherFactory codes: bj bj23 bgs rbgs sbgs gs rgs sgs j j23;
                                  not in the reference: cheb1..cheb4, a Chebyshev polynomial
                                  of degree 1-4 of D^-1 A on [0.25, 1.1] x lambda_max */
    const char* post_smoother;
    double relax;              /* RelaxationParameter */
    double tol;                /* relative residual ||r||/||b|| */
    int max_iter;              /* MaxIterations; 0 stops before the first iteration
                                  (IterativeSolver.h:229-230); -1: the default 200 */
    int max_coarse;            /* MatrixMaxSizeForCoarsestLevel (ProgramArguments.h:61) */
    double theta;              /* strong coupling threshold (SolverFactory.h:95) */
    int mg_prec;               /* working precision of the multigrid vectors: FH_F64 | FH_F32 */
    int mg_store;              /* storage of the level matrices: FH_F64 | FH_F32 | FH_BF16 */
    int galerkin;              /* geometric multigrid: coarse operators R A P (-g 1, Multigrid.h:21) */
    /* Not in the reference (one smoother for all levels there): smoother code
     * and relaxation of the levels below the finest.  NULL or "" keeps the
     * finest level's smoothers; coarse_relax <= 0 keeps relax. */
    const char* coarse_smoother;
    double coarse_relax;
    int fcg_truncation;        /* FlexibleConjugateGradient::Truncation (<= 0: 1), <= 32 */
    int fcg_restart;           /* FlexibleConjugateGradient::Restart (FlexibleConjugateGradient.h:18-25) */
    /* Not in the reference: storage of the level matrices below the finest
     * (0 = as mg_store, FH_F32, FH_BF16; needs FP32 multigrid vectors). */
    int coarse_store;
    /* Not in the reference: K-cycle (cycle 'K') only on levels 1..kcycle_levels
     * (<= 0: all, as Multigrid.h:611-656), V-cycle recursion below. */
    int kcycle_levels;
    /* Multigrid::CoarseLevelChangeSmoothingOperator / Coeff (Multigrid.h:224-242,
     * -cycle V,nu1,nu2,<op><coeff>): each level's smoothing iterations are its
     * finer level's changed by op ('+', '-', '*'; 0 = unchanged) with coeff;
     * a count that falls below 0 is 0. */
    char nu_change_op;
    int nu_change_coeff;
    /* Not in the reference: storage of the operator in FCG's own product A p
     * (FH_F64, 0: the operator itself; FH_F32: the FP32 copy of the
     * multigrid's finest level, FP64 vectors and accumulation, with the
     * recursive residual replaced by the FP64 b - A x every replace_every
     * iterations (<= 0: 10) and before convergence is declared, so the
     * stopping test always holds for the FP64 residual). */
    int outer_prec;
    int replace_every;
    /* Not in the reference: when coarse_nu_given is nonzero, the levels below
     * the finest take coarse_nu1 / coarse_nu2 (>= 0) pre/post smoothing
     * iterations instead of nu1 / nu2 changed by nu_change_op; the finest
     * level keeps nu1 / nu2 (e.g. a pre-smoothing-only finest level). */
    int coarse_nu_given;
    int coarse_nu1, coarse_nu2;
} fh_solver_args;

typedef struct fh_solve_stats {
    int iterations;
    int converged;
    int levels;
    int graph_mode;            /* multigrid cycle execution: 0 eager (undecided), 1 eager
                                  (faster than replay), 2 HIP graph replay */
    double final_res;
    double setup_s;
    double solve_s;
    double op_complexity;
    int nn_used;               /* learned post-smoother: corrections accepted */
    int nn_fallback;           /* learned post-smoother: block-Jacobi fallbacks */
} fh_solve_stats;

/* Creation from string codes (SolverFactory<Dim>::CreateSolver,
 * src/Solver/SolverFactory.h:126-310). */
int fh_solver_create(fh_ctx* ctx, const fh_solver_args* args, fh_solver** out);
int fh_solver_destroy(fh_solver* s);
/* Geometric multigrid hierarchy input: operator of level l (l >= 1) and the
 * prolongation from level l+1 to level l (Level::OperatorMatrix / P,
 * src/Solver/Multigrid/Level.h:11,32; LevelForHHO.h:182-213).  Matrices must
 * outlive the solver (Solver.h:9,25-28). */
int fh_solver_set_level(fh_solver* s, int level, const fh_mat* A_level, const fh_mat* P_level);
/* p-multigrid level: the transfer from level+1 to level is the natural
 * injection of hierarchical face bases keeping the first nfc modes of every
 * face (P_LevelForHHO::SetupProlongation, src/Solver/Multigrid/
 * MultigridForHHO/P_MultigridForHHO.h:62-86; R = P^T, :89-92); the operator
 * of level+1 is its Galerkin operator R A P, the leading nfc x nfc sub-block
 * of every block (Level::ComputeGalerkinOperator, Level.h:133-137), unless
 * fh_solver_set_level gives one.  Levels below may be geometric (h) levels. */
int fh_solver_set_level_p(fh_solver* s, int level, int nfc);
/* Solver::Setup (Solver.h:25-28; Multigrid::Setup, Multigrid.h:88-313). */
int fh_solver_setup(fh_solver* s, const fh_mat* A);
/* IterativeSolver::Solve(b, "0") (IterativeSolver.h:55-73): x (device, FP64)
 * receives the solution from a zero initial guess. */
int fh_solver_solve(fh_solver* s, const double* b, double* x, fh_solve_stats* stats);
/* IterativeSolver::Solve(b, x, xEquals0) (IterativeSolver.h:75-80): x
 * (device, FP64) holds the initial guess on input, unless x_is_zero, and the
 * solution on output.  With a guess the first residual is b - A x and the
 * first history entry its norm over ||b|| (ConjugateGradient.h:36-45,
 * IterationResult.h:121-139); the initial-guess codes "1" and "rand" of
 * IterativeSolver::Solve(b, code) (:55-73) are this call with the caller's
 * vector of ones or random values. */
int fh_solver_solve_x0(fh_solver* s, const double* b, double* x, int x_is_zero, fh_solve_stats* stats);
/* ------------------------------------------------- learned post-smoother */
/* No counterpart in the reference.  A message-passing network over the face
 * graph (hidden width 32, `layers` layers; a few thousand parameters)
 * predicts a correction of the preconditioned vector from the residual left
 * by the multigrid cycle and two diffusion features per face.  It runs inside
 * FCG (FlexibleConjugateGradient.h:47-171 accepts a varying preconditioner)
 * after the multigrid cycle of SolverPreconditioner (Preconditioner.h:100-199);
 * block Jacobi (BlockJacobi.h:77-127) replaces it whenever it would not lower
 * the energy norm of the error. */
typedef struct fh_nn fh_nn;
typedef struct fh_nngraph fh_nngraph;
int fh_nn_create(fh_ctx* ctx, int nf, int layers, uint64_t seed, fh_nn** out);
int fh_nn_destroy(fh_nn* nn);
int fh_nn_num_params(const fh_nn* nn, int64_t* n);
int fh_nn_set_params(fh_nn* nn, const float* params);
int fh_nn_get_params(const fh_nn* nn, float* params);
/* Face graph of an FP64 operator: edge weights |tr A_ij| / sum_k |tr A_ik|,
 * block inverses, and 2 features per block row (host, nbr x 2: mean and
 * half-difference of log kappa of the face's cells).  A must outlive it. */
int fh_nn_graph_create(fh_ctx* ctx, const fh_mat* A, const double* face_features, fh_nngraph** out);
int fh_nn_graph_destroy(fh_nngraph* g);
/* delta = network correction of r1 (device FP64 vectors); bf16 != 0 uses the
 * BF16 weights (inference), else the FP32 training weights. */
int fh_nn_forward(fh_ctx* ctx, fh_nn* nn, const fh_nngraph* g, const double* r1, double* delta, int bf16);
/* Relative energy error (e1-delta)^T A (e1-delta) / e1^T r1 of one training
 * pair (r1 = A e1, device FP64) and its gradient (host FP32, may be NULL). */
int fh_nn_loss_grad(fh_ctx* ctx, fh_nn* nn, const fh_nngraph* g, const double* r1, const double* e1, double* loss,
                    float* grad);
/* One Adam step (lr) on one pair; *loss = the loss before the step. */
int fh_nn_train_step(fh_ctx* ctx, fh_nn* nn, const fh_nngraph* g, const double* r1, const double* e1, float lr,
                     double* loss);
/* Enables the learned post-smoother in an FCG solver (before setup; the graph
 * must be built on the solver's operator); omega: fallback relaxation.
 * unit_step 0: z0 + (delta.r1 / delta.A delta) delta, accepted when
 * delta.r1 > 0; 1: z0 + delta, accepted when delta.A delta < 2 delta.r1. */
int fh_solver_set_nn(fh_solver* s, fh_nn* nn, const fh_nngraph* g, double omega, int unit_step);

/* Level::Setup with ExportComponents (src/Solver/Multigrid/Level.h:116-125,
 * ExportModule prefix "level<l>", Level.h:50-54): after setup, writes
 * <dir>/level<l>_A.dat for every level and level<l>_P.dat / level<l>_R.dat
 * for every level but the coarsest (exported-system format). */
int fh_solver_export_levels(fh_solver* s, const char* dir);
/* IterativeSolver::Tolerance and MaxIterations (IterativeSolver.h:11-23). */
int fh_solver_set_stop(fh_solver* s, double tol, int max_iter);
/* z = M r with the solver's preconditioner alone (Preconditioner::Apply,
 * Preconditioner.h:100-118: one multigrid cycle from 0); device FP64. */
int fh_solver_apply_prec(fh_solver* s, const double* r, double* z);
/* Normalised residual at each stopping test (IterationResult.h:121-139);
 * returns the number of entries written (<= cap). */
int fh_solver_history(const fh_solver* s, double* out, int cap);
/* Unknowns of multigrid level l. */
int fh_solver_level_size(const fh_solver* s, int level);

/* ------------------------------------------------------------ HHO system */
/* Producer of the condensed face system on structured affine meshes:
 * mesh + local HHO operators + static condensation + assembly
 * (Program_Diffusion_HHO<Dim>::Execute up to Assemble,
 * src/Program/Program_Diffusion_HHO.h:16-238; Diffusion_HHO::Assemble,
 * src/Discretizations/HHO/Diffusion_HHO.h:252-704). */
typedef struct fh_hho fh_hho;
typedef struct fh_hho_args {
    int dim;                /* 2 or 3 */
    const char* mesh;       /* "stetra" (cube -> 6 tets, Cube_CartesianTetrahedralMesh.h) | "stri" (square ->
                               2 triangles, Square_TriangularMesh.h) | "cart" (Cube_CartesianMesh.h,
                               Square_CartesianMesh.h) | "hex" / "quad": the cells and faces of "cart" as
                               trilinear / bilinear maps of their (perturbable) vertices */
    int n;                  /* background cells per direction */
    int k;                  /* face polynomial degree 0..3 */
    double perturb;         /* interior vertex perturbation (fraction of h) */
    uint64_t seed;
    const char* source;     /* "pisine" | "sine" (DiffusionTestCase.h:62-131) | "modes" */
    const char* kappa;      /* "one" | "lognormal" | "subdomain" */
    double sigma;           /* lognormal standard deviation */
    int subdomain;          /* cells per subdomain edge (on the reference grid) */
    double aniso;           /* z-scaling of the anisotropic subdomains */
    int ref_n;              /* finest grid of a hierarchy (0: n); perturbation and subdomains live on it */
    uint64_t source_seed;   /* "modes" source: seed and number (1..64) of random sinusoid modes */
    int source_modes;
    int defer_assembly;     /* 1: no host assembly at creation (use fh_hho_assemble_gpu) */
} fh_hho_args;
typedef struct fh_hho_info {
    int nb, nf, ncells, nfaces;
    int64_t nnzb;
    double mesh_s, local_s, assemble_s;
    /* vertex perturbation: draws rejected because a cell of the vertex would
     * fall below the quality floor 0.1 (Jacobian determinant / h^dim, 1 on
     * the unperturbed grid), vertices left on the grid after 64 rejected
     * draws, and the smallest cell quality of the mesh */
    int64_t redraws, kept_vertices;
    double min_quality;
} fh_hho_info;
int fh_hho_create(const fh_hho_args* args, fh_hho** out);
int fh_hho_destroy(fh_hho* h);
int fh_hho_get_info(const fh_hho* h, fh_hho_info* info);
/* Device BCSR of the condensed operator A (Diffusion_HHO::A). */
int fh_hho_matrix(fh_ctx* ctx, const fh_hho* h, int keep_host, fh_mat** out);
/* Host copies: condensed right-hand side b (nb*nf) and block arrays. */
int fh_hho_rhs(const fh_hho* h, double* b);
int fh_hho_blocks(const fh_hho* h, int* rowptr, int* colidx, double* vals);
/* Relative L2 error of the k+1 reconstruction of the face solution x (host)
 * against the exact solution (Diffusion_HHO::L2Error, Diffusion_HHO.h:120-149). */
int fh_hho_l2_error(const fh_hho* h, const double* x, double* err);
/* The same on the GPU: x_dev is the face solution in device memory (nb*nf
 * FP64).  Cell recovery, k+1 reconstruction and the quadrature of the error
 * run one wavefront per cell (Diffusion_HHO.h:120-149, 855-913, 960-966);
 * the per-cell integrals are summed on the host in cell order. */
int fh_hho_l2_error_gpu(fh_ctx* ctx, const fh_hho* h, const double* x_dev, double* err);
/* Geometric multigrid prolongation from the problem on the once-coarsened
 * mesh (coarse->n = fine->n / 2, same ref_n) to the fine one: nf x nf block
 * matrix (fine faces x coarse faces).  Replaces LevelForHHO::SetupProlongation_H
 * with CellInterp_Trace (src/Solver/Multigrid/MultigridForHHO/LevelForHHO.h:182-213). */
/* Local HHO operators (Diff_HHOElement::InitHHO, Diff_HHOElement.h:104-150,
 * 312-509), static condensation (Diffusion_HHO.h:624-652) and assembly on
 * the GPU: one workgroup per cell, the cell matrices in LDS.  *out receives
 * the device operator (host copy of the values with keep_host), b_dev (device,
 * nb*nf FP64) the condensed right-hand side. */
int fh_hho_assemble_gpu(fh_ctx* ctx, const fh_hho* h, int keep_host, fh_mat** out, double* b_dev);
int fh_hho_prolongation(fh_ctx* ctx, const fh_hho* fine, const fh_hho* coarse, int keep_host, fh_mat** out);
/* Cells of every interior face (2 ints per block row) and cell kappa. */
int fh_hho_face_cells(const fh_hho* h, int* out);
int fh_hho_cell_kappa(const fh_hho* h, double* out);
/* Uncondensed hybrid blocks kept for U-AMG (Diffusion_HHO::A_T_T, A_T_ndF,
 * A_ndF_ndF; Diffusion_HHO.h:52-53,410-470,564-624, kept when the solver is
 * uamg, Program_Diffusion_HHO.h:180-186): which = 0 A_TT (cells x cells,
 * cell_unknowns DOFs per cell), 1 A_TF (cells x non-Dirichlet face DOFs),
 * 2 A_FF.  Scalar CSR, ascending columns; sizes always, arrays when non-NULL
 * (rowptr nrows+1, colidx and vals nnz).  Assembled at the first call. */
int fh_hho_uncondensed(fh_hho* h, int which, int* nrows, int* ncols, int64_t* nnz, int* rowptr, int* colidx,
                       double* vals);
/* Number of cell unknowns per cell (HHO->nCellUnknowns). */
int fh_hho_cell_unknowns(const fh_hho* h, int* nc);

/* ------------------------------------- batched condensation and recovery */
/* Static condensation of the caller's element matrices on the GPU: the face
 * system A = sum over cells (A_FF - A_TF^T A_TT^-1 A_TF) and b = B_F - sum
 * over cells A_TF^T A_TT^-1 B_T.  Replaces the condensation step of
 * Diffusion_HHO::Assemble (Discretizations/HHO/Diffusion_HHO.h:624-637,
 * A = (A_ndF_ndF - A_T_ndF^T Solve_A_T_T(A_T_ndF)).pruned(), and
 * CondensedRHS :1026-1032) applied cell by cell to Diff_HHOElement::A
 * (Diff_HHOElement.h:144).
 *   ncells cells of nfc faces (nf unknowns each) and nc cell unknowns;
 *   nfaces system faces (block rows of the result);
 *   cell_faces: HOST int32 [ncells*nfc], the system number of each face of
 *     each cell in element order, -1 for a face whose unknowns are
 *     eliminated (Dirichlet faces, Diffusion_HHO.h:596-602);
 *   local: DEVICE FP64 [ncells*nl*nl], nl = nc + nfc*nf, each cell's local
 *     matrix row-major with the cell unknowns first, then the faces in
 *     element order (FirstDOFNumber, Diff_HHOElement.h:522-525); only the
 *     A_TT, A_TF and A_FF parts are read;
 *   B_T: DEVICE [ncells*nc] or NULL (zero); B_F: DEVICE [nfaces*nf] or NULL
 *     (zero); b_out: DEVICE [nfaces*nf], or NULL not to form b.
 * *out: a new FP64 device matrix of nfaces x nfaces blocks of nf x nf whose
 * pattern is every pair of faces of one cell (ascending columns); blocks
 * that sum to zero are kept (the reference prunes exact zeros of its scalar
 * pattern).  -4 if an A_TT is not positive definite, -3 if a cell's matrices
 * do not fit one workgroup's LDS. */
int fh_condense_batched(fh_ctx* ctx, int ncells, int nfc, int nc, int nf, int nfaces, const int* cell_faces,
                        const double* local, const double* B_T, const double* B_F, double* b_out, fh_mat** out);
/* Cell unknowns from the face solution, x_T = A_TT^-1 (B_T - A_TF x_F) per
 * cell (Diffusion_HHO::SolveCellUnknowns, Diffusion_HHO.h:960-966); faces
 * numbered -1 contribute nothing.  cell_faces, local and B_T as for
 * fh_condense_batched; x_F: DEVICE [nfaces*nf]; x_T: DEVICE [ncells*nc]. */
int fh_cell_recover_batched(fh_ctx* ctx, int ncells, int nfc, int nc, int nf, int nfaces, const int* cell_faces,
                            const double* local, const double* B_T, const double* x_F, double* x_T);

/* --------------------------------------------- uncondensed AMG (U-AMG) */
/* Host scalar CSR (RowMajor, int32 indices, strictly ascending columns in
 * every row): the reference's SparseMatrix (src/Utils/Types.h:11-22). */
typedef struct fh_csr {
    int nrows, ncols;
    const int* rowptr;
    const int* colidx;
    const double* vals;
} fh_csr;
/* IterativeSolver::Setup(A, A_T_T, A_T_F, A_F_F) (src/Solver/IterativeSolver.h:34-37,
 * FlexibleConjugateGradient.h:39-44) for a solver whose multigrid is "uamg"
 * (UncondensedAMG::Setup, src/Solver/Multigrid/UncondensedAMG/UncondensedAMG.h:84-98):
 * the hierarchy is built on the host from the uncondensed blocks with the
 * reference's default U-AMG options (main.cpp:1847-1868), then the levels run
 * on the GPU.  A must be created with keep_host. */
int fh_solver_setup_uncondensed(fh_solver* s, const fh_mat* A, const fh_csr* A_TT, const fh_csr* A_TF,
                                const fh_csr* A_FF, int cell_bs);
/* The same hierarchy without a solver (host only; no device needed):
 * UncondensedLevel::CoarsenMesh for every level (UncondensedLevel.h:74-185). */
typedef struct fh_uamg fh_uamg;
int fh_uamg_build(const fh_csr* A, const fh_csr* A_TT, const fh_csr* A_TF, const fh_csr* A_FF, int cell_bs,
                  int face_bs, double theta, int max_coarse, fh_uamg** out);
int fh_uamg_destroy(fh_uamg* u);
int fh_uamg_info(const fh_uamg* u, int* nlevels, int* npasses);
/* which = 0: operator of level l (Level::OperatorMatrix); which = 1:
 * prolongation from level l+1 to level l (Level::P, chained over the passes).
 * Sizes always, arrays when non-NULL. */
int fh_uamg_level(const fh_uamg* u, int l, int which, int* nrows, int* ncols, int64_t* nnz, int* rowptr, int* colidx,
                  double* vals);
/* Coarsening pass k (in order over all levels): its level and pass number,
 * counts = {elements, faces, coarse elements, coarse faces}, the aggregate of
 * every element (HybridElementAggregate::Number) and the coarse face of
 * every face (-1 for a face removed inside an aggregate). */
int fh_uamg_pass(const fh_uamg* u, int k, int* level, int* pass, int* counts, int* elem_agg, int* face_coarse);

#ifdef __cplusplus
}
#endif
#endif /* FHGPU_H */
