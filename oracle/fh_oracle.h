/* CPU oracle of the condensed-HHO solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU and in plain
 * C++, the reference algorithms of the solve path so that tests, the smoke
 * check and the CPU-baseline leg of the benchmark can check the HIP library
 * against it.  Nothing in the product links, loads or calls it.
 *
 * Matrices are scalar CSR exactly as the reference stores its condensed
 * operator (RowMajor, int32 indices, src/Utils/Types.h:11-22); `nf` is the
 * block size of the block smoothers (face-polynomial dimension).
 *
 * Every function that has a check pinning it is listed with that check in
 * tests/test_oracle.py.  All functions are pinned (golden fixtures, closed
 * forms or brute force); none is "parity unpinned".
 */
#pragma once
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* y = A x  (SparseDenseProduct of Eigen; order of summation: row entries in
 * column order). */
void or_spmv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y);

/* Inverse of every nf x nf diagonal block (BlockJacobi::Setup,
 * BlockJacobi.h:57-74 factorises them with FullPivLU; the inverse is the
 * same operator).  invD has (n/nf)*nf*nf entries, row-major per block.
 * Returns 0, or -(1+i) if block i is singular. */
int or_block_inverse(int n, int nf, const int* rp, const int* ci, const double* v, double* invD);

/* One sweep of damped block Jacobi (BlockJacobi.h:77-127):
 * x_i <- D_i^-1 [ w b_i + (1-w) D_i x_i - w sum_{j!=i} A_ij x_j ]. */
void or_bj_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                 double omega, const double* b, double* x);

/* One lexicographic block Gauss-Seidel sweep (BlockGaussSeidel.h:84-123,
 * 193-223): dir 0 forward, 1 backward, 2 symmetric (forward then backward). */
void or_bgs_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                  int dir, const double* b, double* x);

/* One multicolour block Gauss-Seidel sweep: colours visited in order
 * 0..ncol-1 (dir 0) or ncol-1..0 (dir 1); inside a colour the update uses the
 * values of the previous colours only (rows of one colour are not coupled). */
void or_mcbgs_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                    int ncol, const int* colour_ptr, const int* colour_rows, int dir,
                    const double* b, double* x);
/* Block SOR (BlockSOR.h:193-223): the Gauss-Seidel sweeps with relaxation
 * omega (omega = 1 is Gauss-Seidel), lexicographic and multicolour. */
void or_bsor_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                   int dir, double omega, const double* b, double* x);
void or_mcsor_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                    int ncol, const int* colour_ptr, const int* colour_rows, int dir, double omega,
                    const double* b, double* x);

/* Greedy distance-1 colouring of the block graph: block rows in increasing
 * order take the smallest colour unused by their already-coloured
 * neighbours.  Returns the number of colours; colour[i] per block row. */
int or_greedy_colouring(int nb, const int* brp, const int* bci, int* colour);

/* Algebraic mesh + one pairwise aggregation pass (AlgebraicMesh.h:54-153,
 * PairwiseAggregation.h:9-205).  Coupling of block rows i != j is
 * trace(A_ij); strong iff coupling < 0 and < theta * (most negative
 * coupling of i).  agg[i] = aggregate of block row i.  Returns the number of
 * aggregates; *coarsest set when the coarsening ratio is <= 1.01. */
int or_pairwise_aggregate(int n, int nf, const int* rp, const int* ci, const double* v, double theta,
                          int* agg, int* coarsest);

/* ---------------------------------------------------------------- solver */
typedef struct or_level_desc {
    /* operator of this level (level 0: the system matrix) */
    int n; const int* rp; const int* ci; const double* v;
    /* prolongation to this level from the next coarser one (NULL for the
     * coarsest level or for aggregation levels): n x nc scalar CSR */
    int nc; const int* prp; const int* pci; const double* pv;
    /* block size of this level (0: the configuration's nf), and, instead of
     * a prolongation matrix, a p-transfer to the next level keeping the first
     * pmodes modes of every block (0: use prp/pci/pv) */
    int nf; int pmodes;
} or_level_desc;

typedef struct or_solver_cfg {
    int krylov;        /* 0 none (stand-alone preconditioner iterations), 1 CG, 2 FCG(1) */
    int prec;          /* 0 identity, 1 block Jacobi, 2 block GS, 3 GMG (given levels), 4 aggregation AMG */
    int nf;            /* block size */
    double omega;      /* relaxation of block Jacobi */
    int pre_kind, post_kind;   /* 1 bj, 2 bgs forward, 3 bgs backward, 4 bgs symmetric,
                                  5/6/7 multicolour bgs forward/backward/symmetric,
                                  8/9/10/11 Chebyshev of degree 2/3/4/1 of D^-1 A
                                  on [0.25, 1.1] x its power-iteration lambda_max,
                                  12/13 alternating bgs (forward / backward first),
                                  14/15 the same, multicolour; the Gauss-Seidel
                                  kinds relax with pre/post_omega (SOR) */
    double pre_omega, post_omega;
    int nu1, nu2;
    char cycle;        /* 'V', 'W' or 'K' */
    int wloops;
    int max_coarse;    /* MatrixMaxSizeForCoarsestLevel (AMG) */
    double theta;      /* strong coupling threshold (AMG) */
    double tol;
    int max_iter;
    int nlevels;       /* GMG: number of levels given in `levels` */
    int coarse_kind;   /* > 0: smoother kind of the levels below the finest (both sweeps) */
    double coarse_omega;
    int fcg_truncation;  /* FCG: number of kept directions (<= 0: 1) */
    int fcg_restart;     /* FCG: clear instead of dropping the oldest */
    int kcycle_levels;   /* K-cycle only on levels 1..kcycle_levels (<= 0: all), V recursion below */
    int x_given;         /* nonzero: x holds the initial guess (IterativeSolver::Solve(b, x, false)) */
    char nu_change_op;   /* Multigrid.h:224-242: '+', '-', '*' or 0; each level's smoothing */
    int nu_change_coeff; /* iterations are its finer level's changed by op with coeff (>= 0) */
    int coarse_nu_given; /* nonzero: levels below the finest take coarse_nu1 / coarse_nu2 */
    int coarse_nu1, coarse_nu2;
} or_solver_cfg;

typedef struct or_solver_out {
    int iterations;
    int nlevels;
    double final_res;
    int level_n[32];
} or_solver_out;

/* Solves A x = b from x = 0 with the configured method.  hist receives the
 * normalised residual at every stopping test (length iterations+1, capacity
 * max_iter+1).  Returns 0 on convergence, 1 when max_iter is reached,
 * negative on error. */
int or_solve(const or_solver_cfg* cfg, const or_level_desc* levels, const double* b, double* x,
             double* hist, or_solver_out* out);
/* The same with the system matrix in block storage (nb block rows of nf x nf
 * dense blocks, cfg->nf; int32 block indices, 64-bit value offsets), for
 * systems whose scalar entries overflow int32: no preconditioner, or the
 * p-first aggregation AMG (prec 4) with Chebyshev smoothing. */
int or_solve_bcsr(const or_solver_cfg* cfg, int nb, const int* brp, const int* bci, const double* bv, const double* b,
                  double* x, double* hist, or_solver_out* out);

/* Aggregation AMG hierarchy only: number of levels, unknowns per level, and
 * the first-level aggregate map (length n/nf). */
int or_amg_hierarchy(int n, int nf, const int* rp, const int* ci, const double* v, double theta,
                     int max_coarse, int* agg0, int* level_n, int max_levels);

/* Seconds spent by the last or_solve in its setup and in its iterations
 * (wall clock), for the CPU baseline. */
void or_last_times(double* setup_s, double* solve_s);

/* ------------------------------------------------ learned post-smoother
 * Oracle of the message-passing network of the HIP library (FP64): the
 * operator is given in block-CSR form (nb block rows of nf x nf row-major
 * blocks); feat holds 2 features per block row; params are laid out as
 * We (32 x (nf+2)), be (32), then per layer S (32 x 32), N (32 x 32), c (32),
 * then Wo (nf x 32), bo (nf).
 *   v = D^-1 r1, sigma = RMS(v), u = v / sigma, x0 = [u, feat]
 *   h = relu(We x0 + be); per layer m_i = sum_j w_ij h_j (w_ij =
 *   |tr A_ij| / sum_{k != i} |tr A_ik|, w_ii = 0), h = relu(S h + N m + c)
 *   delta = sigma (u + Wo h + bo)
 * Loss of a pair (r1 = A e1, e1): (e1 - delta)^T A (e1 - delta) / (e1^T r1).
 * No reference counterpart: pinned by central finite differences of the
 * loss (tests/test_oracle.py). */
int or_mpnn_num_params(int nf, int layers);
void or_mpnn_forward(int nb, int nf, const int* brp, const int* bci, const double* bval, const double* feat,
                     int layers, const double* params, const double* r1, double* delta);
double or_mpnn_loss_grad(int nb, int nf, const int* brp, const int* bci, const double* bval, const double* feat,
                         int layers, const double* params, const double* r1, const double* e1, double* grad);

#ifdef __cplusplus
}
#endif
