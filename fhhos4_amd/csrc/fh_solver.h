// Host-side Solver API of the MI355X solve path.
//
// It keeps the reference's solver objects and their semantics -- Solver /
// IterativeSolver (src/Solver/Solver.h:6-44, IterativeSolver.h:6-248),
// Preconditioner (Preconditioner.h:100-199), ConjugateGradient
// (Krylov/ConjugateGradient.h:27-104), FlexibleConjugateGradient
// (Krylov/FlexibleConjugateGradient.h:47-171), Multigrid and Level
// (Multigrid/Multigrid.h:88-656, Level.h:7-222), the block smoothers
// (FixedPoint/BlockJacobi.h, BlockGaussSeidel.h) and the aggregation AMG
// (Multigrid/AggregAMG/*) -- while every data-parallel step runs as a gfx950
// kernel on the context's stream.  Vectors live in HBM; the host reads back
// one scalar per Krylov iteration (the residual norm for the stopping test)
// and, in the K-cycle, the two norms that decide its branch.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "fh_common.h"
#include "fh_matrix.h"
#include "fh_uamg.h"

namespace fh {

enum class SmootherKind { BlockJacobi, BlockGS, Chebyshev };
// Alternating sweeps go forward on even smoothing iterations and backward
// on odd ones (AltForwardFirst), or the other way (BlockGaussSeidel.h:105-120,
// BlockSOR.h:153-182: IterationCount % 2).
enum class Sweep { Forward = 0, Backward = 1, Symmetric = 2, AltForwardFirst = 3, AltBackwardFirst = 4 };

// Smoother codes of SmootherFactory::Create (SmootherFactory.h:9-104)
// restricted to the block smoothers: bj, bj23, j, j23 (block Jacobi); bgs,
// rbgs, sbgs, abgs, rabgs, gs, rgs, sgs, ags, rags (Gauss-Seidel) and bsor,
// rbsor, sbsor, absor, rabsor, sor, rsor (SOR: Gauss-Seidel with the
// relaxation parameter; bgs with relax != 1 is SOR too, as in the
// reference); hbgs, hrbgs (the reference's chunk-parallel sweep).  Every
// Gauss-Seidel / SOR sweep runs as a multicolour block sweep.
struct SmootherSpec {
    SmootherKind kind = SmootherKind::BlockJacobi;
    Sweep sweep = Sweep::Forward;
    double omega = 1.0;
    int iterations = 1;
    int degree = 2;  // Chebyshev: polynomial degree per smoothing iteration
    static SmootherSpec parse(const std::string& code, int iterations, double relax);
};

// Chebyshev smoothing targets [kChebLow, kChebHigh] x the power-iteration
// estimate of lambda_max(D^-1 A) (kChebPowerIts iterations from a fixed
// start vector).
constexpr double kChebLow = 0.25, kChebHigh = 1.1;
constexpr int kChebPowerIts = 20;

// Per-level smoother: block inverses of the diagonal (FP64 -> dprec) and,
// for Gauss-Seidel, a colouring of the block graph.
struct LevelSmoother {
    SmootherSpec spec;
    std::shared_ptr<DevBuf> invD;  // block inverses (shared by a level's pre- and post-smoother)
    Prec dprec = Prec::F64;
    int ncolours = 0;
    std::vector<int> colour_ptr;   // host: colour c = rows[colour_ptr[c] .. colour_ptr[c+1])
    DevBuf colour_rows;            // device row lists, grouped by colour
    double lmax = 0;               // Chebyshev: estimate of lambda_max(D^-1 A)
    DVec d;                        // Chebyshev direction vector
};

// Power-iteration estimate of lambda_max(D^-1 A) in FP64.
double estimate_lmax(Ctx& c, const Bcsr& A64, const void* invD64);

struct SolveStats {
    int nn_used = 0, nn_fallback = 0;  // learned corrections accepted / block-Jacobi fallbacks
    int iterations = 0;
    double final_res = 0;
    double setup_s = 0;
    double solve_s = 0;
    int levels = 0;
    int converged = 0;
    double op_complexity = 0;
    int graph_mode = 0;  // multigrid cycle: 0 eager (undecided), 1 eager (faster), 2 graph replay
};

class Multigrid;
class Mpnn;
struct NnGraph;

// Krylov / fixed-point solver created by string codes (SolverFactory.h:126-233).
struct SolverArgs {
    std::string solver = "cg";       // cg | fcg | mg | aggregamg | bj | bgs
    std::string prec = "mg";         // "" | none | bj | bj23 | bgs | mg | aggregamg
    char cycle = 'V';
    int wloops = 1;
    int nu1 = 1, nu2 = 1;
    std::string pre_smoother = "bj", post_smoother = "bj";
    double relax = 1.0;
    double tol = 1e-8;
    int max_iter = 200;
    int max_coarse = 1000;           // MatrixMaxSizeForCoarsestLevel
    double theta = 0.25;             // strong coupling threshold
    Prec mg_prec = Prec::F64;        // working precision of the preconditioner vectors
    Prec mg_store = Prec::F64;       // storage of level matrices (F64 / F32 / BF16)
    Prec coarse_store = Prec::F64;   // storage of the levels below the finest (set to mg_store unless given)
    bool galerkin = false;           // geometric multigrid: coarse operators P^T A P (UseGalerkinOperator)
    // Smoother of the levels below the finest ("" = the same as the finest)
    // and its relaxation (<= 0: relax).  Galerkin coarse levels are dense
    // (50-100 blocks per row), so a multicolour sweep there costs 50-100
    // launches; block Jacobi costs one.
    std::string coarse_smoother;
    double coarse_relax = 0.0;
    int fcg_truncation = 1;          // FlexibleConjugateGradient::Truncation (FCG(1) by default)
    bool fcg_restart = false;        // FlexibleConjugateGradient::Restart
    // K-cycle only on levels 1 .. kcycle_levels (0: every level below the
    // finest, the reference's K-cycle); V-cycle recursion below.  Not in the
    // reference: bounds the number of visits of the coarsest levels, which
    // doubles with every K level.
    int kcycle_levels = 0;
    // Multigrid.h:224-242: smoothing iterations of each coarser level are its
    // finer level's changed by op ('+', '-', '*'; 0 = unchanged) with coeff
    char nu_change_op = 0;
    int nu_change_coeff = 0;
    // FCG's own product A p with the FP32 copy of the finest multigrid level
    // (FP64 vectors), the recursive residual replaced by the FP64 b - A x
    // every replace_every iterations and before convergence is declared
    Prec outer_prec = Prec::F64;
    int replace_every = 10;
    // smoothing iterations of the levels below the finest, when >= 0 (not in
    // the reference); -1: the finest level's, changed by nu_change_op
    int coarse_nu1 = -1, coarse_nu2 = -1;
    // smoothing iterations of level l from the finest level's nu
    int level_nu(int nu, int l) const {
        for (int k = 0; k < l; k++) {
            if (nu_change_op == '+') nu += nu_change_coeff;
            else if (nu_change_op == '-') nu -= nu_change_coeff;
            else if (nu_change_op == '*') nu *= nu_change_coeff;
        }
        return nu < 0 ? 0 : nu;
    }
};

class Solver {
public:
    explicit Solver(Ctx& c, const SolverArgs& a);
    ~Solver();
    // GMG hierarchy input: operator and prolongation of level l (l >= 1 for
    // the operator; P of level l maps level l+1 to level l).
    void SetLevel(int l, const Bcsr* A, const Bcsr* P);
    // The transfer from level l + 1 to level l is the p-injection keeping the
    // first nfc modes of every face block (hierarchical bases; the natural
    // injection of P_LevelForHHO::SetupProlongation, P_MultigridForHHO.h:62-86);
    // the operator of level l + 1 is then its Galerkin operator, the leading
    // nfc x nfc sub-blocks (unless SetLevel gives one).
    void SetLevelP(int l, int nfc);
    // Learned post-smoother inside the preconditioner of FCG: after the
    // multigrid cycle z0 = M r, r1 = r - A z0, delta = net(r1); z = z0 +
    // (delta.r1 / delta.A delta) delta when delta.r1 > 0 (unit_step: z0 +
    // delta when delta.A delta < 2 delta.r1), else the block-Jacobi step
    // z0 + omega D^-1 r1.  The network and graph must outlive the solver.
    void SetLearnedSmoother(Mpnn* nn, const NnGraph* g, double omega, bool unit_step = false);
    void Setup(const Bcsr& A);
    // IterativeSolver::Setup(A, A_T_T, A_T_F, A_F_F) (IterativeSolver.h:34-37)
    // for the "uamg" multigrid (UncondensedAMG::Setup, UncondensedAMG.h:84-98):
    // builds the hierarchy from the uncondensed blocks on the host, then sets
    // up as with given levels.  A needs host values (keep_host).
    void SetupUncondensed(const Bcsr& A, const Csr& ATT, const Csr& ATF, const Csr& AFF, int cell_bs);
    const UamgHierarchy* Uamg() const { return uamg_.get(); }
    // Solves A x = b from x = 0 (FP64 device vectors of length A.rows()).
    // x holds the initial guess unless x_is_zero (then it is set to zero)
    void Solve(const double* b, double* x, bool x_is_zero = true);
    // IterativeSolver::Tolerance / MaxIterations (IterativeSolver.h:11-23).
    void SetStop(double tol, int max_iter);
    // The configured preconditioner alone: z = M r (Preconditioner::Apply,
    // Preconditioner.h:100-118; one cycle from 0 for a multigrid).
    void ApplyPreconditioner(const double* r, double* z);
    const SolveStats& Stats() const { return stats_; }
    const std::vector<double>& History() const { return hist_; }
    int NumLevels() const;
    // Level::Setup with ExportComponents (Level.h:116-125): writes
    // <dir>/level<l>_A.dat for every level and level<l>_P.dat, level<l>_R.dat
    // for every level but the coarsest, in the exported-system format.
    void ExportLevels(const std::string& dir);
    int LevelUnknowns(int l) const;

private:
    Ctx& c_;
    SolverArgs args_;
    const Bcsr* A_ = nullptr;
    std::unique_ptr<Multigrid> mg_;
    std::unique_ptr<LevelSmoother> prec_sm_;   // bj / bgs as preconditioner
    std::vector<std::pair<const Bcsr*, const Bcsr*>> given_levels_;
    std::vector<int> given_p_;  // kept modes of a p-transfer below level l (0: none)
    int nn_fallbacks_seen_ = 0;  // block-Jacobi fallbacks of the learned post-smoother in this solve
    bool nn_f32_products_ = false;  // the learned post-smoother's products used FP32-stored values
    // U-AMG hierarchy and the level matrices made from it
    std::unique_ptr<UamgHierarchy> uamg_;
    std::vector<std::unique_ptr<Bcsr>> uamg_mats_;
    SolveStats stats_;
    std::vector<double> hist_;
    DVec r_, z_, d_, Ad_, Az_, dk_, Adk_, t_;
    std::vector<DVec> fd_, fAd_;  // FCG(m) direction ring
    bool fused_ = false;
    Mpnn* nn_ = nullptr;
    const NnGraph* nng_ = nullptr;
    double nn_omega_ = 1.0;
    bool nn_unit_step_ = false;
    DVec r1_, delta_, Adelta_;
    DevBuf nn_counters_;
    void apply_prec(const DVec& r, DVec& z, DVec* Az);
    double read_scalar(int slot);
    void solve_cg(const double* b, double* x, bool xz);
    void solve_fcg(const double* b, double* x, bool xz);
    void solve_stationary(const double* b, double* x, bool xz);
    // r = b (x = 0) or b - A x; returns ||b|| and the first normalised residual
    double initial_residual(const double* b, double* x, bool xz, double& res);
};

// Smoother application on one level (in the level's working precision).
void smoother_setup(Ctx& c, const Bcsr& A64, LevelSmoother& S, Prec dprec);
void smoother_apply(Ctx& c, const Bcsr& A, LevelSmoother& S, Prec vp, const void* b, DVec& x, DVec& work,
                    bool& x_is_zero);

}  // namespace fh
