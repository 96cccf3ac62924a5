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
#include <string>
#include <vector>

#include "fh_common.h"
#include "fh_matrix.h"

namespace fh {

enum class SmootherKind { BlockJacobi, BlockGS };
enum class Sweep { Forward = 0, Backward = 1, Symmetric = 2 };

// Smoother codes of SmootherFactory::Create (SmootherFactory.h:9-104)
// restricted to the block smoothers: bj, bj23, j, j23, bgs, rbgs, sbgs, gs,
// rgs, sgs (Gauss-Seidel runs as multicolour block Gauss-Seidel).
struct SmootherSpec {
    SmootherKind kind = SmootherKind::BlockJacobi;
    Sweep sweep = Sweep::Forward;
    double omega = 1.0;
    int iterations = 1;
    static SmootherSpec parse(const std::string& code, int iterations, double relax);
};

// Per-level smoother: block inverses of the diagonal (FP64 -> dprec) and,
// for Gauss-Seidel, a colouring of the block graph.
struct LevelSmoother {
    SmootherSpec spec;
    DevBuf invD;
    Prec dprec = Prec::F64;
    int ncolours = 0;
    std::vector<int> colour_ptr;   // host: colour c = rows[colour_ptr[c] .. colour_ptr[c+1])
    DevBuf colour_rows;            // device row lists, grouped by colour
};

struct SolveStats {
    int iterations = 0;
    double final_res = 0;
    double setup_s = 0;
    double solve_s = 0;
    int levels = 0;
    int converged = 0;
    double op_complexity = 0;
};

class Multigrid;

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
    bool galerkin = false;           // geometric multigrid: coarse operators P^T A P (UseGalerkinOperator)
};

class Solver {
public:
    explicit Solver(Ctx& c, const SolverArgs& a);
    ~Solver();
    // GMG hierarchy input: operator and prolongation of level l (l >= 1 for
    // the operator; P of level l maps level l+1 to level l).
    void SetLevel(int l, const Bcsr* A, const Bcsr* P);
    void Setup(const Bcsr& A);
    // Solves A x = b from x = 0 (FP64 device vectors of length A.rows()).
    void Solve(const double* b, double* x);
    const SolveStats& Stats() const { return stats_; }
    const std::vector<double>& History() const { return hist_; }
    int NumLevels() const;
    int LevelUnknowns(int l) const;

private:
    Ctx& c_;
    SolverArgs args_;
    const Bcsr* A_ = nullptr;
    std::unique_ptr<Multigrid> mg_;
    std::unique_ptr<LevelSmoother> prec_sm_;   // bj / bgs as preconditioner
    std::vector<std::pair<const Bcsr*, const Bcsr*>> given_levels_;
    SolveStats stats_;
    std::vector<double> hist_;
    DVec r_, z_, d_, Ad_, Az_, dk_, Adk_, t_;
    bool fused_ = false;
    void apply_prec(const DVec& r, DVec& z, DVec* Az);
    double read_scalar(int slot);
    void solve_cg(const double* b, double* x);
    void solve_fcg(const double* b, double* x);
    void solve_stationary(const double* b, double* x);
};

// Smoother application on one level (in the level's working precision).
void smoother_setup(Ctx& c, const Bcsr& A64, LevelSmoother& S, Prec dprec);
void smoother_apply(Ctx& c, const Bcsr& A, LevelSmoother& S, Prec vp, const void* b, DVec& x, DVec& work,
                    bool& x_is_zero);

}  // namespace fh
