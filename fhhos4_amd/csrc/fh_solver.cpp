// Host-side Solver API: Krylov drivers, multigrid hierarchy and cycles.
#include "fh_solver.h"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <utility>

#include "fh_io.h"
#include "fh_kernels.h"
#include "fh_nn.h"

namespace fh {

// Device scalar slots.
enum : int {
    S_RZ = 0, S_RZOLD = 1, S_DAD = 2, S_RR = 3, S_RD = 4, S_ZADK = 5, S_DKADK = 6, S_BB = 7,
    S_KBASE = 32,  // K-cycle: 8 slots per level
    S_FCG = 400,   // FCG(m): 2 slots per kept direction (m <= 32)
};

Ctx::Ctx(int dev) : device(dev) {
    if (dev < 0) {
        host_only = true;
        return;
    }
    FH_HIP(hipSetDevice(dev));
    FH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    FH_HIP(hipMalloc(&d_partials, sizeof(double) * kMaxPartials * kPartialSlots));
    FH_HIP(hipMalloc(&d_scalars, sizeof(double) * kScalars));
    FH_HIP(hipHostMalloc(&h_scalars, sizeof(double) * kScalars, hipHostMallocDefault));
    FH_HIP(hipMemsetAsync(d_scalars, 0, sizeof(double) * kScalars, stream));
    FH_HIP(hipStreamSynchronize(stream));
    noalloc_scratch_init(dev);  // this device is current
}

Ctx::~Ctx() {
    if (host_only) return;
    if (stream) (void)hipStreamSynchronize(stream);
    if (d_partials) (void)hipFree(d_partials);
    if (d_scalars) (void)hipFree(d_scalars);
    if (h_scalars) (void)hipHostFree(h_scalars);
    if (stream) (void)hipStreamDestroy(stream);
}

static void read_scalars(Ctx& c, int first, int count) {
    FH_HIP(hipMemcpyAsync(c.h_scalars + first, c.d_scalars + first, sizeof(double) * count, hipMemcpyDeviceToHost,
                          c.stream));
    FH_HIP(hipStreamSynchronize(c.stream));
}

static void dot1(Ctx& c, int64_t n, Prec p, const void* x, const void* y, int slot) {
    const void* xs[1] = {x};
    const void* ys[1] = {y};
    int sl[1] = {slot};
    launch_dots(c, n, p, 1, xs, ys, sl);
}
static void dot2(Ctx& c, int64_t n, Prec p, const void* x0, const void* y0, int s0, const void* x1, const void* y1,
                 int s1) {
    const void* xs[2] = {x0, x1};
    const void* ys[2] = {y0, y1};
    int sl[2] = {s0, s1};
    launch_dots(c, n, p, 2, xs, ys, sl);
}
static void dot3(Ctx& c, int64_t n, Prec p, const void* x0, const void* y0, int s0, const void* x1, const void* y1,
                 int s1, const void* x2, const void* y2, int s2) {
    const void* xs[3] = {x0, x1, x2};
    const void* ys[3] = {y0, y1, y2};
    int sl[3] = {s0, s1, s2};
    launch_dots(c, n, p, 3, xs, ys, sl);
}
static void copy(Ctx& c, int64_t n, Prec p, const void* s, void* d) {
    FH_HIP(hipMemcpyAsync(d, s, (size_t)n * prec_bytes(p), hipMemcpyDeviceToDevice, c.stream));
}

// --------------------------------------------------------------- smoothers
SmootherSpec SmootherSpec::parse(const std::string& code, int iterations, double relax) {
    SmootherSpec s;
    s.iterations = iterations;
    if (code == "bj" || code == "j") { s.kind = SmootherKind::BlockJacobi; s.omega = relax; }
    else if (code == "bj23" || code == "j23") { s.kind = SmootherKind::BlockJacobi; s.omega = 2.0 / 3.0; }
    else if (code == "bgs" || code == "gs" || code == "bsor" || code == "sor" || code == "hbgs") {
        s.kind = SmootherKind::BlockGS; s.sweep = Sweep::Forward;
    } else if (code == "rbgs" || code == "rgs" || code == "rbsor" || code == "rsor" || code == "hrbgs") {
        s.kind = SmootherKind::BlockGS; s.sweep = Sweep::Backward;
    } else if (code == "sbgs" || code == "sgs" || code == "sbsor") {
        s.kind = SmootherKind::BlockGS; s.sweep = Sweep::Symmetric;
    } else if (code == "abgs" || code == "ags" || code == "absor") {
        s.kind = SmootherKind::BlockGS; s.sweep = Sweep::AltForwardFirst;
    } else if (code == "rabgs" || code == "rags" || code == "rabsor") {
        s.kind = SmootherKind::BlockGS; s.sweep = Sweep::AltBackwardFirst;
    }
    else if (code == "cheb1" || code == "cheb2" || code == "cheb3" || code == "cheb4") {
        s.kind = SmootherKind::Chebyshev;
        s.degree = code[4] - '0';
        return s;
    }
    else throw Error(-2, "Unmanaged smoother code: '" + code + "'");
    if (s.kind == SmootherKind::BlockGS) {
        // SmootherFactory.h:27-80: the Gauss-Seidel codes with omega != 1 are
        // BlockSOR (x_i <- x_i + omega D_i^-1 (b_i - sum_j A_ij x_j))
        FH_CHECK(relax > 0 && relax < 2, -2, "SOR relaxation must lie in (0, 2)");
        s.omega = relax;
    }
    return s;
}

double estimate_lmax(Ctx& c, const Bcsr& A64, const void* invD64) {
    const int64_t n = A64.rows();
    std::vector<double> h(n);
    for (int64_t i = 0; i < n; i++) h[i] = 1.0 + (double)((i * 7919) % 13) / 13.0;
    DVec v(n, Prec::F64), t(n, Prec::F64), w(n, Prec::F64), zero(n, Prec::F64);
    FH_HIP(hipMemcpyAsync(v.data(), h.data(), sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    launch_fill(c, n, zero.data(), Prec::F64, 0.0);
    double lam = 0;
    const int slot = Ctx::kScalars - 8;
    for (int it = 0; it < kChebPowerIts; it++) {
        // t = v + D^-1 (0 - A v) ; w = v - t = D^-1 A v
        launch_bcsr(c, A64, OP_SMOOTH, Prec::F64, v.data(), zero.data(), t.data(), 1.0, 0.0, 1.0, invD64, Prec::F64,
                    nullptr, 0);
        FH_HIP(hipMemcpyAsync(w.data(), v.data(), sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
        launch_axpby(c, n, Prec::F64, t.data(), w.data(), -1.0, -1, -1, 1.0, -1, -1);
        dot2(c, n, Prec::F64, w.data(), w.data(), slot, v.data(), v.data(), slot + 1);
        read_scalars(c, slot, 2);
        const double nw = std::sqrt(c.h_scalars[slot]), nv = std::sqrt(c.h_scalars[slot + 1]);
        FH_CHECK(nw > 0 && nv > 0, -6, "power iteration met a zero vector");
        lam = nw / nv;
        launch_axpby(c, n, Prec::F64, w.data(), v.data(), 1.0 / nw, -1, -1, 0.0, -1, -1);
    }
    c.sync();
    return lam;
}

void smoother_setup(Ctx& c, const Bcsr& A64, LevelSmoother& S, Prec dprec) {
    S.dprec = dprec;
    const int nf = A64.bs_r;
    std::vector<int> diag = diag_blocks(A64);
    DevBuf d_diag(sizeof(int) * diag.size());
    FH_HIP(hipMemcpyAsync(d_diag.p, diag.data(), sizeof(int) * diag.size(), hipMemcpyHostToDevice, c.stream));
    S.invD = std::make_shared<DevBuf>((size_t)A64.nbr * nf * nf * prec_bytes(dprec));
    DevBuf flag(sizeof(int));
    FH_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), c.stream));
    launch_block_inverse(c, A64, d_diag.as<int>(), S.invD->p, dprec, flag.as<int>());
    int h_flag = 0;
    FH_HIP(hipMemcpyAsync(&h_flag, flag.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    FH_CHECK(h_flag == 0, -6, "singular diagonal block in block row " + std::to_string(h_flag - 1));
    if (S.spec.kind == SmootherKind::Chebyshev) {
        if (dprec == Prec::F64)
            S.lmax = estimate_lmax(c, A64, S.invD->p);
        else {
            LevelSmoother S64;
            smoother_setup(c, A64, S64, Prec::F64);
            S.lmax = estimate_lmax(c, A64, S64.invD->p);
        }
        S.d.resize(A64.rows(), dprec);
    }
    if (S.spec.kind == SmootherKind::BlockGS) {
        std::vector<int> col;
        S.ncolours = greedy_colouring(A64, col);
        S.colour_ptr.assign(S.ncolours + 1, 0);
        for (int cc : col) S.colour_ptr[cc + 1]++;
        for (int k = 0; k < S.ncolours; k++) S.colour_ptr[k + 1] += S.colour_ptr[k];
        std::vector<int> rows(A64.nbr), pos(S.colour_ptr.begin(), S.colour_ptr.end() - 1);
        for (int i = 0; i < A64.nbr; i++) rows[pos[col[i]]++] = i;
        S.colour_rows.alloc(sizeof(int) * rows.size());
        FH_HIP(hipMemcpyAsync(S.colour_rows.p, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, c.stream));
        c.sync();
    }
}

// One multicolour block SOR sweep (omega = 1: Gauss-Seidel): the rows of a
// colour are uncoupled, so x_i <- x_i + omega D_i^-1 (b_i - sum_j A_ij x_j)
// over one colour at a time, in place, is the sequential sweep in the order
// of the colours (BlockSOR::ProcessBlockRow, BlockSOR.h:193-223).
static void gs_sweep(Ctx& c, const Bcsr& A, LevelSmoother& S, Prec vp, const void* b, DVec& x, bool backward) {
    for (int k = 0; k < S.ncolours; k++) {
        int cc = backward ? S.ncolours - 1 - k : k;
        int cnt = S.colour_ptr[cc + 1] - S.colour_ptr[cc];
        const int* rows = S.colour_rows.as<int>() + S.colour_ptr[cc];
        launch_bcsr(c, A, OP_SMOOTH, vp, x.data(), b, x.data(), 1.0, 0.0, S.spec.omega, S.invD->p, S.dprec, rows, cnt);
    }
}

void smoother_apply(Ctx& c, const Bcsr& A, LevelSmoother& S, Prec vp, const void* b, DVec& x, DVec& work,
                    bool& x_is_zero) {
    const int nf = A.bs_r;
    for (int it = 0; it < S.spec.iterations; it++) {
        if (S.spec.kind == SmootherKind::BlockJacobi) {
            if (x_is_zero)
                launch_bjac_zero(c, A.nbr, nf, vp, b, x.data(), S.spec.omega, S.invD->p, S.dprec, nullptr, 0);
            else {
                launch_bcsr(c, A, OP_SMOOTH, vp, x.data(), b, work.data(), 1.0, 0.0, S.spec.omega, S.invD->p, S.dprec,
                            nullptr, 0);
                std::swap(x.buf, work.buf);
            }
        } else if (S.spec.kind == SmootherKind::Chebyshev) {
            // Chebyshev polynomial of D^-1 A on [lo, hi] (three-term recurrence)
            const double hi = kChebHigh * S.lmax, lo = kChebLow * S.lmax;
            const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
            double rho = 1.0 / sigma;
            for (int k = 0; k < S.spec.degree; k++) {
                if (k == 0 && x_is_zero) {
                    // from x = 0: d = D^-1 b / theta and x = d in one pass
                    launch_bjac_zero(c, A.nbr, nf, vp, b, x.data(), 1.0 / theta, S.invD->p, S.dprec, nullptr, 0,
                                     S.d.data());
                } else {
                    // d = a d + c D^-1 (b - A x) and x + d in one block-row pass
                    // into work, which then becomes x
                    double a = 0.0, cc = 1.0 / theta;
                    if (k > 0) {
                        const double rn = 1.0 / (2.0 * sigma - rho);
                        a = rn * rho;
                        cc = 2.0 * rn / delta;
                        rho = rn;
                    }
                    launch_bcsr_cheb(c, A, vp, x.data(), b, work.data(), S.d.data(), a, cc, S.invD->p, S.dprec);
                    std::swap(x.buf, work.buf);
                }
                x_is_zero = false;
            }
        } else {
            if (x_is_zero) launch_fill(c, x.n, x.data(), vp, 0.0);
            Sweep sw = S.spec.sweep;
            if (sw == Sweep::AltForwardFirst || sw == Sweep::AltBackwardFirst)
                sw = (it % 2 == (sw == Sweep::AltForwardFirst ? 0 : 1)) ? Sweep::Forward : Sweep::Backward;
            if (sw == Sweep::Forward || sw == Sweep::Symmetric) gs_sweep(c, A, S, vp, b, x, false);
            if (sw == Sweep::Backward || sw == Sweep::Symmetric) gs_sweep(c, A, S, vp, b, x, true);
        }
        x_is_zero = false;
    }
}

// --------------------------------------------------------------- multigrid
struct MgLevel {
    const Bcsr* A64 = nullptr;
    std::unique_ptr<Bcsr> A64_own, Aw_own;
    const Bcsr* Aw = nullptr;
    LevelSmoother pre, post;
    // transfer to the next coarser level
    bool agg = false;
    const Bcsr* P64 = nullptr;
    std::unique_ptr<Bcsr> P_own, R64_own, Pw_own, Rw_own;
    const Bcsr* Pw = nullptr;
    const Bcsr* Rw = nullptr;
    DevBuf agg_map, agg_ptr, agg_members;
    int nagg = 0;
    bool ptrunc = false;  // p-transfer keeping the first nfc modes of every block
    int nfc = 1;
    // work vectors (working precision)
    DVec b, x, r, work, kc, kv, kd, kw;
    int64_t n = 0;
};

class Multigrid {
public:
    Multigrid(Ctx& c, const SolverArgs& a) : c_(c), a_(a), vp_(a.mg_prec) {}
    std::vector<MgLevel> L;
    DevBuf coarse_inv;
    int coarse_n = 0;
    double op_complexity = 0;

    void setup(const Bcsr& A, const std::vector<std::pair<const Bcsr*, const Bcsr*>>& given, bool amg,
               const std::vector<int>& pmodes = {}) {
        L.clear();
        L.emplace_back();
        L[0].A64 = &A;
        if (amg) build_amg();
        else {
            auto pm = [&pmodes](size_t l) { return l < pmodes.size() ? pmodes[l] : 0; };
            for (size_t l = 1; l < given.size(); l++) {
                FH_CHECK(a_.galerkin || pm(l - 1) > 0 || given[l].first != nullptr, -2,
                         "missing operator of level " + std::to_string(l));
                L.emplace_back();
                L[l].A64 = given[l].first;
            }
            for (size_t l = 0; l + 1 < given.size(); l++) {
                if (pm(l) > 0) {
                    // p-injection of hierarchical bases: keep the first nfc
                    // modes; Galerkin operator = leading sub-blocks
                    FH_CHECK(pm(l) < L[l].A64->bs_r, -2, "a p-transfer must keep fewer modes than the level has");
                    L[l].ptrunc = true;
                    L[l].nfc = pm(l);
                    if (!L[l + 1].A64) {
                        L[l + 1].A64_own = bcsr_mode_submatrix(c_, *L[l].A64, pm(l));
                        L[l + 1].A64 = L[l + 1].A64_own.get();
                    }
                    FH_CHECK(L[l + 1].A64->nbr == L[l].A64->nbr && L[l + 1].A64->bs_r == pm(l), -2,
                             "the operator below a p-transfer must have the kept modes of the same faces");
                    continue;
                }
                FH_CHECK(given[l].second != nullptr, -2, "missing prolongation of level " + std::to_string(l));
                L[l].P64 = given[l].second;
                if (a_.galerkin) {
                    // coarse operator R A P (Level::ComputeGalerkinOperator, Level.h:133-137)
                    L[l + 1].A64_own = galerkin_rap(c_, *L[l].A64, *L[l].P64);
                    L[l + 1].A64 = L[l + 1].A64_own.get();
                }
                FH_CHECK(L[l].P64->nbr == L[l].A64->nbr && L[l].P64->nbc == L[l + 1].A64->nbr, -2,
                         "prolongation shape does not match the level operators");
                L[l].R64_own = bcsr_transpose(c_, *L[l].P64);  // R = P^T (LevelForHHO.h:535-543)
            }
        }
        finish();
    }

    ~Multigrid() {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    }

    // Preconditioner application on FP64 vectors: z = M r (one cycle from 0).
    // No cycle has a host decision inside (the K-cycle's coefficients and
    // its early exit are computed on the device), so the whole cycle can be
    // captured once per (r, z) pair into a HIP graph and replayed.  Whether
    // replay beats eager launching depends on the cycle (measured on MI355X:
    // faster for ~250 launches per cycle, slower for ~1000), so the first
    // application of a pair times one eager cycle and one replay and keeps
    // the faster.  Cycles run eagerly while launches are being timed.  The
    // graph holds no memset node (vectors are cleared by kernels).
    void apply(const double* r, double* z) {
        if (!graphs_enabled() || c_.prof_matrix != nullptr) {
            apply_eager(r, z);
            return;
        }
        if (graph_r_ == r && graph_z_ == z && graph_mode_ != 0 && graph_epoch_ == layout_epoch()) {
            if (graph_mode_ == 2) FH_HIP(hipGraphLaunch(graph_exec_, c_.stream));
            else apply_eager(r, z);
            return;
        }
        if (graph_exec_) {
            FH_HIP(hipGraphExecDestroy(graph_exec_));
            graph_exec_ = nullptr;
        }
        graph_mode_ = 0;
        // eager reference time, after one warm-up application (first launches
        // load code objects); each application leaves the same result in z
        apply_eager(r, z);
        c_.sync();
        auto t0 = std::chrono::steady_clock::now();
        apply_eager(r, z);
        c_.sync();
        const double t_eager = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        hipGraph_t graph = nullptr;
        FH_HIP(hipStreamBeginCapture(c_.stream, hipStreamCaptureModeThreadLocal));
        try {
            apply_eager(r, z);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c_.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        FH_HIP(hipStreamEndCapture(c_.stream, &graph));
        hipError_t e = hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        FH_HIP(e);
        graph_r_ = r;
        graph_z_ = z;
        graph_epoch_ = layout_epoch();
        FH_HIP(hipGraphLaunch(graph_exec_, c_.stream));  // first replay (uploads the graph)
        c_.sync();
        t0 = std::chrono::steady_clock::now();
        FH_HIP(hipGraphLaunch(graph_exec_, c_.stream));
        c_.sync();
        const double t_graph = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        graph_mode_ = t_graph < t_eager ? 2 : 1;
        if (graph_mode_ == 1) {
            FH_HIP(hipGraphExecDestroy(graph_exec_));
            graph_exec_ = nullptr;
        }
        t_eager_ = t_eager;
        t_graph_ = t_graph;
    }
    // 0 undecided, 1 eager, 2 graph replay; timings of the decision (s)
    int graph_mode_ = 0;
    double t_eager_ = 0, t_graph_ = 0;

    void apply_eager(const double* r, double* z) {
        MgLevel& F = L[0];
        if (vp_ == Prec::F64) {
            cycle(0, r, F.x, true, nullptr);  // the cycle only reads its right-hand side
            copy(c_, F.n, Prec::F64, F.x.data(), z);
        } else {
            launch_convert(c_, F.n, r, Prec::F64, F.b.data(), vp_);
            cycle(0, F.b.data(), F.x, true, nullptr);
            launch_convert(c_, F.n, F.x.data(), vp_, z, Prec::F64);
        }
    }
    // Stand-alone cycle on FP64 vectors (x updated in place).
    void iterate(const double* b, double* x) {
        MgLevel& F = L[0];
        FH_CHECK(vp_ == Prec::F64, -2, "stand-alone multigrid runs in FP64");
        copy(c_, F.n, Prec::F64, x, F.x.data());
        cycle(0, b, F.x, false, nullptr);
        copy(c_, F.n, Prec::F64, F.x.data(), x);
    }

private:
    Ctx& c_;
    SolverArgs a_;
    Prec vp_;
    hipGraphExec_t graph_exec_ = nullptr;
    const double* graph_r_ = nullptr;
    double* graph_z_ = nullptr;
    // layout epoch at capture: a matrix whose sliced / chunked copy was
    // dropped or rebuilt since then makes the graph stale
    uint64_t graph_epoch_ = 0;

    // FHGPU_GRAPH=0 runs every cycle eagerly (for timing comparisons).
    static bool graphs_enabled() {
        static const bool on = [] {
            const char* e = std::getenv("FHGPU_GRAPH");
            return !(e && std::atoi(e) == 0);
        }();
        return on;
    }

    void build_amg() {
        // Blocks of hierarchical face bases: the first coarsening keeps the
        // lowest mode of every face (p-injection, LevelForHHO.h:586-620, as
        // HighOrderAggregAMG coarsens blocks to their first component,
        // HighOrderAggregAMG.h:28-45); scalar pairwise aggregation follows.
        if (L[0].A64->bs_r > 1 && L[0].A64->rows() > a_.max_coarse) {
            L[0].ptrunc = true;
            L[0].nfc = 1;
            MgLevel C;
            C.A64_own = bcsr_mode_submatrix(c_, *L[0].A64, 1);
            C.A64 = C.A64_own.get();
            L.push_back(std::move(C));
        }
        // AggregAMG with double pairwise aggregation (AggregAMG.h:32-139,
        // Multigrid.h:121-283): coarsen while the level has more than
        // max_coarse unknowns and the coarsening ratio stays above 1.01.
        while (L.back().A64->rows() > a_.max_coarse && L.size() < 40) {
            MgLevel& F = L.back();
            const Bcsr& A = *F.A64;
            std::vector<int> agg1, agg2;
            std::vector<std::vector<int>> mem1, mem2;
            bool coarsest = false;
            int n1 = pairwise_aggregate(A, a_.theta, agg1, mem1, coarsest);
            if (coarsest) break;
            auto A1 = galerkin_aggregate(c_, A, agg1, n1);
            int n2 = pairwise_aggregate(*A1, a_.theta, agg2, mem2, coarsest);
            if (coarsest) break;
            // final aggregates = members of members, in order (AggregAMG.h:142-154)
            std::vector<int> ptr(n2 + 1, 0), mem;
            mem.reserve(A.nbr);
            for (int J = 0; J < n2; J++) {
                for (int e : mem2[J])
                    for (int k : mem1[e]) mem.push_back(k);
                ptr[J + 1] = (int)mem.size();
            }
            std::vector<int> agg(A.nbr);
            for (int i = 0; i < A.nbr; i++) agg[i] = agg2[agg1[i]];
            auto Ac = galerkin_aggregate(c_, *A1, agg2, n2);
            F.agg = true;
            F.nagg = n2;
            F.agg_map.alloc(sizeof(int) * agg.size());
            F.agg_ptr.alloc(sizeof(int) * ptr.size());
            F.agg_members.alloc(sizeof(int) * mem.size());
            FH_HIP(hipMemcpyAsync(F.agg_map.p, agg.data(), sizeof(int) * agg.size(), hipMemcpyHostToDevice, c_.stream));
            FH_HIP(hipMemcpyAsync(F.agg_ptr.p, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice, c_.stream));
            FH_HIP(hipMemcpyAsync(F.agg_members.p, mem.data(), sizeof(int) * mem.size(), hipMemcpyHostToDevice, c_.stream));
            c_.sync();
            MgLevel C;
            C.A64_own = std::move(Ac);
            C.A64 = C.A64_own.get();
            L.push_back(std::move(C));
        }
    }

    void finish() {
        const Prec dprec = vp_ == Prec::F64 ? Prec::F64 : Prec::F32;
        // smoothing iterations of level l: the finest level's nu, changed
        // level by level by nu_change_op, or the given coarse count
        auto nu_of = [&](int nu, int coarse_nu, int l) { return l > 0 && coarse_nu >= 0 ? coarse_nu : a_.level_nu(nu, l); };
        double fine_nnz = (double)L[0].A64->nnzb;
        op_complexity = 0;
        for (size_t l = 0; l < L.size(); l++) {
            MgLevel& V = L[l];
            V.n = V.A64->rows();
            op_complexity += (double)V.A64->nnzb / fine_nnz;
            const Prec store = l > 0 ? a_.coarse_store : a_.mg_store;
            if (store != Prec::F64) {
                V.Aw_own = bcsr_convert(c_, *V.A64, store);
                V.Aw = V.Aw_own.get();
            } else
                V.Aw = V.A64;
            bool coarsest = l + 1 == L.size();
            V.b.resize(V.n, vp_);
            V.x.resize(V.n, vp_);
            V.r.resize(V.n, vp_);
            V.work.resize(V.n, vp_);
            if (a_.cycle == 'K' && l > 0 && !coarsest) {
                V.kc.resize(V.n, vp_); V.kv.resize(V.n, vp_); V.kd.resize(V.n, vp_); V.kw.resize(V.n, vp_);
            }
            if (!coarsest) {
                if (l > 0 && !a_.coarse_smoother.empty()) {
                    const double w = a_.coarse_relax > 0 ? a_.coarse_relax : a_.relax;
                    V.pre.spec = SmootherSpec::parse(a_.coarse_smoother, nu_of(a_.nu1, a_.coarse_nu1, (int)l), w);
                    V.post.spec = SmootherSpec::parse(a_.coarse_smoother, nu_of(a_.nu2, a_.coarse_nu2, (int)l), w);
                } else {
                    V.pre.spec = SmootherSpec::parse(a_.pre_smoother, nu_of(a_.nu1, a_.coarse_nu1, (int)l), a_.relax);
                    V.post.spec = SmootherSpec::parse(a_.post_smoother, nu_of(a_.nu2, a_.coarse_nu2, (int)l), a_.relax);
                }
                smoother_setup(c_, *V.A64, V.pre, dprec);
                if (V.post.spec.kind == V.pre.spec.kind && V.pre.spec.kind != SmootherKind::BlockGS) {
                    // the same block inverses and lambda_max estimate serve both
                    V.post.dprec = V.pre.dprec;
                    V.post.invD = V.pre.invD;
                    V.post.lmax = V.pre.lmax;
                    if (V.post.spec.kind == SmootherKind::Chebyshev) V.post.d.resize(V.A64->rows(), dprec);
                } else {
                    smoother_setup(c_, *V.A64, V.post, dprec);
                }
                if (!V.agg && !V.ptrunc) {
                    Prec tp = vp_ == Prec::F64 ? Prec::F64 : Prec::F32;
                    if (tp != Prec::F64) {
                        V.Pw_own = bcsr_convert(c_, *V.P64, tp);
                        V.Rw_own = bcsr_convert(c_, *V.R64_own, tp);
                        V.Pw = V.Pw_own.get();
                        V.Rw = V.Rw_own.get();
                    } else {
                        V.Pw = V.P64;
                        V.Rw = V.R64_own.get();
                    }
                }
            } else {
                std::vector<double> inv = dense_spd_inverse(*V.A64);
                coarse_n = (int)V.n;
                coarse_inv.alloc(sizeof(double) * inv.size());
                FH_HIP(hipMemcpyAsync(coarse_inv.p, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice,
                                      c_.stream));
                c_.sync();
            }
        }
    }

    void restrict_to(int l, const DVec& r, DVec& rc) {
        MgLevel& V = L[l];
        if (V.ptrunc)
            launch_p_restrict(c_, V.A64->nbr, V.A64->bs_r, V.nfc, vp_, r.data(), rc.data());
        else if (V.agg)
            launch_agg_restrict(c_, V.nagg, V.A64->bs_r, vp_, V.agg_ptr.as<int>(), V.agg_members.as<int>(), r.data(),
                                rc.data());
        else
            launch_bcsr(c_, *V.Rw, OP_SPMV, vp_, r.data(), nullptr, rc.data(), 1.0, 0.0, 0.0, nullptr, vp_, nullptr, 0);
    }
    void prolong_add(int l, const DVec& ec, DVec& x) {
        MgLevel& V = L[l];
        if (V.ptrunc)
            launch_p_prolong_add(c_, V.A64->nbr, V.A64->bs_r, V.nfc, vp_, ec.data(), x.data());
        else if (V.agg)
            launch_agg_prolong_add(c_, V.A64->nbr, V.A64->bs_r, vp_, V.agg_map.as<int>(), ec.data(), x.data());
        else
            launch_bcsr(c_, *V.Pw, OP_SPMV, vp_, ec.data(), nullptr, x.data(), 1.0, 1.0, 0.0, nullptr, vp_, nullptr, 0);
    }

public:
    // Multigrid::MultigridCycle (Multigrid.h:438-608).
    void cycle(int l, const void* b, DVec& x, bool x_is_zero, DVec* Ax) {
        MgLevel& V = L[l];
        if (l + 1 == (int)L.size()) {
            launch_dense_symv(c_, coarse_n, coarse_inv.p, Prec::F64, vp_, b, x.data());
            if (Ax) launch_bcsr(c_, *V.Aw, OP_SPMV, vp_, x.data(), nullptr, Ax->data(), 1.0, 0.0, 0.0, nullptr, vp_, nullptr, 0);
            return;
        }
        bool xz = x_is_zero;
        smoother_apply(c_, *V.Aw, V.pre, vp_, b, x, V.work, xz);
        MgLevel& C = L[l + 1];
        // A p-transfer to one mode per face needs only entry 0 of each block
        // row of the residual: one pass over entry 0 of every block
        if (xz || !V.ptrunc || V.nfc != 1 || !launch_resid_head(c_, *V.Aw, vp_, x.data(), b, C.b.data())) {
            if (xz) copy(c_, V.n, vp_, b, V.r.data());
            else launch_bcsr(c_, *V.Aw, OP_RESID, vp_, x.data(), b, V.r.data(), 1.0, 0.0, 0.0, nullptr, vp_, nullptr, 0);
            restrict_to(l, V.r, C.b);
        }
        bool coarse_is_last = l + 2 == (int)L.size();
        if (a_.cycle == 'K' && !coarse_is_last && (a_.kcycle_levels <= 0 || l + 1 <= a_.kcycle_levels)) {
            kcycle(l + 1, C.b, C.x);
        } else {
            int loops = a_.cycle == 'W' ? std::max(2, a_.wloops) : 1;
            bool ez = true;
            for (int w = 0; w < loops; w++) {
                cycle(l + 1, C.b.data(), C.x, ez, nullptr);
                ez = false;
                if (coarse_is_last) break;
            }
        }
        if (xz) { launch_fill(c_, V.n, x.data(), vp_, 0.0); xz = false; }
        prolong_add(l, C.x, x);
        smoother_apply(c_, *V.Aw, V.post, vp_, b, x, V.work, xz);
        if (Ax) launch_bcsr(c_, *V.Aw, OP_SPMV, vp_, x.data(), nullptr, Ax->data(), 1.0, 0.0, 0.0, nullptr, vp_, nullptr, 0);
    }

    // Multigrid::FCGForKCycle (Multigrid.h:611-656); r is overwritten.
    // Everything stays on the device: the second cycle, on the updated
    // residual, is always run, and the coefficient kernel drops it
    // (cb = 0) when |r_new| <= t |r|, as the reference's early return does.
    // With no host decision inside, the whole preconditioner can be
    // captured into a graph.
    void kcycle(int l, DVec& r, DVec& out) {
        MgLevel& V = L[l];
        const int s = S_KBASE + 8 * l;
        const double t = 0.25;
        cycle(l, r.data(), V.kc, true, &V.kv);
        // rho1 = c.v, alpha1 = c.r, |r|^2
        dot3(c_, V.n, vp_, V.kc.data(), V.kv.data(), s + 0, V.kc.data(), r.data(), s + 1, r.data(), r.data(), s + 2);
        // r <- r - (alpha1 / rho1) v, then |r_new|^2
        launch_axpby(c_, V.n, vp_, V.kv.data(), r.data(), -1.0, s + 1, s + 0, 1.0, -1, -1);
        dot1(c_, V.n, vp_, r.data(), r.data(), s + 3);
        cycle(l, r.data(), V.kd, true, &V.kw);
        // gamma = d.v, beta = d.w, alpha2 = d.r
        dot3(c_, V.n, vp_, V.kd.data(), V.kv.data(), s + 4, V.kd.data(), V.kw.data(), s + 5, V.kd.data(), r.data(), s + 6);
        launch_kcycle_coeffs(c_, s, t);
        launch_axpby(c_, V.n, vp_, V.kc.data(), out.data(), 1.0, s + 2, -1, 0.0, -1, -1);
        launch_axpby(c_, V.n, vp_, V.kd.data(), out.data(), 1.0, s + 3, -1, 1.0, -1, -1);
    }
};

// ------------------------------------------------------------------ Solver
Solver::Solver(Ctx& c, const SolverArgs& a) : c_(c), args_(a) {}
Solver::~Solver() = default;

void Solver::SetLevel(int l, const Bcsr* A, const Bcsr* P) {
    if ((int)given_levels_.size() <= l) given_levels_.resize(l + 1, {nullptr, nullptr});
    given_levels_[l] = {A, P};
}

void Solver::SetLevelP(int l, int nfc) {
    FH_CHECK(l >= 0 && nfc >= 1, -2, "p-transfer: level >= 0 and at least one kept mode");
    if ((int)given_p_.size() <= l) given_p_.resize(l + 1, 0);
    given_p_[l] = nfc;
    if ((int)given_levels_.size() <= l + 1) given_levels_.resize(l + 2, {nullptr, nullptr});
}

void Solver::SetLearnedSmoother(Mpnn* nn, const NnGraph* g, double omega, bool unit_step) {
    FH_CHECK((nn == nullptr) == (g == nullptr), -2, "network and graph go together");
    nn_ = nn;
    nng_ = g;
    nn_omega_ = omega;
    nn_unit_step_ = unit_step;
}

int Solver::NumLevels() const { return mg_ ? (int)mg_->L.size() : 1; }
void Solver::ExportLevels(const std::string& dir) {
    FH_CHECK(mg_ != nullptr, -2, "the solver has no multigrid hierarchy");
    const auto& L = mg_->L;
    for (size_t l = 0; l < L.size(); l++) {
        const MgLevel& V = L[l];
        const std::string pre = dir + "/level" + std::to_string(l) + "_";
        market_write_matrix(c_, *V.A64, pre + "A.dat");
        if (l + 1 == L.size()) continue;
        if (V.P64) {
            market_write_matrix(c_, *V.P64, pre + "P.dat");
            market_write_matrix(c_, *V.R64_own, pre + "R.dat");
            continue;
        }
        // transfers without a stored matrix: p-injection or aggregation
        const int nf = V.A64->bs_r, nb = V.A64->nbr;
        const int nfc = V.ptrunc ? V.nfc : nf;
        std::vector<int> rows, cols;
        if (V.ptrunc) {
            for (int i = 0; i < nb; i++)
                for (int a = 0; a < nfc; a++) { rows.push_back(i * nf + a); cols.push_back(i * nfc + a); }
        } else {
            std::vector<int> agg(nb);
            FH_HIP(hipMemcpy(agg.data(), V.agg_map.p, sizeof(int) * nb, hipMemcpyDeviceToHost));
            for (int i = 0; i < nb; i++)
                for (int a = 0; a < nf; a++) { rows.push_back(i * nf + a); cols.push_back(agg[i] * nf + a); }
        }
        const int nr = nb * nf, nc = (int)L[l + 1].A64->rows();
        std::vector<int> prp(nr + 1, 0), pci(rows.size()), rrp(nc + 1, 0), rci(rows.size());
        std::vector<double> one(rows.size(), 1.0);
        for (int r : rows) prp[r + 1]++;
        for (int c : cols) rrp[c + 1]++;
        for (int i = 0; i < nr; i++) prp[i + 1] += prp[i];
        for (int i = 0; i < nc; i++) rrp[i + 1] += rrp[i];
        std::vector<int> fp(prp.begin(), prp.end() - 1), fr(rrp.begin(), rrp.end() - 1);
        for (size_t e = 0; e < rows.size(); e++) {  // rows ascending, one entry per row: already sorted
            pci[fp[rows[e]]++] = cols[e];
            rci[fr[cols[e]]++] = rows[e];
        }
        market_write_csr(nr, nc, prp.data(), pci.data(), one.data(), pre + "P.dat");
        market_write_csr(nc, nr, rrp.data(), rci.data(), one.data(), pre + "R.dat");
    }
}

int Solver::LevelUnknowns(int l) const {
    if (!mg_) return A_ ? (int)A_->rows() : 0;
    return (int)mg_->L.at(l).A64->rows();
}

void Solver::Setup(const Bcsr& A) {
    auto t0 = std::chrono::steady_clock::now();
    FH_CHECK(A.bs_r == A.bs_c && A.nbr == A.nbc, -2, "the operator must be square with square blocks");
    A_ = &A;
    const std::string& p = args_.solver == "cg" || args_.solver == "fcg" ? args_.prec : args_.solver;
    mg_.reset();
    prec_sm_.reset();
    if (p == "uamg") {
        // levels from SetupUncondensed; coarse operators are the hierarchy's
        FH_CHECK(uamg_ && !given_levels_.empty() && given_levels_[0].first == &A, -2,
                 "the uamg multigrid needs the uncondensed blocks (SetupUncondensed)");
        SolverArgs a = args_;
        a.galerkin = false;
        mg_ = std::make_unique<Multigrid>(c_, a);
        mg_->setup(A, given_levels_, false);
    } else if (p == "mg" || p == "aggregamg") {
        mg_ = std::make_unique<Multigrid>(c_, args_);
        std::vector<std::pair<const Bcsr*, const Bcsr*>> g = given_levels_;
        if (p == "mg") {
            if (g.empty()) g.push_back({&A, nullptr});
            g[0].first = &A;
        }
        mg_->setup(A, g, p == "aggregamg", given_p_);
    } else if (p == "bj" || p == "bj23" || p == "bgs" || p == "rbgs" || p == "sbgs" || p == "bsor" || p == "rbsor" ||
               p == "sbsor") {
        // SolverFactory.h:57-84: the bgs codes are BlockSOR with omega = 1,
        // the bsor codes take the relaxation parameter
        prec_sm_ = std::make_unique<LevelSmoother>();
        const bool gs = p == "bgs" || p == "rbgs" || p == "sbgs";
        prec_sm_->spec = SmootherSpec::parse(p, 1, gs ? 1.0 : args_.relax);
        smoother_setup(c_, A, *prec_sm_, Prec::F64);
    } else if (!(p.empty() || p == "none" || p == "default"))
        throw Error(-2, "Unknown solver or preconditioner '" + p + "'");
    // Multigrid::CanOptimizeResidualComputation (Multigrid.h:74-77): the
    // reference obtains A z from a point Gauss-Seidel post-smoother.
    fused_ = mg_ && mg_->L.size() > 1 && A.bs_r == 1 && args_.nu2 > 0 &&
             SmootherSpec::parse(args_.post_smoother, 1, 1.0).kind == SmootherKind::BlockGS;
    const int64_t n = A.rows();
    if (nn_) {
        FH_CHECK(args_.solver == "fcg", -2, "the learned post-smoother runs inside FCG (flexible preconditioner)");
        FH_CHECK(nng_->A == &A, -2, "the network graph was built for another operator");
        nn_->reserve(nng_->n);
        r1_.resize(n, Prec::F64); delta_.resize(n, Prec::F64); Adelta_.resize(n, Prec::F64);
        nn_counters_.alloc(2 * sizeof(int));
    }
    r_.resize(n, Prec::F64); z_.resize(n, Prec::F64); d_.resize(n, Prec::F64); Ad_.resize(n, Prec::F64);
    Az_.resize(n, Prec::F64); dk_.resize(n, Prec::F64); Adk_.resize(n, Prec::F64); t_.resize(n, Prec::F64);
    c_.sync();
    stats_ = SolveStats();
    stats_.levels = NumLevels();
    stats_.op_complexity = mg_ ? mg_->op_complexity : 1.0;
    stats_.setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Scalar CSR of the stored entries of A (host values; the stored-entry mask
// when A was made from a scalar CSR, else every non-zero entry).
static Csr scalar_csr(const Bcsr& A) {
    FH_CHECK(!A.h_vals.empty() && A.bs_r == A.bs_c, -3, "the operator needs host values (keep_host)");
    const int bs = A.bs_r;
    const size_t bsz = (size_t)bs * bs;
    Csr S;
    S.nr = (int)A.rows();
    S.nc = (int)A.cols();
    S.rp.assign(S.nr + 1, 0);
    std::vector<int> order;
    for (int I = 0; I < A.nbr; I++) {
        order.assign(A.h_colidx.begin() + A.h_rowptr[I], A.h_colidx.begin() + A.h_rowptr[I + 1]);
        std::vector<int> pos(order.size());
        for (size_t t = 0; t < pos.size(); t++) pos[t] = A.h_rowptr[I] + (int)t;
        std::sort(pos.begin(), pos.end(), [&A](int x, int y) { return A.h_colidx[x] < A.h_colidx[y]; });
        for (int a = 0; a < bs; a++) {
            for (int p : pos)
                for (int b = 0; b < bs; b++) {
                    const size_t e = (size_t)p * bsz + (size_t)a * bs + b;
                    const bool stored = A.h_mask.empty() ? A.h_vals[e] != 0.0 : A.h_mask[e] != 0;
                    if (!stored) continue;
                    S.ci.push_back(A.h_colidx[p] * bs + b);
                    S.v.push_back(A.h_vals[e]);
                }
            S.rp[I * bs + a + 1] = (int)S.ci.size();
        }
    }
    return S;
}

void Solver::SetupUncondensed(const Bcsr& A, const Csr& ATT, const Csr& ATF, const Csr& AFF, int cell_bs) {
    const std::string& p = args_.solver == "cg" || args_.solver == "fcg" ? args_.prec : args_.solver;
    FH_CHECK(p == "uamg", -2, "uncondensed blocks are only used by the uamg multigrid");
    UamgArgs ua;
    ua.cell_bs = cell_bs;
    ua.face_bs = A.bs_r;
    ua.theta = args_.theta;
    ua.max_coarse = args_.max_coarse;
    uamg_ = std::make_unique<UamgHierarchy>(uamg_build(scalar_csr(A), ATT, ATF, AFF, ua));
    uamg_mats_.clear();
    given_levels_.clear();
    const int nl = (int)uamg_->A.size();
    std::vector<const Bcsr*> Al(nl, nullptr), Pl(nl, nullptr);
    Al[0] = &A;
    for (int l = 0; l < nl; l++) {
        if (l > 0) {
            const Csr& M = uamg_->A[l];
            uamg_mats_.push_back(bcsr_from_csr(c_, M.nr, M.nc, A.bs_r, M.rp.data(), M.ci.data(), M.v.data(), true));
            Al[l] = uamg_mats_.back().get();
        }
        if (l + 1 < nl) {
            const Csr& M = uamg_->P[l];
            uamg_mats_.push_back(bcsr_from_csr(c_, M.nr, M.nc, A.bs_r, M.rp.data(), M.ci.data(), M.v.data(), true));
            Pl[l] = uamg_mats_.back().get();
        }
    }
    for (int l = 0; l < nl; l++) SetLevel(l, Al[l], Pl[l]);
    Setup(A);
}

void Solver::SetStop(double tol, int max_iter) {
    FH_CHECK(tol > 0 && max_iter >= 0, -2, "tolerance must be positive and max_iter >= 0");
    args_.tol = tol;
    args_.max_iter = max_iter;
}

void Solver::ApplyPreconditioner(const double* r, double* z) {
    FH_CHECK(A_ != nullptr, -2, "Setup() must be called before ApplyPreconditioner()");
    const int64_t n = A_->rows();
    // through the solver's own work vectors, so that a captured cycle is reused
    copy(c_, n, Prec::F64, r, r_.data());
    Mpnn* keep = nn_;
    nn_ = nullptr;  // the plain preconditioner, without the learned post-smoother
    apply_prec(r_, z_, nullptr);
    nn_ = keep;
    copy(c_, n, Prec::F64, z_.data(), z);
    c_.sync();
}

double Solver::read_scalar(int slot) {
    read_scalars(c_, slot, 1);
    return c_.h_scalars[slot];
}

void Solver::apply_prec(const DVec& r, DVec& z, DVec* Az) {
    const int64_t n = A_->rows();
    if (mg_) mg_->apply((const double*)r.data(), (double*)z.data());
    else if (prec_sm_) {
        bool xz = true;
        smoother_apply(c_, *A_, *prec_sm_, Prec::F64, r.data(), z, t_, xz);
    } else
        copy(c_, n, Prec::F64, r.data(), z.data());
    if (nn_) {
        const Prec P = Prec::F64;
        // with outer_prec FP32, the network's two products use the FP32 copy
        // of the finest level when it has an element layout (FP64 vectors
        // and accumulation); FCG then replaces its residual as for its own
        // FP32-stored product
        const Bcsr* A32 = nullptr;
        if (args_.outer_prec == Prec::F32 && mg_ && !mg_->L.empty() && mg_->L[0].Aw &&
            mg_->L[0].Aw->prec == Prec::F32 && mg_->L[0].Aw->elem_on && elem_enabled())
            A32 = mg_->L[0].Aw;
        nn_f32_products_ = A32 != nullptr;
        launch_bcsr(c_, A32 ? *A32 : *A_, OP_RESID, P, z.data(), r.data(), r1_.data(), 1.0, 0.0, 0.0, nullptr, P,
                    nullptr, 0);
        nn_->forward(*nng_, (const double*)r1_.data(), (double*)delta_.data(), true);
        if (!(A32 && launch_spmv_f32vals(c_, *A32, delta_.data(), Adelta_.data())))
            launch_bcsr(c_, *A_, OP_SPMV, P, delta_.data(), nullptr, Adelta_.data(), 1.0, 0.0, 0.0, nullptr, P, nullptr,
                        0);
        const void* xs[2] = {delta_.data(), delta_.data()};
        const void* ys[2] = {r1_.data(), Adelta_.data()};
        int sl[2] = {kNnSlotS1, kNnSlotS2};
        launch_dots(c_, n, P, 2, xs, ys, sl);
        // an accepted correction also gives A z = (r - r1) + alpha A delta;
        // a block-Jacobi fallback (rare) needs the product
        nn_select(c_, n, A_->bs_r, kNnSlotS1, kNnSlotS2, (const double*)delta_.data(), (const double*)r1_.data(),
                  nng_->invD64.as<double>(), nn_omega_, nn_unit_step_, (double*)z.data(), nn_counters_.as<int>(),
                  Az ? (const double*)r.data() : nullptr, (const double*)Adelta_.data(),
                  Az ? (double*)Az->data() : nullptr);
        if (Az) {
            int cnt[2];
            FH_HIP(hipMemcpyAsync(cnt, nn_counters_.p, sizeof(cnt), hipMemcpyDeviceToHost, c_.stream));
            c_.sync();
            if (cnt[1] != nn_fallbacks_seen_) {
                nn_fallbacks_seen_ = cnt[1];
                launch_bcsr(c_, *A_, OP_SPMV, Prec::F64, z.data(), nullptr, Az->data(), 1.0, 0.0, 0.0, nullptr,
                            Prec::F64, nullptr, 0);
            }
        }
        return;
    }
    if (Az) launch_bcsr(c_, *A_, OP_SPMV, Prec::F64, z.data(), nullptr, Az->data(), 1.0, 0.0, 0.0, nullptr, Prec::F64, nullptr, 0);
}

void Solver::Solve(const double* b, double* x, bool x_is_zero) {
    FH_CHECK(A_ != nullptr, -2, "Setup() must be called before Solve()");
    c_.sync();
    auto t0 = std::chrono::steady_clock::now();
    hist_.clear();
    if (nn_) {
        launch_fill(c_, 2, nn_counters_.p, Prec::F32, 0.0);  // two int zeros (same bit pattern as 0.0f)
        nn_fallbacks_seen_ = 0;
    }
    if (args_.solver == "cg") solve_cg(b, x, x_is_zero);
    else if (args_.solver == "fcg") solve_fcg(b, x, x_is_zero);
    else solve_stationary(b, x, x_is_zero);
    c_.sync();
    stats_.graph_mode = mg_ ? mg_->graph_mode_ : 0;
    if (nn_) {
        int cnt[2];
        FH_HIP(hipMemcpy(cnt, nn_counters_.p, sizeof(cnt), hipMemcpyDeviceToHost));
        stats_.nn_used = cnt[0];
        stats_.nn_fallback = cnt[1];
    }
    stats_.solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// IterativeSolver::StoppingCriteriaReached (IterativeSolver.h:224-247)
#define FH_STOP_TEST()                                                                        \
    [&]() -> bool {                                                                           \
        if (args_.max_iter == 0) return true;                                                 \
        if (it == 0) return false;                                                            \
        if (it >= args_.max_iter) {                                                           \
            stats_.converged = res < args_.tol ? 1 : 0;                                       \
            return true;                                                                      \
        }                                                                                     \
        FH_CHECK(!std::isinf(res), -7, "divergence of the solver");                           \
        FH_CHECK(!std::isnan(res), -8, "the residual is NaN.");                               \
        if (res < args_.tol) { stats_.converged = 1; return true; }                           \
        return false;                                                                         \
    }

// The first residual of the Krylov loops (ConjugateGradient.h:36-45,
// FlexibleConjugateGradient.h:58-66): r = b from x = 0, else b - A x, and
// the normalised norm of the first iteration result (IterationResult.h:121-139).
double Solver::initial_residual(const double* b, double* x, bool xz, double& res) {
    const int64_t n = A_->rows();
    const Prec P = Prec::F64;
    if (xz) {
        launch_fill(c_, n, x, P, 0.0);
        copy(c_, n, P, b, r_.data());
    } else {
        launch_bcsr(c_, *A_, OP_RESID, P, x, b, r_.data(), 1.0, 0.0, 0.0, nullptr, P, nullptr, 0);
    }
    dot2(c_, n, P, b, b, S_BB, r_.data(), r_.data(), S_RR);
    const double bn = std::sqrt(read_scalar(S_BB));
    const double rn = std::sqrt(read_scalar(S_RR));
    res = xz ? 1.0 : (bn > 0 ? rn / bn : rn);
    return bn;
}

void Solver::solve_cg(const double* b, double* x, bool xz) {
    // ConjugateGradient::Solve (ConjugateGradient.h:27-104)
    const int64_t n = A_->rows();
    const Prec P = Prec::F64;
    double res = 1.0;
    const double bn = initial_residual(b, x, xz, res);
    apply_prec(r_, z_, nullptr);
    copy(c_, n, P, z_.data(), d_.data());
    dot1(c_, n, P, r_.data(), z_.data(), S_RZ);
    int it = 0;
    hist_.push_back(res);
    stats_.converged = 0;
    auto stop = FH_STOP_TEST();
    while (!stop()) {
        if (it > 0) launch_axpby(c_, n, P, z_.data(), d_.data(), 1.0, -1, -1, 1.0, S_RZ, S_RZOLD);  // d = z + beta d
        launch_bcsr(c_, *A_, OP_SPMV, P, d_.data(), nullptr, Ad_.data(), 1.0, 0.0, 0.0, nullptr, P, nullptr, 0);
        dot1(c_, n, P, d_.data(), Ad_.data(), S_DAD);
        launch_cg_update(c_, n, P, x, d_.data(), r_.data(), Ad_.data(), S_RZ, S_DAD, 1.0);
        launch_scalar_op(c_, 4, S_RZOLD, S_RZ, 0);
        apply_prec(r_, z_, nullptr);
        dot2(c_, n, P, r_.data(), z_.data(), S_RZ, r_.data(), r_.data(), S_RR);
        it++;
        double rr = read_scalar(S_RR);
        res = bn > 0 ? std::sqrt(rr) / bn : std::sqrt(rr);
        hist_.push_back(res);
    }
    stats_.iterations = it;
    stats_.final_res = res;
}

void Solver::solve_fcg(const double* b, double* x, bool xz) {
    // FlexibleConjugateGradient::Solve (FlexibleConjugateGradient.h:47-171).
    // z is A-orthogonalised against the kept previous directions,
    // oldest first; when `Truncation` directions are kept the oldest is
    // dropped (or all are, with Restart).  Directions live in a ring of
    // Truncation + 1 buffers, so no vector is copied.
    const int64_t n = A_->rows();
    const Prec P = Prec::F64;
    const int m = std::max(1, args_.fcg_truncation);
    if ((int)fd_.size() != m + 1) {
        fd_.clear();
        fAd_.clear();
        fd_.resize(m + 1);
        fAd_.resize(m + 1);
    }
    for (int k = 0; k <= m; k++) {
        fd_[k].resize(n, P);
        fAd_[k].resize(n, P);
    }
    auto slot_dad = [](int k) { return S_FCG + 2 * k; };
    auto slot_zad = [](int k) { return S_FCG + 2 * k + 1; };
    // FCG's own product with the FP32 copy of the finest multigrid level
    const Bcsr* A32 = nullptr;
    if (args_.outer_prec == Prec::F32 && mg_ && !mg_->L.empty() && mg_->L[0].Aw && mg_->L[0].Aw->prec == Prec::F32)
        A32 = mg_->L[0].Aw;
    double res = 1.0;
    const double bn = initial_residual(b, x, xz, res);
    int it = 0;
    hist_.push_back(res);
    stats_.converged = 0;
    std::vector<int> kept;  // ring indices, oldest first
    int since_replace = 0;  // FP32-stored products since the last residual replacement
    auto stop = FH_STOP_TEST();
    while (!stop()) {
        int cur = 0;
        while (std::find(kept.begin(), kept.end(), cur) != kept.end()) cur++;
        DVec& d = fd_[cur];
        DVec& Ad = fAd_[cur];
        const bool haz = fused_ || nn_ != nullptr;  // the preconditioner returns A z
        apply_prec(r_, z_, haz ? &Az_ : nullptr);
        // z . Adk for every kept direction (three per launch)
        for (size_t q = 0; q < kept.size(); q += 3) {
            const void* xs[3];
            const void* ys[3];
            int sl[3];
            const int cnt = (int)std::min<size_t>(3, kept.size() - q);
            for (int t = 0; t < cnt; t++) {
                xs[t] = z_.data();
                ys[t] = fAd_[kept[q + t]].data();
                sl[t] = slot_zad(kept[q + t]);
            }
            launch_dots(c_, n, P, cnt, xs, ys, sl);
        }
        // d = z - sum_k (z.Adk / dk.Adk) dk over the kept directions, oldest
        // first, in one pass (and A d the same way when the preconditioner
        // returned A z)
        {
            std::vector<const double*> vs, avs;
            std::vector<int> num, den;
            for (int k : kept) {
                vs.push_back((const double*)fd_[k].data());
                avs.push_back((const double*)fAd_[k].data());
                num.push_back(slot_zad(k));
                den.push_back(slot_dad(k));
            }
            launch_fcg_orth(c_, n, (const double*)z_.data(), (int)kept.size(), vs.data(), num.data(), den.data(),
                            (double*)d.data());
            if (haz)
                launch_fcg_orth(c_, n, (const double*)Az_.data(), (int)kept.size(), avs.data(), num.data(), den.data(),
                                (double*)Ad.data());
        }
        bool approx = nn_ != nullptr && nn_f32_products_;  // A d formed with FP32-stored values
        if (!haz) {
            approx = A32 && launch_spmv_f32vals(c_, *A32, d.data(), Ad.data());
            if (!approx)
                launch_bcsr(c_, *A_, OP_SPMV, P, d.data(), nullptr, Ad.data(), 1.0, 0.0, 0.0, nullptr, P, nullptr, 0);
        }
        since_replace = approx ? since_replace + 1 : since_replace;
        // d.Ad and d.r (= r.d) in one pass over d
        dot2(c_, n, P, d.data(), Ad.data(), S_DAD, d.data(), r_.data(), S_RD);
        // x += a d, r -= a Ad and |r|^2 in one pass
        launch_cg_update_rr(c_, n, x, (const double*)d.data(), (double*)r_.data(), (const double*)Ad.data(), S_RD,
                            S_DAD, 1.0, S_RR);
        if ((int)kept.size() == m) {
            if (args_.fcg_restart) kept.clear();
            else kept.erase(kept.begin());
        }
        launch_scalar_op(c_, 4, slot_dad(cur), S_DAD, 0);
        kept.push_back(cur);
        it++;
        double rr = read_scalar(S_RR);
        res = bn > 0 ? std::sqrt(rr) / bn : std::sqrt(rr);
        if (since_replace > 0 && (res < args_.tol || since_replace >= args_.replace_every)) {
            // residual replacement: the recursive residual carries the
            // rounding of the FP32-stored products; the stopping test and the
            // next directions use the FP64 b - A x
            launch_bcsr(c_, *A_, OP_RESID, P, x, b, r_.data(), 1.0, 0.0, 0.0, nullptr, P, nullptr, 0);
            dot1(c_, n, P, r_.data(), r_.data(), S_RR);
            rr = read_scalar(S_RR);
            res = bn > 0 ? std::sqrt(rr) / bn : std::sqrt(rr);
            since_replace = 0;
        }
        hist_.push_back(res);
    }
    stats_.iterations = it;
    stats_.final_res = res;
}

void Solver::solve_stationary(const double* b, double* x, bool xz) {
    // IterativeSolver::Solve (IterativeSolver.h:82-191) of a multigrid or a
    // smoother used as the solver.
    const int64_t n = A_->rows();
    const Prec P = Prec::F64;
    double res = 1.0;
    const double bn = initial_residual(b, x, xz, res);
    int it = 0;
    hist_.push_back(res);
    stats_.converged = 0;
    auto stop = FH_STOP_TEST();
    while (!stop()) {
        if (mg_) mg_->iterate(b, x);
        else {
            copy(c_, n, P, x, z_.data());
            smoother_apply(c_, *A_, *prec_sm_, P, b, z_, t_, xz);
            copy(c_, n, P, z_.data(), x);
        }
        launch_bcsr(c_, *A_, OP_RESID, P, x, b, r_.data(), 1.0, 0.0, 0.0, nullptr, P, nullptr, 0);
        dot1(c_, n, P, r_.data(), r_.data(), S_RR);
        it++;
        double rr = read_scalar(S_RR);
        res = bn > 0 ? std::sqrt(rr) / bn : std::sqrt(rr);
        hist_.push_back(res);
    }
    stats_.iterations = it;
    stats_.final_res = res;
}

}  // namespace fh
