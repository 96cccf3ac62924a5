// extern "C" boundary of the MI355X solve path (declared in include/fhgpu.h).
#include <cstring>
#include <string>

#include "../../include/fhgpu.h"
#include "fh_io.h"
#include "fh_kernels.h"
#include "fh_matrix.h"
#include "fh_nn.h"
#include "fh_solver.h"
#include "fh_hho.h"
#include "fh_condense.h"
#include "fh_uamg.h"

struct fh_ctx {
    fh::Ctx c;
    explicit fh_ctx(int dev) : c(dev) {}
};
struct fh_mat {
    std::unique_ptr<fh::Bcsr> A;
};
struct fh_bjac {
    fh::LevelSmoother S;
};
struct fh_coloring {
    int ncolours = 0;
    std::vector<int> colour;
    fh::LevelSmoother S;  // colour lists on the device
};
struct fh_solver {
    fh_ctx* ctx;
    std::unique_ptr<fh::Solver> s;
};
struct fh_nn {
    std::unique_ptr<fh::Mpnn> m;
};
struct fh_nngraph {
    std::unique_ptr<fh::NnGraph> g;
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const fh::Error& e) {
        g_err = e.what();
        return e.code < 0 ? e.code : -1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

fh::Prec P(int p) {
    if (p < 0 || p > 2) throw fh::Error(-2, "invalid precision code " + std::to_string(p));
    return (fh::Prec)p;
}
fh::Csr csr_in(const fh_csr* m, const char* name) {
    FH_CHECK(m && m->rowptr && m->nrows >= 0 && m->ncols >= 0, -2, std::string("missing matrix ") + name);
    FH_CHECK(m->rowptr[m->nrows] == 0 || (m->colidx && m->vals), -2, std::string("null arrays in ") + name);
    return fh::csr_copy(m->nrows, m->ncols, m->rowptr, m->colidx, m->vals);
}
void csr_out(const fh::Csr& M, int* nrows, int* ncols, int64_t* nnz, int* rowptr, int* colidx, double* vals) {
    if (nrows) *nrows = M.nr;
    if (ncols) *ncols = M.nc;
    if (nnz) *nnz = M.nnz();
    if (rowptr) std::memcpy(rowptr, M.rp.data(), sizeof(int) * M.rp.size());
    if (colidx) std::memcpy(colidx, M.ci.data(), sizeof(int) * M.ci.size());
    if (vals) std::memcpy(vals, M.v.data(), sizeof(double) * M.v.size());
}
#define NEED(ptr) FH_CHECK((ptr) != nullptr, -2, #ptr " is null")
#define NEED_DEV(ctx) FH_CHECK(!(ctx)->c.host_only, -9, "this entry point needs a device context")
}  // namespace

extern "C" {

const char* fh_last_error(void) { return g_err.c_str(); }

int fh_ctx_create(int device, fh_ctx** out) {
    return guard([&] {
        NEED(out);
        *out = new fh_ctx(device);
    });
}
int fh_ctx_destroy(fh_ctx* ctx) {
    return guard([&] { delete ctx; });
}
int fh_ctx_sync(fh_ctx* ctx) {
    return guard([&] { NEED(ctx); ctx->c.sync(); });
}

int fh_mat_from_csr(fh_ctx* ctx, int nrows, int ncols, int bs, const int* rowptr, const int* colidx,
                    const double* vals, int keep_host, fh_mat** out) {
    return guard([&] {
        NEED(ctx); NEED(rowptr); NEED(out);
        FH_CHECK(rowptr[nrows] == 0 || (colidx && vals), -2, "null column or value array");
        auto m = new fh_mat;
        m->A = fh::bcsr_from_csr(ctx->c, nrows, ncols, bs, rowptr, colidx, vals, keep_host != 0);
        *out = m;
    });
}
int fh_mat_from_blocks(fh_ctx* ctx, int nbr, int nbc, int bs, const int* rowptr, const int* colidx,
                       const double* vals, int keep_host, fh_mat** out) {
    return guard([&] {
        NEED(ctx); NEED(rowptr); NEED(out);
        FH_CHECK(rowptr[nbr] == 0 || (colidx && vals), -2, "null column or value array");
        for (int p = 0; p < rowptr[nbr]; p++)
            FH_CHECK(colidx[p] >= 0 && colidx[p] < nbc, -2, "block column index out of range");
        auto m = new fh_mat;
        m->A = fh::bcsr_from_blocks(ctx->c, nbr, nbc, bs, bs, rowptr, colidx, vals, keep_host != 0);
        *out = m;
    });
}
int fh_mat_convert(fh_ctx* ctx, const fh_mat* A, int prec, fh_mat** out) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(A); NEED(out);
        auto m = new fh_mat;
        m->A = fh::bcsr_convert(ctx->c, *A->A, P(prec));
        *out = m;
    });
}
int fh_mat_destroy(fh_mat* A) {
    return guard([&] { delete A; });
}
int fh_mat_info(const fh_mat* A, int* nbr, int* nbc, int* bs, int64_t* nnzb, int* prec) {
    return guard([&] {
        NEED(A);
        if (nbr) *nbr = A->A->nbr;
        if (nbc) *nbc = A->A->nbc;
        if (bs) *bs = A->A->bs_r;
        if (nnzb) *nnzb = A->A->nnzb;
        if (prec) *prec = (int)A->A->prec;
    });
}
int fh_mat_blocks(const fh_mat* A, int* rowptr, int* colidx, double* vals) {
    return guard([&] {
        NEED(A);
        const fh::Bcsr& M = *A->A;
        if (rowptr) std::memcpy(rowptr, M.h_rowptr.data(), sizeof(int) * M.h_rowptr.size());
        if (colidx) std::memcpy(colidx, M.h_colidx.data(), sizeof(int) * M.h_colidx.size());
        if (vals) {
            if (!M.h_vals.empty())
                std::memcpy(vals, M.h_vals.data(), sizeof(double) * M.h_vals.size());
            else {
                // device-only FP64 values (GPU assembly): copied from HBM after
                // the device's pending work
                FH_CHECK(M.prec == fh::Prec::F64 && M.vals.p, -3, "matrix without FP64 values");
                FH_HIP(hipDeviceSynchronize());
                FH_HIP(hipMemcpy(vals, M.vals.p, M.value_bytes(), hipMemcpyDeviceToHost));
            }
        }
    });
}
int fh_mat_device_arrays(const fh_mat* A, const int** rowptr, const int** colidx, const void** vals) {
    return guard([&] {
        NEED(A);
        if (rowptr) *rowptr = A->A->rowptr.as<int>();
        if (colidx) *colidx = A->A->colidx.as<int>();
        if (vals) *vals = A->A->vals.p;
    });
}

int fh_mat_set_sliced(fh_ctx* ctx, const fh_mat* A, int on, int* active) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(A);
        const bool r = fh::set_sliced(ctx->c, *A->A, on != 0);
        if (active) *active = r ? 1 : 0;
    });
}

int fh_mat_set_elem(fh_ctx* ctx, const fh_mat* A, int on, int* active) {
    return guard([&] {
        NEED(ctx); NEED(A);
        const bool r = fh::set_elem(*A->A, on != 0) && fh::elem_enabled();
        if (active) *active = r ? 1 : 0;
    });
}

int fh_mat_layout(const fh_mat* A, int* layout) {
    return guard([&] {
        NEED(A); NEED(layout);
        const fh::Bcsr& M = *A->A;
        int v = 0;
        if (M.elem_on && fh::elem_enabled())
            v = 5;
        else if (M.sliced_state == 1 && M.sliced)
            v = M.sliced->kind == 2 ? 4 : (M.sliced->kind == 1 ? 3 : (M.sliced->next ? 2 : 1));
        *layout = v;
    });
}

int fh_mat_export_market(fh_ctx* ctx, const fh_mat* A, const char* path) {
    return guard([&] {
        NEED(ctx); NEED(A); NEED(path);
        fh::market_write_matrix(ctx->c, *A->A, path);
    });
}
int fh_mat_import_market(fh_ctx* ctx, const char* path, int bs, int keep_host, fh_mat** out) {
    return guard([&] {
        NEED(ctx); NEED(path); NEED(out);
        int n = 0, m = 0;
        std::vector<int> rp, ci;
        std::vector<double> v;
        fh::market_read_csr(path, n, m, rp, ci, v);
        auto mat = new fh_mat;
        try {
            mat->A = fh::bcsr_from_csr(ctx->c, n, m, bs, rp.data(), ci.data(), v.data(), keep_host != 0);
        } catch (...) {
            delete mat;
            throw;
        }
        *out = mat;
    });
}
int fh_vec_export_market(int64_t n, const double* v, const char* path) {
    return guard([&] {
        NEED(v); NEED(path);
        FH_CHECK(n > 0, -2, "vector length must be positive");
        fh::market_write_vector(n, v, path);
    });
}
int fh_vec_import_market(const char* path, int64_t cap, double* v, int64_t* n) {
    return guard([&] {
        NEED(path); NEED(n);
        std::vector<double> x = fh::market_read_vector(path);
        *n = (int64_t)x.size();
        if (v) {
            FH_CHECK(cap >= (int64_t)x.size(), -2, "output buffer too small");
            std::memcpy(v, x.data(), sizeof(double) * x.size());
        }
    });
}

int fh_spmv(fh_ctx* ctx, const fh_mat* A, double alpha, const void* x, double beta, void* y, int vprec) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(A); NEED(x); NEED(y);
        fh::launch_bcsr(ctx->c, *A->A, fh::OP_SPMV, P(vprec), x, nullptr, y, alpha, beta, 0.0, nullptr, P(vprec),
                        nullptr, 0);
    });
}
int fh_residual(fh_ctx* ctx, const fh_mat* A, const void* b, const void* x, void* r, int vprec) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(A); NEED(b); NEED(x); NEED(r);
        FH_CHECK(A->A->nbr == A->A->nbc, -2, "residual needs a square operator");
        fh::launch_bcsr(ctx->c, *A->A, fh::OP_RESID, P(vprec), x, b, r, 1.0, 0.0, 0.0, nullptr, P(vprec), nullptr, 0);
    });
}
int fh_dot(fh_ctx* ctx, int64_t n, const void* x, const void* y, int vprec, double* out) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(x); NEED(y); NEED(out);
        const void* xs[1] = {x};
        const void* ys[1] = {y};
        int slot[1] = {fh::Ctx::kScalars - 1};
        fh::launch_dots(ctx->c, n, P(vprec), 1, xs, ys, slot);
        FH_HIP(hipMemcpyAsync(out, ctx->c.d_scalars + slot[0], sizeof(double), hipMemcpyDeviceToHost, ctx->c.stream));
        ctx->c.sync();
    });
}
int fh_axpby(fh_ctx* ctx, int64_t n, double a, const void* x, double b, void* y, int vprec) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(x); NEED(y);
        fh::launch_axpby(ctx->c, n, P(vprec), x, y, a, -1, -1, b, -1, -1);
    });
}
int fh_convert(fh_ctx* ctx, int64_t n, const void* src, int sprec, void* dst, int dprec) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(src); NEED(dst);
        fh::launch_convert(ctx->c, n, src, P(sprec), dst, P(dprec));
    });
}

int fh_profile_begin(fh_ctx* ctx, const fh_mat* A, int mode) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(A);
        FH_CHECK(mode >= 0 && mode <= 4, -2, "mode must be 0, 1, 2, 3 or 4");
        ctx->c.prof_matrix = A->A.get();
        ctx->c.prof_mode = mode;
        ctx->c.prof_events.clear();
    });
}
int fh_profile_end(fh_ctx* ctx, int* launches, double* mean_ms) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx);
        ctx->c.sync();
        double tot = 0;
        for (auto& e : ctx->c.prof_events) {
            float ms = 0;
            FH_HIP(hipEventElapsedTime(&ms, e.first, e.second));
            tot += ms;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        int n = (int)ctx->c.prof_events.size();
        if (launches) *launches = n;
        if (mean_ms) *mean_ms = n ? tot / n : 0.0;
        ctx->c.prof_events.clear();
        ctx->c.prof_matrix = nullptr;
        ctx->c.prof_mode = -1;
    });
}

int fh_bjac_create(fh_ctx* ctx, const fh_mat* A, double omega, int dprec, fh_bjac** out) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(A); NEED(out);
        FH_CHECK(A->A->prec == fh::Prec::F64, -2, "block inverses are computed from an FP64 matrix");
        FH_CHECK(omega > 0 && omega < 2, -2, "omega must lie in (0, 2)");
        auto b = new fh_bjac;
        b->S.spec.kind = fh::SmootherKind::BlockJacobi;
        b->S.spec.omega = omega;
        try {
            fh::smoother_setup(ctx->c, *A->A, b->S, P(dprec));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}
int fh_bjac_destroy(fh_bjac* bj) {
    return guard([&] { delete bj; });
}
int fh_bjac_inverses(fh_ctx* ctx, const fh_bjac* bj, double* host_out) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(bj); NEED(host_out);
        FH_CHECK(bj->S.dprec == fh::Prec::F64, -2, "download available for FP64 inverses");
        FH_HIP(hipMemcpyAsync(host_out, bj->S.invD->p, bj->S.invD->bytes, hipMemcpyDeviceToHost, ctx->c.stream));
        ctx->c.sync();
    });
}
int fh_bjac_sweep(fh_ctx* ctx, const fh_bjac* bj, const fh_mat* A, const void* b, void* x, void* work, int nsweeps,
                  int x_is_zero, int vprec) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(bj); NEED(A); NEED(b); NEED(x); NEED(work);
        const fh::Bcsr& M = *A->A;
        fh::Prec vp = P(vprec);
        const int nf = M.bs_r;
        void* cur = x;
        void* nxt = work;
        bool xz = x_is_zero != 0;
        for (int s = 0; s < nsweeps; s++) {
            if (xz)
                fh::launch_bjac_zero(ctx->c, M.nbr, nf, vp, b, cur, bj->S.spec.omega, bj->S.invD->p, bj->S.dprec,
                                     nullptr, 0);
            else {
                fh::launch_bcsr(ctx->c, M, fh::OP_SMOOTH, vp, cur, b, nxt, 1.0, 0.0, bj->S.spec.omega, bj->S.invD->p,
                                bj->S.dprec, nullptr, 0);
                std::swap(cur, nxt);
            }
            xz = false;
        }
        if (cur != x)
            FH_HIP(hipMemcpyAsync(x, cur, (size_t)M.rows() * fh::prec_bytes(vp), hipMemcpyDeviceToDevice,
                                  ctx->c.stream));
    });
}
int fh_bjac_sweep_noalloc(fh_ctx* ctx, const fh_bjac* bj, const fh_mat* A, const void* b, void* x, int nsweeps,
                          int x_is_zero, int vprec) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(bj); NEED(A); NEED(b); NEED(x);
        const fh::Bcsr& M = *A->A;
        fh::Prec vp = P(vprec);
        FH_CHECK(M.nbr == M.nbc && M.bs_r == M.bs_c, -2, "block Jacobi needs a square operator");
        const size_t need = (size_t)M.rows() * fh::prec_bytes(vp);
        FH_CHECK(need <= fh::kNoAllocScratchDoubles * sizeof(double), -5,
                 "fh_bjac_sweep_noalloc: " + std::to_string(M.rows()) + " unknowns exceed the scratch of " +
                     std::to_string(fh::kNoAllocScratchDoubles * sizeof(double) / fh::prec_bytes(vp)) +
                     " entries; use fh_bjac_sweep with a work vector");
        void* work = fh::noalloc_scratch(ctx->c.device);
        struct Flag {
            fh::Ctx& c;
            explicit Flag(fh::Ctx& c_) : c(c_) { c.no_alloc = true; }
            ~Flag() { c.no_alloc = false; }
        } flag(ctx->c);
        const int nf = M.bs_r;
        void* cur = x;
        void* nxt = work;
        bool xz = x_is_zero != 0;
        for (int s2 = 0; s2 < nsweeps; s2++) {
            if (xz)
                fh::launch_bjac_zero(ctx->c, M.nbr, nf, vp, b, cur, bj->S.spec.omega, bj->S.invD->p, bj->S.dprec,
                                     nullptr, 0);
            else {
                fh::launch_bcsr(ctx->c, M, fh::OP_SMOOTH, vp, cur, b, nxt, 1.0, 0.0, bj->S.spec.omega, bj->S.invD->p,
                                bj->S.dprec, nullptr, 0);
                std::swap(cur, nxt);
            }
            xz = false;
        }
        if (cur != x)
            FH_HIP(hipMemcpyAsync(x, cur, (size_t)M.rows() * fh::prec_bytes(vp), hipMemcpyDeviceToDevice,
                                  ctx->c.stream));
    });
}
int fh_coloring_create(fh_ctx* ctx, const fh_mat* A, fh_coloring** out) {
    return guard([&] {
        NEED(ctx); NEED(A); NEED(out);
        auto col = new fh_coloring;
        col->ncolours = fh::greedy_colouring(*A->A, col->colour);
        auto& S = col->S;
        S.ncolours = col->ncolours;
        S.colour_ptr.assign(S.ncolours + 1, 0);
        for (int cc : col->colour) S.colour_ptr[cc + 1]++;
        for (int k = 0; k < S.ncolours; k++) S.colour_ptr[k + 1] += S.colour_ptr[k];
        std::vector<int> rows(A->A->nbr), pos(S.colour_ptr.begin(), S.colour_ptr.end() - 1);
        for (int i = 0; i < A->A->nbr; i++) rows[pos[col->colour[i]]++] = i;
        if (!ctx->c.host_only) {
            S.colour_rows.alloc(sizeof(int) * rows.size());
            FH_HIP(hipMemcpyAsync(S.colour_rows.p, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice,
                                  ctx->c.stream));
            ctx->c.sync();
        }
        *out = col;
    });
}
int fh_coloring_destroy(fh_coloring* col) {
    return guard([&] { delete col; });
}
int fh_coloring_get(const fh_coloring* col, int* ncolours, int* colour_of_row) {
    return guard([&] {
        NEED(col);
        if (ncolours) *ncolours = col->ncolours;
        if (colour_of_row) std::memcpy(colour_of_row, col->colour.data(), sizeof(int) * col->colour.size());
    });
}
int fh_mcgs_sweep(fh_ctx* ctx, const fh_mat* A, const fh_bjac* bj, const fh_coloring* col, int dir, const void* b,
                  void* x, int vprec) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(A); NEED(bj); NEED(col); NEED(b); NEED(x);
        FH_CHECK(dir >= 0 && dir <= 2, -2, "dir must be 0, 1 or 2");
        const auto& S = col->S;
        auto sweep = [&](bool backward) {
            for (int k = 0; k < S.ncolours; k++) {
                int cc = backward ? S.ncolours - 1 - k : k;
                int cnt = S.colour_ptr[cc + 1] - S.colour_ptr[cc];
                fh::launch_bcsr(ctx->c, *A->A, fh::OP_SMOOTH, P(vprec), x, b, x, 1.0, 0.0, 1.0, bj->S.invD->p,
                                bj->S.dprec, S.colour_rows.as<int>() + S.colour_ptr[cc], cnt);
            }
        };
        if (dir == 0 || dir == 2) sweep(false);
        if (dir == 1 || dir == 2) sweep(true);
    });
}

int fh_aggregate_pairwise(fh_ctx* ctx, const fh_mat* A, double theta, int* agg_out, int* nagg, int* coarsest) {
    return guard([&] {
        NEED(ctx); NEED(A); NEED(agg_out);
        std::vector<int> agg;
        std::vector<std::vector<int>> mem;
        bool c = false;
        int na = fh::pairwise_aggregate(*A->A, theta, agg, mem, c);
        std::memcpy(agg_out, agg.data(), sizeof(int) * agg.size());
        if (nagg) *nagg = na;
        if (coarsest) *coarsest = c ? 1 : 0;
    });
}
int fh_galerkin_aggregate(fh_ctx* ctx, const fh_mat* A, const int* agg, int nagg, fh_mat** out) {
    return guard([&] {
        NEED(ctx); NEED(A); NEED(agg); NEED(out);
        std::vector<int> a(agg, agg + A->A->nbr);
        for (int v : a) FH_CHECK(v >= 0 && v < nagg, -2, "aggregate index out of range");
        auto m = new fh_mat;
        m->A = fh::galerkin_aggregate(ctx->c, *A->A, a, nagg);
        *out = m;
    });
}

int fh_galerkin_rap(fh_ctx* ctx, const fh_mat* A, const fh_mat* P, fh_mat** out) {
    return guard([&] {
        NEED(ctx); NEED(A); NEED(P); NEED(out);
        auto m = new fh_mat;
        m->A = fh::galerkin_rap(ctx->c, *A->A, *P->A);
        *out = m;
    });
}

int fh_solver_create(fh_ctx* ctx, const fh_solver_args* a, fh_solver** out) {
    return guard([&] {
        NEED(ctx); NEED(a); NEED(out);
        fh::SolverArgs s;
        if (a->solver) s.solver = a->solver;
        if (a->prec) s.prec = a->prec;
        s.cycle = a->cycle ? a->cycle : 'V';
        FH_CHECK(s.cycle == 'V' || s.cycle == 'W' || s.cycle == 'K', -2, "cycle must be V, W or K");
        s.wloops = a->wloops > 0 ? a->wloops : (s.cycle == 'W' ? 2 : 1);
        s.nu1 = a->nu1;
        s.nu2 = a->nu2;
        if (a->pre_smoother) s.pre_smoother = a->pre_smoother;
        if (a->post_smoother) s.post_smoother = a->post_smoother;
        s.relax = a->relax > 0 ? a->relax : 1.0;
        s.tol = a->tol > 0 ? a->tol : 1e-8;
        s.max_iter = a->max_iter >= 0 ? a->max_iter : 200;  // -1: the reference's default (ProgramArguments.h:95)
        s.max_coarse = a->max_coarse > 0 ? a->max_coarse : 1000;
        s.theta = a->theta > 0 ? a->theta : 0.25;
        s.mg_prec = P(a->mg_prec);
        s.mg_store = P(a->mg_store);
        FH_CHECK(a->coarse_store == 0 || a->coarse_store == 1 || a->coarse_store == 2, -2,
                 "coarse_store: 0 (as mg_store), FH_F32 or FH_BF16");
        s.coarse_store = a->coarse_store == 0 ? s.mg_store : P(a->coarse_store);
        s.galerkin = a->galerkin != 0;
        if (a->coarse_smoother) s.coarse_smoother = a->coarse_smoother;
        s.coarse_relax = a->coarse_relax;
        s.fcg_truncation = a->fcg_truncation > 0 ? a->fcg_truncation : 1;
        s.kcycle_levels = a->kcycle_levels > 0 ? a->kcycle_levels : 0;
        FH_CHECK(a->nu_change_op == 0 || a->nu_change_op == '+' || a->nu_change_op == '-' || a->nu_change_op == '*',
                 -2, "smoothing change operator must be '+', '-' or '*'");
        s.nu_change_op = a->nu_change_op;
        s.nu_change_coeff = a->nu_change_coeff;
        FH_CHECK(a->outer_prec == 0 || a->outer_prec == 1, -2, "outer_prec: FH_F64 or FH_F32");
        s.outer_prec = P(a->outer_prec);
        s.replace_every = a->replace_every > 0 ? a->replace_every : 10;
        if (a->coarse_nu_given) {
            FH_CHECK(a->coarse_nu1 >= 0 && a->coarse_nu2 >= 0, -2, "coarse smoothing iterations must be >= 0");
            s.coarse_nu1 = a->coarse_nu1;
            s.coarse_nu2 = a->coarse_nu2;
        }
        FH_CHECK(s.outer_prec == fh::Prec::F64 || (s.solver == "fcg" && s.mg_store == fh::Prec::F32), -2,
                 "an FP32 outer product needs FCG and an FP32-stored multigrid");
        FH_CHECK(s.fcg_truncation <= 32, -2, "FCG truncation is limited to 32 directions");
        s.fcg_restart = a->fcg_restart != 0;
        if (!s.coarse_smoother.empty())
            (void)fh::SmootherSpec::parse(s.coarse_smoother, 1, s.coarse_relax > 0 ? s.coarse_relax : s.relax);
        FH_CHECK(s.mg_prec != fh::Prec::BF16, -2, "multigrid vectors are FP64 or FP32");
        FH_CHECK(s.mg_prec == fh::Prec::F32 || s.mg_store == fh::Prec::F64, -2,
                 "FP32/BF16 level matrices need FP32 multigrid vectors");
        FH_CHECK(s.mg_prec == fh::Prec::F64 || s.mg_store != fh::Prec::F64, -2,
                 "FP32 multigrid vectors need FP32 or BF16 level matrices");
        FH_CHECK(s.coarse_store == s.mg_store || s.mg_prec == fh::Prec::F32, -2,
                 "a separate coarse-level storage needs FP32 multigrid vectors");
        auto sv = new fh_solver;
        sv->ctx = ctx;
        sv->s = std::make_unique<fh::Solver>(ctx->c, s);
        *out = sv;
    });
}
int fh_solver_destroy(fh_solver* s) {
    return guard([&] { delete s; });
}
int fh_solver_set_level(fh_solver* s, int level, const fh_mat* A_level, const fh_mat* P_level) {
    return guard([&] {
        NEED(s);
        FH_CHECK(level >= 0 && level < 64, -2, "level out of range");
        s->s->SetLevel(level, A_level ? A_level->A.get() : nullptr, P_level ? P_level->A.get() : nullptr);
    });
}
int fh_solver_setup(fh_solver* s, const fh_mat* A) {
    return guard([&] {
        NEED(s); NEED(A);
        FH_CHECK(!s->ctx->c.host_only, -9, "solver set-up needs a device context");
        FH_CHECK(A->A->prec == fh::Prec::F64, -2, "the system matrix must be FP64");
        s->s->Setup(*A->A);
    });
}
int fh_solver_setup_uncondensed(fh_solver* s, const fh_mat* A, const fh_csr* A_TT, const fh_csr* A_TF,
                                const fh_csr* A_FF, int cell_bs) {
    return guard([&] {
        NEED(s); NEED(A);
        FH_CHECK(!s->ctx->c.host_only, -9, "solver set-up needs a device context");
        FH_CHECK(A->A->prec == fh::Prec::F64, -2, "the system matrix must be FP64");
        FH_CHECK(cell_bs > 0, -2, "cell block size must be positive");
        s->s->SetupUncondensed(*A->A, csr_in(A_TT, "A_TT"), csr_in(A_TF, "A_TF"), csr_in(A_FF, "A_FF"), cell_bs);
    });
}
int fh_solver_solve(fh_solver* s, const double* b, double* x, fh_solve_stats* st) {
    return fh_solver_solve_x0(s, b, x, 1, st);
}
int fh_solver_solve_x0(fh_solver* s, const double* b, double* x, int x_is_zero, fh_solve_stats* st) {
    return guard([&] {
        NEED(s); NEED(b); NEED(x);
        s->s->Solve(b, x, x_is_zero != 0);
        if (st) {
            const auto& t = s->s->Stats();
            st->iterations = t.iterations;
            st->converged = t.converged;
            st->levels = t.levels;
            st->final_res = t.final_res;
            st->setup_s = t.setup_s;
            st->solve_s = t.solve_s;
            st->op_complexity = t.op_complexity;
            st->nn_used = t.nn_used;
            st->nn_fallback = t.nn_fallback;
            st->graph_mode = t.graph_mode;
        }
    });
}

int fh_nn_create(fh_ctx* ctx, int nf, int layers, uint64_t seed, fh_nn** out) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(out);
        auto n = new fh_nn;
        try {
            n->m = std::make_unique<fh::Mpnn>(ctx->c, nf, layers, seed);
        } catch (...) {
            delete n;
            throw;
        }
        *out = n;
    });
}
int fh_nn_destroy(fh_nn* nn) {
    return guard([&] { delete nn; });
}
int fh_nn_num_params(const fh_nn* nn, int64_t* n) {
    return guard([&] { NEED(nn); NEED(n); *n = nn->m->shape().total(); });
}
int fh_nn_set_params(fh_nn* nn, const float* params) {
    return guard([&] { NEED(nn); NEED(params); nn->m->set_params(params); });
}
int fh_nn_get_params(const fh_nn* nn, float* params) {
    return guard([&] { NEED(nn); NEED(params); nn->m->get_params(params); });
}
int fh_nn_graph_create(fh_ctx* ctx, const fh_mat* A, const double* face_features, fh_nngraph** out) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(A); NEED(out);
        auto g = new fh_nngraph;
        try {
            g->g = fh::nn_graph_create(ctx->c, *A->A, face_features);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}
int fh_nn_graph_destroy(fh_nngraph* g) {
    return guard([&] { delete g; });
}
int fh_nn_forward(fh_ctx* ctx, fh_nn* nn, const fh_nngraph* g, const double* r1, double* delta, int bf16) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(nn); NEED(g); NEED(r1); NEED(delta);
        nn->m->reserve(g->g->n);
        nn->m->forward(*g->g, r1, delta, bf16 != 0);
    });
}
int fh_nn_loss_grad(fh_ctx* ctx, fh_nn* nn, const fh_nngraph* g, const double* r1, const double* e1, double* loss,
                    float* grad) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(nn); NEED(g); NEED(r1); NEED(e1);
        double l = nn->m->loss_grad(*g->g, r1, e1, grad);
        if (loss) *loss = l;
    });
}
int fh_nn_train_step(fh_ctx* ctx, fh_nn* nn, const fh_nngraph* g, const double* r1, const double* e1, float lr,
                     double* loss) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(nn); NEED(g); NEED(r1); NEED(e1);
        FH_CHECK(lr > 0, -2, "learning rate must be positive");
        double l = nn->m->train_step(*g->g, r1, e1, lr);
        if (loss) *loss = l;
    });
}
int fh_solver_set_nn(fh_solver* s, fh_nn* nn, const fh_nngraph* g, double omega, int unit_step) {
    return guard([&] {
        NEED(s);
        FH_CHECK((nn == nullptr) == (g == nullptr), -2, "network and graph go together");
        s->s->SetLearnedSmoother(nn ? nn->m.get() : nullptr, g ? g->g.get() : nullptr, omega > 0 ? omega : 1.0,
                                 unit_step != 0);
    });
}
int fh_solver_export_levels(fh_solver* s, const char* dir) {
    return guard([&] {
        NEED(s); NEED(dir);
        s->s->ExportLevels(dir);
    });
}
int fh_solver_set_level_p(fh_solver* s, int level, int nfc) {
    return guard([&] {
        NEED(s);
        s->s->SetLevelP(level, nfc);
    });
}
int fh_solver_set_stop(fh_solver* s, double tol, int max_iter) {
    return guard([&] { NEED(s); s->s->SetStop(tol, max_iter); });
}
int fh_solver_apply_prec(fh_solver* s, const double* r, double* z) {
    return guard([&] {
        NEED(s); NEED(r); NEED(z);
        s->s->ApplyPreconditioner(r, z);
    });
}
int fh_solver_history(const fh_solver* s, double* out, int cap) {
    int n = 0;
    int rc = guard([&] {
        NEED(s);
        const auto& h = s->s->History();
        n = (int)std::min<size_t>(h.size(), (size_t)cap);
        if (out) std::memcpy(out, h.data(), sizeof(double) * n);
    });
    return rc ? rc : n;
}
struct fh_hho {
    std::unique_ptr<fh::HhoProblem> p;
    // uncondensed blocks, assembled at the first request
    std::unique_ptr<fh::Csr> unc[3];
};
struct fh_uamg {
    fh::UamgHierarchy h;
};

int fh_hho_uncondensed(fh_hho* h, int which, int* nrows, int* ncols, int64_t* nnz, int* rowptr, int* colidx,
                       double* vals) {
    return guard([&] {
        NEED(h);
        FH_CHECK(which >= 0 && which < 3, -2, "which: 0 = A_TT, 1 = A_TF, 2 = A_FF");
        if (!h->unc[0]) {
            fh::Csr tt, tf, ff;
            h->p->Uncondensed(tt, tf, ff);
            h->unc[0] = std::make_unique<fh::Csr>(std::move(tt));
            h->unc[1] = std::make_unique<fh::Csr>(std::move(tf));
            h->unc[2] = std::make_unique<fh::Csr>(std::move(ff));
        }
        csr_out(*h->unc[which], nrows, ncols, nnz, rowptr, colidx, vals);
    });
}
int fh_hho_cell_unknowns(const fh_hho* h, int* nc) {
    return guard([&] {
        NEED(h); NEED(nc);
        *nc = h->p->cell_unknowns();
    });
}

int fh_uamg_build(const fh_csr* A, const fh_csr* A_TT, const fh_csr* A_TF, const fh_csr* A_FF, int cell_bs,
                  int face_bs, double theta, int max_coarse, fh_uamg** out) {
    return guard([&] {
        NEED(out);
        fh::UamgArgs a;
        a.cell_bs = cell_bs;
        a.face_bs = face_bs;
        a.theta = theta;
        a.max_coarse = max_coarse;
        auto u = std::make_unique<fh_uamg>();
        u->h = fh::uamg_build(csr_in(A, "A"), csr_in(A_TT, "A_TT"), csr_in(A_TF, "A_TF"), csr_in(A_FF, "A_FF"), a);
        *out = u.release();
    });
}
int fh_uamg_destroy(fh_uamg* u) {
    return guard([&] { delete u; });
}
int fh_uamg_info(const fh_uamg* u, int* nlevels, int* npasses) {
    return guard([&] {
        NEED(u);
        if (nlevels) *nlevels = (int)u->h.A.size();
        if (npasses) *npasses = (int)u->h.passes.size();
    });
}
int fh_uamg_level(const fh_uamg* u, int l, int which, int* nrows, int* ncols, int64_t* nnz, int* rowptr, int* colidx,
                  double* vals) {
    return guard([&] {
        NEED(u);
        FH_CHECK(which == 0 || which == 1, -2, "which: 0 = operator, 1 = prolongation");
        const auto& v = which == 0 ? u->h.A : u->h.P;
        FH_CHECK(l >= 0 && l < (int)v.size(), -2, "level out of range");
        csr_out(v[l], nrows, ncols, nnz, rowptr, colidx, vals);
    });
}
int fh_uamg_pass(const fh_uamg* u, int k, int* level, int* pass, int* counts, int* elem_agg, int* face_coarse) {
    return guard([&] {
        NEED(u);
        FH_CHECK(k >= 0 && k < (int)u->h.passes.size(), -2, "pass out of range");
        const fh::UamgPass& p = u->h.passes[k];
        if (level) *level = p.level;
        if (pass) *pass = p.pass;
        if (counts) {
            counts[0] = p.n_elems; counts[1] = p.n_faces;
            counts[2] = p.n_coarse_elems; counts[3] = p.n_coarse_faces;
        }
        if (elem_agg && !p.elem_agg.empty()) std::memcpy(elem_agg, p.elem_agg.data(), sizeof(int) * p.elem_agg.size());
        if (face_coarse && !p.face_coarse.empty())
            std::memcpy(face_coarse, p.face_coarse.data(), sizeof(int) * p.face_coarse.size());
    });
}

int fh_hho_create(const fh_hho_args* a, fh_hho** out) {
    return guard([&] {
        NEED(a); NEED(out);
        fh::HhoArgs h;
        h.dim = a->dim;
        if (a->mesh) h.mesh = a->mesh;
        h.n = a->n;
        h.k = a->k;
        h.perturb = a->perturb;
        h.seed = a->seed;
        if (a->source) h.source = a->source;
        if (a->kappa) h.kappa = a->kappa;
        h.sigma = a->sigma;
        if (a->subdomain > 0) h.subdomain = a->subdomain;
        h.aniso = a->aniso > 0 ? a->aniso : 1e-2;
        h.ref_n = a->ref_n;
        h.source_seed = a->source_seed;
        h.source_modes = a->source_modes > 0 ? a->source_modes : 8;
        h.host_assembly = a->defer_assembly == 0;
        FH_CHECK(h.dim == 2 || h.dim == 3, -2, "dim must be 2 or 3");
        auto o = new fh_hho;
        try {
            o->p = std::make_unique<fh::HhoProblem>(h);
        } catch (...) {
            delete o;
            throw;
        }
        *out = o;
    });
}
int fh_hho_destroy(fh_hho* h) {
    return guard([&] { delete h; });
}
int fh_hho_get_info(const fh_hho* h, fh_hho_info* info) {
    return guard([&] {
        NEED(h); NEED(info);
        const auto& p = *h->p;
        info->nb = p.nb;
        info->nf = p.nf;
        info->ncells = p.ncells();
        info->nfaces = p.nfaces();
        info->nnzb = p.rowptr.empty() ? 0 : p.rowptr.back();
        info->mesh_s = p.times.mesh_s;
        info->local_s = p.times.local_s;
        info->assemble_s = p.times.assemble_s;
        info->redraws = p.redraws();
        info->kept_vertices = p.kept_vertices();
        info->min_quality = p.min_quality();
    });
}
int fh_hho_matrix(fh_ctx* ctx, const fh_hho* h, int keep_host, fh_mat** out) {
    return guard([&] {
        NEED(ctx); NEED(h); NEED(out);
        const auto& p = *h->p;
        FH_CHECK(p.host_assembled, -2, "problem created with deferred assembly: use fh_hho_assemble_gpu");
        auto m = new fh_mat;
        m->A = fh::bcsr_from_blocks(ctx->c, p.nb, p.nb, p.nf, p.nf, p.rowptr.data(), p.colidx.data(), p.vals.data(),
                                    keep_host != 0);
        *out = m;
    });
}
int fh_hho_assemble_gpu(fh_ctx* ctx, const fh_hho* h, int keep_host, fh_mat** out, double* b_dev) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(h); NEED(out); NEED(b_dev);
        auto m = new fh_mat;
        try {
            m->A = h->p->AssembleDevice(ctx->c, b_dev, keep_host != 0);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}
int fh_hho_rhs(const fh_hho* h, double* b) {
    return guard([&] {
        NEED(h); NEED(b);
        FH_CHECK(h->p->host_assembled, -2, "problem created with deferred assembly: the right-hand side is on the device");
        std::memcpy(b, h->p->b.data(), sizeof(double) * h->p->b.size());
    });
}
int fh_hho_blocks(const fh_hho* h, int* rowptr, int* colidx, double* vals) {
    return guard([&] {
        NEED(h);
        const auto& p = *h->p;
        if (rowptr) std::memcpy(rowptr, p.rowptr.data(), sizeof(int) * p.rowptr.size());
        if (colidx) std::memcpy(colidx, p.colidx.data(), sizeof(int) * p.colidx.size());
        if (vals) {
            FH_CHECK(p.host_assembled, -2, "problem created with deferred assembly: no host values");
            std::memcpy(vals, p.vals.data(), sizeof(double) * p.vals.size());
        }
    });
}
int fh_hho_l2_error(const fh_hho* h, const double* x, double* err) {
    return guard([&] {
        NEED(h); NEED(x); NEED(err);
        *err = h->p->L2Error(x);
    });
}
int fh_hho_l2_error_gpu(fh_ctx* ctx, const fh_hho* h, const double* x_dev, double* err) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(h); NEED(x_dev); NEED(err);
        *err = h->p->L2ErrorDevice(ctx->c, x_dev);
    });
}
int fh_hho_prolongation(fh_ctx* ctx, const fh_hho* fine, const fh_hho* coarse, int keep_host, fh_mat** out) {
    return guard([&] {
        NEED(ctx); NEED(fine); NEED(coarse); NEED(out);
        std::vector<int> rp, ci;
        std::vector<double> v;
        fine->p->Prolongation(*coarse->p, rp, ci, v);
        auto m = new fh_mat;
        m->A = fh::bcsr_from_blocks(ctx->c, fine->p->nb, coarse->p->nb, fine->p->nf, fine->p->nf, rp.data(), ci.data(),
                                    v.data(), keep_host != 0);
        *out = m;
    });
}
int fh_hho_face_cells(const fh_hho* h, int* out) {
    return guard([&] {
        NEED(h); NEED(out);
        std::vector<int> v;
        h->p->FaceCells(v);
        std::memcpy(out, v.data(), sizeof(int) * v.size());
    });
}
int fh_condense_batched(fh_ctx* ctx, int ncells, int nfc, int nc, int nf, int nfaces, const int* cell_faces,
                        const double* local, const double* B_T, const double* B_F, double* b_out, fh_mat** out) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(cell_faces); NEED(local); NEED(out);
        auto A = fh::condense_batched(ctx->c, ncells, nfc, nc, nf, nfaces, cell_faces, local, B_T, B_F, b_out);
        auto m = new fh_mat;
        m->A = std::move(A);
        *out = m;
    });
}
int fh_cell_recover_batched(fh_ctx* ctx, int ncells, int nfc, int nc, int nf, int nfaces, const int* cell_faces,
                            const double* local, const double* B_T, const double* x_F, double* x_T) {
    return guard([&] {
        NEED(ctx); NEED_DEV(ctx); NEED(cell_faces); NEED(local); NEED(x_F); NEED(x_T);
        fh::cell_recover_batched(ctx->c, ncells, nfc, nc, nf, nfaces, cell_faces, local, B_T, x_F, x_T);
    });
}
int fh_hho_cell_kappa(const fh_hho* h, double* out) {
    return guard([&] {
        NEED(h); NEED(out);
        auto k = h->p->CellKappa();
        std::memcpy(out, k.data(), sizeof(double) * k.size());
    });
}

int fh_solver_level_size(const fh_solver* s, int level) {
    int n = 0;
    int rc = guard([&] {
        NEED(s);
        FH_CHECK(level >= 0 && level < s->s->NumLevels(), -2, "level out of range");
        n = s->s->LevelUnknowns(level);
    });
    return rc ? rc : n;
}

}  // extern "C"
