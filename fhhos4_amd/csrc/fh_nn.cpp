// Host side of the learned post-smoother (network state, forward, loss,
// backward, Adam); the arithmetic runs in fh_nn_kernels.hip.
#include "fh_nn.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fh_kernels.h"
#include "fh_solver.h"

namespace fh {

std::unique_ptr<NnGraph> nn_graph_create(Ctx& c, const Bcsr& A, const double* face_features) {
    FH_CHECK(!c.host_only, -9, "the network graph lives on the device");
    FH_CHECK(A.bs_r == A.bs_c && A.nbr == A.nbc && A.prec == Prec::F64, -2, "square FP64 block operator expected");
    auto g = std::make_unique<NnGraph>();
    g->A = &A;
    g->n = A.nbr;
    g->nf = A.bs_r;
    g->w.alloc(sizeof(float) * std::max<int64_t>(1, A.nnzb));
    g->wT.alloc(sizeof(float) * std::max<int64_t>(1, A.nnzb));
    nn_edge_weights(c, A, g->w.as<float>());
    // index of the reverse block (j, i) of every stored block (i, j)
    std::vector<int> tpos(A.nnzb, 0);
    std::vector<std::pair<int, int>> rowcols;
    std::vector<std::vector<std::pair<int, int>>> sorted(A.nbr);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < A.nbr; i++) {
        auto& v = sorted[i];
        for (int p = A.h_rowptr[i]; p < A.h_rowptr[i + 1]; p++) v.push_back({A.h_colidx[p], p});
        std::sort(v.begin(), v.end());
    }
    int missing = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : missing)
    for (int i = 0; i < A.nbr; i++)
        for (int p = A.h_rowptr[i]; p < A.h_rowptr[i + 1]; p++) {
            const auto& v = sorted[A.h_colidx[p]];
            auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(i, -1));
            if (it == v.end() || it->first != i) missing++;
            else tpos[p] = it->second;
        }
    FH_CHECK(missing == 0, -2, "the block pattern of the operator is not symmetric");
    DevBuf d_tpos(sizeof(int) * std::max<size_t>(1, tpos.size()));
    FH_HIP(hipMemcpyAsync(d_tpos.p, tpos.data(), sizeof(int) * tpos.size(), hipMemcpyHostToDevice, c.stream));
    nn_gather_index(c, A.nnzb, d_tpos.as<int>(), g->w.as<float>(), g->wT.as<float>());
    std::vector<float> f(2 * (size_t)A.nbr, 0.f);
    if (face_features)
        for (size_t k = 0; k < f.size(); k++) f[k] = (float)face_features[k];
    g->feat.alloc(sizeof(float) * f.size());
    FH_HIP(hipMemcpyAsync(g->feat.p, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice, c.stream));
    LevelSmoother s32, s64;
    smoother_setup(c, A, s32, Prec::F32);
    smoother_setup(c, A, s64, Prec::F64);
    g->invD32 = std::move(*s32.invD);
    g->invD64 = std::move(*s64.invD);
    c.sync();
    return g;
}

namespace {
// Deterministic uniform(-1, 1) stream (splitmix64).
struct Rng {
    uint64_t s;
    double next() {
        uint64_t x = (s += 0x9E3779B97F4A7C15ull);
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
};
}  // namespace

Mpnn::Mpnn(Ctx& c, int nf, int layers, uint64_t seed) : c_(c) {
    FH_CHECK(!c.host_only, -9, "the network lives on the device");
    FH_CHECK(nf >= 1 && nf <= 10, -2, "face block size must be 1..10");
    FH_CHECK(layers >= 1 && layers <= kNnMaxLayers, -2, "1..6 message-passing layers");
    sh_.nf = nf;
    sh_.layers = layers;
    const int64_t P = sh_.total();
    // He-uniform weights (relu layers), zero biases, zero output layer: the
    // untrained network returns the block-Jacobi step.
    std::vector<float> p(P, 0.f);
    Rng r{seed};
    const int H = kNnHidden;
    const double ae = std::sqrt(6.0 / sh_.din()), am = std::sqrt(6.0 / (2 * H));
    for (int64_t k = 0; k < (int64_t)H * sh_.din(); k++) p[sh_.We() + k] = (float)(ae * r.next());
    for (int l = 0; l < layers; l++)
        for (int k = 0; k < H * H; k++) {
            p[sh_.S(l) + k] = (float)(am * r.next());
            p[sh_.N(l) + k] = (float)(am * r.next());
        }
    p32_.alloc(sizeof(float) * P);
    p16_.alloc(sizeof(uint16_t) * P);
    grad_.alloc(sizeof(float) * P);
    m1_.alloc(sizeof(float) * P);
    m2_.alloc(sizeof(float) * P);
    partials_.alloc(nn_partials_bytes());
    launch_fill(c_, P, m1_.p, Prec::F32, 0.0);
    launch_fill(c_, P, m2_.p, Prec::F32, 0.0);
    set_params(p.data());
}

void Mpnn::set_params(const float* p) {
    const int64_t P = sh_.total();
    FH_HIP(hipMemcpyAsync(p32_.p, p, sizeof(float) * P, hipMemcpyHostToDevice, c_.stream));
    nn_to_bf16(c_, (int)P, p32_.as<float>(), p16_.p);
    c_.sync();
}

void Mpnn::get_params(float* p) const {
    FH_HIP(hipMemcpyAsync(p, p32_.p, sizeof(float) * sh_.total(), hipMemcpyDeviceToHost, c_.stream));
    c_.sync();
}

void Mpnn::reserve(int n) {
    if (n <= cap_) return;
    const int64_t H = kNnHidden, nf = sh_.nf, L = sh_.layers;
    v_.alloc(sizeof(float) * n * nf);
    u_.alloc(sizeof(float) * n * nf);
    gy_.alloc(sizeof(float) * n * nf);
    x0_.alloc(sizeof(float) * n * sh_.din());
    h_.alloc(sizeof(float) * (L + 1) * n * H);
    m_.alloc(sizeof(float) * L * n * H);
    dh_.alloc(sizeof(float) * n * H);
    dh2_.alloc(sizeof(float) * n * H);
    t_.alloc(sizeof(float) * n * H);
    delta_.alloc(sizeof(double) * n * nf);
    diff_.alloc(sizeof(double) * n * nf);
    Adiff_.alloc(sizeof(double) * n * nf);
    g_.alloc(sizeof(double) * n * nf);
    cap_ = n;
}

void Mpnn::forward(const NnGraph& g, const double* r1, double* delta, bool bf16) {
    FH_CHECK(g.nf == sh_.nf, -2, "network and operator have different face block sizes");
    FH_CHECK(g.n <= cap_, -2, "network workspace too small: reserve() first");
    const int n = g.n, nf = g.nf, L = sh_.layers;
    const int64_t H = kNnHidden, nH = (int64_t)n * H;
    const size_t es = bf16 ? 2 : 4;
    const char* P = (const char*)(bf16 ? p16_.p : p32_.p);
    auto par = [&](int64_t off) { return (const void*)(P + off * es); };
    nn_prep(c_, n, nf, r1, g.invD32.as<float>(), v_.as<float>());
    const void* xs[1] = {v_.p};
    const void* ys[1] = {v_.p};
    int sl[1] = {kNnSlotVV};
    launch_dots(c_, (int64_t)n * nf, Prec::F32, 1, xs, ys, sl);
    nn_input(c_, n, nf, v_.as<float>(), g.feat.as<float>(), kNnSlotVV, kNnSlotSigma, u_.as<float>(),
             x0_.as<float>());
    float* h = h_.as<float>();
    nn_encode(c_, n, sh_.din(), x0_.as<float>(), par(sh_.We()), par(sh_.be()), bf16, h);
    for (int l = 0; l < L; l++)
        nn_layer(c_, n, g.A->rowptr.as<int>(), g.A->colidx.as<int>(), g.w.as<float>(), h + l * nH, par(sh_.S(l)),
                 par(sh_.N(l)), par(sh_.c(l)), bf16, h + (l + 1) * nH, bf16 ? nullptr : m_.as<float>() + l * nH);
    nn_out(c_, n, nf, h + L * nH, u_.as<float>(), par(sh_.Wo()), par(sh_.bo()), bf16, kNnSlotSigma, delta);
}

// After forward(train) into delta_: loss = (e1-delta)^T A (e1-delta) / e1^T r1
// and g_ = dloss/ddelta = -2 A (e1 - delta) / (e1^T r1).
double Mpnn::loss_and_seed(const NnGraph& g, const double* e1) {
    const int64_t N = (int64_t)g.n * g.nf;
    FH_HIP(hipMemcpyAsync(diff_.p, e1, sizeof(double) * N, hipMemcpyDeviceToDevice, c_.stream));
    launch_axpby(c_, N, Prec::F64, delta_.p, diff_.p, -1.0, -1, -1, 1.0, -1, -1);
    launch_bcsr(c_, *g.A, OP_SPMV, Prec::F64, diff_.p, nullptr, Adiff_.p, 1.0, 0.0, 0.0, nullptr, Prec::F64, nullptr,
                0);
    const void* xs[1] = {diff_.p};
    const void* ys[1] = {Adiff_.p};
    int sl[1] = {kNnSlotE};
    launch_dots(c_, N, Prec::F64, 1, xs, ys, sl);
    FH_HIP(hipMemcpyAsync(c_.h_scalars + kNnSlotE, c_.d_scalars + kNnSlotE, sizeof(double) * 2, hipMemcpyDeviceToHost,
                          c_.stream));
    c_.sync();
    const double E = c_.h_scalars[kNnSlotE], E0 = c_.h_scalars[kNnSlotE0];
    FH_CHECK(E0 > 0, -2, "training pair with zero energy (e1^T r1 <= 0)");
    FH_HIP(hipMemcpyAsync(g_.p, Adiff_.p, sizeof(double) * N, hipMemcpyDeviceToDevice, c_.stream));
    launch_axpby(c_, N, Prec::F64, Adiff_.p, g_.p, -2.0 / E0, -1, -1, 0.0, -1, -1);
    return E / E0;
}

void Mpnn::backward(const NnGraph& g, const double* gdelta) {
    const int n = g.n, nf = g.nf, L = sh_.layers, din = sh_.din();
    const int64_t H = kNnHidden, nH = (int64_t)n * H;
    const float* P = p32_.as<float>();
    float* G = grad_.as<float>();
    double* part = partials_.as<double>();
    const float* h = h_.as<float>();
    const float* m = m_.as<float>();
    float* dh = dh_.as<float>();
    float* dh2 = dh2_.as<float>();
    nn_out_back(c_, n, nf, gdelta, kNnSlotSigma, P + sh_.Wo(), gy_.as<float>(), dh);
    nn_outer(c_, n, nf, (int)H, gy_.as<float>(), h + L * nH, part, G + sh_.Wo());
    nn_outer(c_, n, nf, 1, gy_.as<float>(), nullptr, part, G + sh_.bo());
    for (int l = L - 1; l >= 0; l--) {
        nn_relu_back(c_, nH, h + (l + 1) * nH, dh);
        nn_outer(c_, n, (int)H, (int)H, dh, h + l * nH, part, G + sh_.S(l));
        nn_outer(c_, n, (int)H, (int)H, dh, m + l * nH, part, G + sh_.N(l));
        nn_outer(c_, n, (int)H, 1, dh, nullptr, part, G + sh_.c(l));
        nn_back_self(c_, n, dh, P + sh_.S(l), P + sh_.N(l), dh2, t_.as<float>());
        nn_back_gather(c_, n, g.A->rowptr.as<int>(), g.A->colidx.as<int>(), g.wT.as<float>(), t_.as<float>(), dh2);
        std::swap(dh, dh2);
    }
    nn_relu_back(c_, nH, h, dh);
    nn_outer(c_, n, (int)H, din, dh, x0_.as<float>(), part, G + sh_.We());
    nn_outer(c_, n, (int)H, 1, dh, nullptr, part, G + sh_.be());
}

double Mpnn::loss_grad(const NnGraph& g, const double* r1, const double* e1, float* grad_host) {
    reserve(g.n);
    const int64_t N = (int64_t)g.n * g.nf;
    const void* xs[1] = {e1};
    const void* ys[1] = {r1};
    int sl[1] = {kNnSlotE0};
    launch_dots(c_, N, Prec::F64, 1, xs, ys, sl);
    forward(g, r1, delta_.as<double>(), false);
    const double loss = loss_and_seed(g, e1);
    backward(g, g_.as<double>());
    if (grad_host) {
        FH_HIP(hipMemcpyAsync(grad_host, grad_.p, sizeof(float) * sh_.total(), hipMemcpyDeviceToHost, c_.stream));
        c_.sync();
    }
    return loss;
}

double Mpnn::train_step(const NnGraph& g, const double* r1, const double* e1, float lr) {
    const double loss = loss_grad(g, r1, e1, nullptr);
    step_++;
    nn_adam(c_, (int)sh_.total(), p32_.as<float>(), grad_.as<float>(), m1_.as<float>(), m2_.as<float>(), lr, 0.9f,
            0.999f, 1e-8f, step_, p16_.p);
    return loss;
}

}  // namespace fh
