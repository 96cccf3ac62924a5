// gfx950 kernels of the learned post-smoother: a message-passing network over
// the face graph of the condensed operator (one node per block row, edges =
// off-diagonal blocks), hidden width 32.
//
// Layout: node features are row-major n x 32 FP32 arrays (one 128-byte row
// per face), so the 32 lanes of a node read or write one row with one
// coalesced access.  A workgroup of 256 threads = 8 nodes x 32 channels; the
// layer weights are staged in LDS once per workgroup, the node's own row and
// its neighbour aggregate are staged per node, and every lane computes one
// output channel (a 64-long dot product from LDS).  Weights are FP32 for
// training and BF16 (FP32 accumulate) for inference.
//
// Weight gradients are sums over nodes of outer products; each workgroup of
// a fixed-size grid accumulates its own share in FP64 and a second kernel
// adds the shares in a fixed order, so gradients are deterministic.
#include "fh_nn.h"

namespace fh {
namespace {

constexpr int H = kNnHidden;  // 32
constexpr int kNodesPerWG = 8;
constexpr int kOuterWGs = 256;

struct bf16w {
    uint16_t u;
};
__device__ __forceinline__ float wload(const float* p) { return *p; }
__device__ __forceinline__ float wload(const bf16w* p) { return __uint_as_float(((uint32_t)p->u) << 16); }

// v = D^-1 r1 (FP32), thread per unknown.
__global__ void k_nn_prep(int n, int nf, const double* __restrict__ r1, const float* __restrict__ invD,
                          float* __restrict__ v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * nf) return;
    const int i = (int)(t / nf), a = (int)(t - (int64_t)i * nf);
    const float* D = invD + ((int64_t)i * nf + a) * nf;
    const double* r = r1 + (int64_t)i * nf;
    float s = 0.f;
    for (int b = 0; b < nf; b++) s += D[b] * (float)r[b];
    v[t] = s;
}

// sigma = sqrt(<v,v> / (n nf)); u = v / sigma; x0 = [u, feat]; sigma -> s[slot_sigma].
__global__ void k_nn_input(int n, int nf, const float* __restrict__ v, const float* __restrict__ feat,
                           double* s, int slot_vv, int slot_sigma, float* __restrict__ u, float* __restrict__ x0) {
    const int din = nf + 2;
    const double sig = sqrt(s[slot_vv] / ((double)n * nf));
    const float inv = sig > 0 ? (float)(1.0 / sig) : 0.f;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) s[slot_sigma] = sig;
    if (t >= (int64_t)n * din) return;
    const int i = (int)(t / din), a = (int)(t - (int64_t)i * din);
    float val;
    if (a < nf) {
        val = v[(int64_t)i * nf + a] * inv;
        u[(int64_t)i * nf + a] = val;
    } else
        val = feat[(int64_t)i * 2 + (a - nf)];
    x0[t] = val;
}

// Encoder: h = relu(We x + be), x of width DIN (<= 12).
template <class TW>
__global__ __launch_bounds__(256) void k_nn_encode(int n, int din, const float* __restrict__ x,
                                                    const TW* __restrict__ We, const TW* __restrict__ be,
                                                    float* __restrict__ h) {
    const int c = threadIdx.x & (H - 1);
    const int64_t i = (int64_t)blockIdx.x * kNodesPerWG + (threadIdx.x >> 5);
    if (i >= n) return;
    float acc = wload(be + c);
    const float* xi = x + i * din;
    for (int k = 0; k < din; k++) acc += wload(We + c * din + k) * xi[k];
    h[i * H + c] = acc > 0.f ? acc : 0.f;
}

// Message-passing layer: m_i = sum_j w_ij h_j ; h'_i = relu(S h_i + N m_i + c).
// m is stored for the backward pass when m_out != nullptr.
template <class TW>
__global__ __launch_bounds__(256) void k_nn_layer(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const float* __restrict__ w, const float* __restrict__ hin,
                                                   const TW* __restrict__ S, const TW* __restrict__ N,
                                                   const TW* __restrict__ cb, float* __restrict__ hout,
                                                   float* __restrict__ m_out) {
    __shared__ float sS[H * H], sN[H * H];
    __shared__ float sx[kNodesPerWG][2 * H];
    for (int e = threadIdx.x; e < H * H; e += blockDim.x) {
        sS[e] = wload(S + e);
        sN[e] = wload(N + e);
    }
    const int c = threadIdx.x & (H - 1), q = threadIdx.x >> 5;
    const int64_t i = (int64_t)blockIdx.x * kNodesPerWG + q;
    const bool ok = i < n;
    float m = 0.f, hs = 0.f;
    if (ok) {
        hs = hin[i * H + c];
        for (int p = rp[i]; p < rp[i + 1]; p++) m += w[p] * hin[(int64_t)ci[p] * H + c];
        if (m_out) m_out[i * H + c] = m;
    }
    sx[q][c] = hs;
    sx[q][H + c] = m;
    __syncthreads();
    if (!ok) return;
    float acc = wload(cb + c);
#pragma unroll 8
    for (int k = 0; k < H; k++) acc += sS[c * H + k] * sx[q][k] + sN[c * H + k] * sx[q][H + k];
    hout[i * H + c] = acc > 0.f ? acc : 0.f;
}

// Output: delta_i = sigma (u_i + Wo h_i + bo), thread per unknown.
template <class TW>
__global__ void k_nn_out(int n, int nf, const float* __restrict__ h, const float* __restrict__ u,
                         const TW* __restrict__ Wo, const TW* __restrict__ bo, const double* s, int slot_sigma,
                         double* __restrict__ delta) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * nf) return;
    const int i = (int)(t / nf), a = (int)(t - (int64_t)i * nf);
    const float* hi = h + (int64_t)i * H;
    float y = wload(bo + a);
#pragma unroll 8
    for (int k = 0; k < H; k++) y += wload(Wo + a * H + k) * hi[k];
    delta[t] = s[slot_sigma] * (double)(u[t] + y);
}

// ---------------------------------------------------------------- backward
// gy = sigma g (FP32) and dh = Wo^T gy, thread per (node, channel).
__global__ __launch_bounds__(256) void k_nn_out_back(int n, int nf, const double* __restrict__ g,
                                                      const double* s, int slot_sigma, const float* __restrict__ Wo,
                                                      float* __restrict__ gy, float* __restrict__ dh) {
    const int c = threadIdx.x & (H - 1);
    const int64_t i = (int64_t)blockIdx.x * kNodesPerWG + (threadIdx.x >> 5);
    if (i >= n) return;
    const double sig = s[slot_sigma];
    float acc = 0.f;
    for (int a = 0; a < nf; a++) {
        const float ga = (float)(sig * g[i * nf + a]);
        if (c == 0) gy[i * nf + a] = ga;
        acc += Wo[a * H + c] * ga;
    }
    dh[i * H + c] = acc;
}

// dpre = dh * (h > 0), in place.
__global__ void k_nn_relu_back(int64_t count, const float* __restrict__ h, float* __restrict__ dh) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count && !(h[t] > 0.f)) dh[t] = 0.f;
}

// dprev_i = S^T dpre_i ; t_i = N^T dpre_i.
__global__ __launch_bounds__(256) void k_nn_back_self(int n, const float* __restrict__ dpre,
                                                       const float* __restrict__ S, const float* __restrict__ N,
                                                       float* __restrict__ dprev, float* __restrict__ tvec) {
    __shared__ float sS[H * H], sN[H * H];
    __shared__ float sd[kNodesPerWG][H];
    for (int e = threadIdx.x; e < H * H; e += blockDim.x) {
        sS[e] = S[e];
        sN[e] = N[e];
    }
    const int c = threadIdx.x & (H - 1), q = threadIdx.x >> 5;
    const int64_t i = (int64_t)blockIdx.x * kNodesPerWG + q;
    const bool ok = i < n;
    sd[q][c] = ok ? dpre[i * H + c] : 0.f;
    __syncthreads();
    if (!ok) return;
    float a = 0.f, b = 0.f;
#pragma unroll 8
    for (int k = 0; k < H; k++) {
        a += sS[k * H + c] * sd[q][k];
        b += sN[k * H + c] * sd[q][k];
    }
    dprev[i * H + c] = a;
    tvec[i * H + c] = b;
}

// dprev_i += sum_{p in row i} wT_p t_{col(p)}  (wT_p = weight of the reverse edge).
__global__ __launch_bounds__(256) void k_nn_back_gather(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                         const float* __restrict__ wT, const float* __restrict__ tvec,
                                                         float* __restrict__ dprev) {
    const int c = threadIdx.x & (H - 1);
    const int64_t i = (int64_t)blockIdx.x * kNodesPerWG + (threadIdx.x >> 5);
    if (i >= n) return;
    float acc = 0.f;
    for (int p = rp[i]; p < rp[i + 1]; p++) acc += wT[p] * tvec[(int64_t)ci[p] * H + c];
    dprev[i * H + c] += acc;
}

// Partial sums over nodes of A_i B_i^T (DA x DB; B == nullptr: a column of
// ones, i.e. the sum of A_i).  Grid of kOuterWGs workgroups, node i handled
// by workgroup i % kOuterWGs; FP64 accumulation.
__global__ __launch_bounds__(256) void k_nn_outer(int n, int da, int db, const float* __restrict__ Am,
                                                   const float* __restrict__ Bm, double* __restrict__ partials) {
    const int ne = da * db;
    double acc[4] = {0, 0, 0, 0};
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const float* a = Am + i * da;
        const float* b = Bm ? Bm + i * db : nullptr;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int e = threadIdx.x + r * 256;
            if (e < ne) {
                const int x = e / db, y = e - x * db;
                acc[r] += (double)a[x] * (b ? (double)b[y] : 1.0);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int e = threadIdx.x + r * 256;
        if (e < ne) partials[(int64_t)blockIdx.x * ne + e] = acc[r];
    }
}

// grad[off + e] = sum over workgroups of partials (fixed order).
__global__ void k_nn_reduce(int ne, int nwg, const double* __restrict__ partials, float* __restrict__ grad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double s = 0;
    for (int k = 0; k < nwg; k++) s += partials[(int64_t)k * ne + e];
    grad[e] = (float)s;
}

// Adam (Kingma & Ba 2015) with bias correction, and the BF16 copy of the
// updated weights (round to nearest even).
__global__ void k_nn_adam(int np, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                          float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
                          bf16w* __restrict__ wb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const float gt = g[t];
    const float a = b1 * m1[t] + (1.f - b1) * gt;
    const float b = b2 * m2[t] + (1.f - b2) * gt * gt;
    m1[t] = a;
    m2[t] = b;
    const float v = w[t] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    w[t] = v;
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    wb[t].u = (uint16_t)(u >> 16);
}

__global__ void k_nn_to_bf16(int np, const float* __restrict__ w, bf16w* __restrict__ wb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    uint32_t u = __float_as_uint(w[t]);
    u += 0x7FFFu + ((u >> 16) & 1u);
    wb[t].u = (uint16_t)(u >> 16);
}

// Edge weights w_p = |tr A_p| / sum_{q in row, q != diag} |tr A_q| (w = 0 on
// the diagonal block), thread per block row.
__global__ void k_nn_edge_weights(int n, int nf, const int* __restrict__ rp, const int* __restrict__ ci,
                                  const double* __restrict__ vals, float* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sum = 0;
    for (int p = rp[i]; p < rp[i + 1]; p++) {
        if (ci[p] == i) continue;
        double tr = 0;
        for (int a = 0; a < nf; a++) tr += vals[((int64_t)p * nf + a) * nf + a];
        sum += fabs(tr);
    }
    for (int p = rp[i]; p < rp[i + 1]; p++) {
        double tr = 0;
        for (int a = 0; a < nf; a++) tr += vals[((int64_t)p * nf + a) * nf + a];
        w[p] = (ci[p] == i || sum == 0) ? 0.f : (float)(fabs(tr) / sum);
    }
}

__global__ void k_nn_gather_index(int64_t nnzb, const int* __restrict__ idx, const float* __restrict__ src,
                                  float* __restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nnzb) dst[t] = src[idx[t]];
}

// Post-smoother selection (device-side, no host decision):
// s1 = delta.r1, s2 = delta.A delta.  If s1 > 0 and s2 > 0 the correction
// alpha delta with alpha = s1 / s2 lowers the energy norm of the error by
// s1^2 / s2; z += alpha delta.  Otherwise z += w D^-1 r1 (block Jacobi).
// With unit_step the correction is delta itself, accepted when it lowers the
// energy norm: delta.A delta - 2 delta.r1 < 0.
__global__ void k_nn_select(int64_t count, int nf, const double* s, int slot_s1, int slot_s2,
                            const double* __restrict__ delta, const double* __restrict__ r1,
                            const double* __restrict__ invD, double omega, int unit_step, double* __restrict__ z,
                            int* __restrict__ counters) {
    const double s1 = s[slot_s1], s2 = s[slot_s2];
    const bool use = unit_step ? (s2 - 2.0 * s1 < 0) : (s1 > 0 && s2 > 0);
    const double alpha = unit_step ? 1.0 : s1 / s2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) counters[use ? 0 : 1] += 1;
    if (t >= count) return;
    if (use) {
        z[t] += alpha * delta[t];
    } else {
        const int64_t i = t / nf;
        const int a = (int)(t - i * nf);
        const double* D = invD + (i * nf + a) * nf;
        const double* r = r1 + i * nf;
        double v = 0;
        for (int b = 0; b < nf; b++) v += D[b] * r[b];
        z[t] += omega * v;
    }
}

inline int g1(int64_t n, int bs = 256) { return (int)((n + bs - 1) / bs); }

}  // namespace

// ------------------------------------------------------------ launchers
void nn_prep(Ctx& c, int n, int nf, const double* r1, const float* invD, float* v) {
    hipLaunchKernelGGL(k_nn_prep, dim3(g1((int64_t)n * nf)), dim3(256), 0, c.stream, n, nf, r1, invD, v);
    FH_HIP(hipGetLastError());
}
void nn_input(Ctx& c, int n, int nf, const float* v, const float* feat, int slot_vv, int slot_sigma, float* u,
              float* x0) {
    hipLaunchKernelGGL(k_nn_input, dim3(g1((int64_t)n * (nf + 2))), dim3(256), 0, c.stream, n, nf, v, feat,
                       c.d_scalars, slot_vv, slot_sigma, u, x0);
    FH_HIP(hipGetLastError());
}
void nn_encode(Ctx& c, int n, int din, const float* x, const void* We, const void* be, bool bf16, float* h) {
    const int g = (n + kNodesPerWG - 1) / kNodesPerWG;
    if (bf16)
        hipLaunchKernelGGL(k_nn_encode<bf16w>, dim3(g), dim3(256), 0, c.stream, n, din, x, (const bf16w*)We,
                           (const bf16w*)be, h);
    else
        hipLaunchKernelGGL(k_nn_encode<float>, dim3(g), dim3(256), 0, c.stream, n, din, x, (const float*)We,
                           (const float*)be, h);
    FH_HIP(hipGetLastError());
}
void nn_layer(Ctx& c, int n, const int* rp, const int* ci, const float* w, const float* hin, const void* S,
              const void* N, const void* cb, bool bf16, float* hout, float* m_out) {
    const int g = (n + kNodesPerWG - 1) / kNodesPerWG;
    if (bf16)
        hipLaunchKernelGGL(k_nn_layer<bf16w>, dim3(g), dim3(256), 0, c.stream, n, rp, ci, w, hin, (const bf16w*)S,
                           (const bf16w*)N, (const bf16w*)cb, hout, m_out);
    else
        hipLaunchKernelGGL(k_nn_layer<float>, dim3(g), dim3(256), 0, c.stream, n, rp, ci, w, hin, (const float*)S,
                           (const float*)N, (const float*)cb, hout, m_out);
    FH_HIP(hipGetLastError());
}
void nn_out(Ctx& c, int n, int nf, const float* h, const float* u, const void* Wo, const void* bo, bool bf16,
            int slot_sigma, double* delta) {
    const int g = g1((int64_t)n * nf);
    if (bf16)
        hipLaunchKernelGGL(k_nn_out<bf16w>, dim3(g), dim3(256), 0, c.stream, n, nf, h, u, (const bf16w*)Wo,
                           (const bf16w*)bo, c.d_scalars, slot_sigma, delta);
    else
        hipLaunchKernelGGL(k_nn_out<float>, dim3(g), dim3(256), 0, c.stream, n, nf, h, u, (const float*)Wo,
                           (const float*)bo, c.d_scalars, slot_sigma, delta);
    FH_HIP(hipGetLastError());
}
void nn_out_back(Ctx& c, int n, int nf, const double* g, int slot_sigma, const float* Wo, float* gy, float* dh) {
    hipLaunchKernelGGL(k_nn_out_back, dim3((n + kNodesPerWG - 1) / kNodesPerWG), dim3(256), 0, c.stream, n, nf, g,
                       c.d_scalars, slot_sigma, Wo, gy, dh);
    FH_HIP(hipGetLastError());
}
void nn_relu_back(Ctx& c, int64_t count, const float* h, float* dh) {
    hipLaunchKernelGGL(k_nn_relu_back, dim3(g1(count)), dim3(256), 0, c.stream, count, h, dh);
    FH_HIP(hipGetLastError());
}
void nn_back_self(Ctx& c, int n, const float* dpre, const float* S, const float* N, float* dprev, float* tvec) {
    hipLaunchKernelGGL(k_nn_back_self, dim3((n + kNodesPerWG - 1) / kNodesPerWG), dim3(256), 0, c.stream, n, dpre, S,
                       N, dprev, tvec);
    FH_HIP(hipGetLastError());
}
void nn_back_gather(Ctx& c, int n, const int* rp, const int* ci, const float* wT, const float* tvec, float* dprev) {
    hipLaunchKernelGGL(k_nn_back_gather, dim3((n + kNodesPerWG - 1) / kNodesPerWG), dim3(256), 0, c.stream, n, rp,
                       ci, wT, tvec, dprev);
    FH_HIP(hipGetLastError());
}
void nn_outer(Ctx& c, int n, int da, int db, const float* A, const float* B, double* partials, float* grad) {
    FH_CHECK(da * db <= 1024, -3, "outer-product gradient larger than 1024 entries");
    hipLaunchKernelGGL(k_nn_outer, dim3(kOuterWGs), dim3(256), 0, c.stream, n, da, db, A, B, partials);
    FH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_nn_reduce, dim3(g1(da * db)), dim3(256), 0, c.stream, da * db, kOuterWGs, partials, grad);
    FH_HIP(hipGetLastError());
}
size_t nn_partials_bytes() { return sizeof(double) * kOuterWGs * 1024; }
void nn_adam(Ctx& c, int np, float* w, const float* g, float* m1, float* m2, float lr, float b1, float b2, float eps,
             int step, void* wb) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_nn_adam, dim3(g1(np)), dim3(256), 0, c.stream, np, w, g, m1, m2, lr, b1, b2, eps, c1, c2,
                       (bf16w*)wb);
    FH_HIP(hipGetLastError());
}
void nn_to_bf16(Ctx& c, int np, const float* w, void* wb) {
    hipLaunchKernelGGL(k_nn_to_bf16, dim3(g1(np)), dim3(256), 0, c.stream, np, w, (bf16w*)wb);
    FH_HIP(hipGetLastError());
}
void nn_edge_weights(Ctx& c, const Bcsr& A, float* w) {
    FH_CHECK(A.prec == Prec::F64, -3, "edge weights need the FP64 operator");
    hipLaunchKernelGGL(k_nn_edge_weights, dim3(g1(A.nbr)), dim3(256), 0, c.stream, A.nbr, A.bs_r, A.rowptr.as<int>(),
                       A.colidx.as<int>(), A.vals.as<double>(), w);
    FH_HIP(hipGetLastError());
}
void nn_gather_index(Ctx& c, int64_t nnzb, const int* idx, const float* src, float* dst) {
    hipLaunchKernelGGL(k_nn_gather_index, dim3(g1(nnzb)), dim3(256), 0, c.stream, nnzb, idx, src, dst);
    FH_HIP(hipGetLastError());
}
void nn_select(Ctx& c, int64_t count, int nf, int slot_s1, int slot_s2, const double* delta, const double* r1,
               const double* invD, double omega, bool unit_step, double* z, int* counters) {
    hipLaunchKernelGGL(k_nn_select, dim3(g1(count)), dim3(256), 0, c.stream, count, nf, c.d_scalars, slot_s1, slot_s2,
                       delta, r1, invD, omega, (int)unit_step, z, counters);
    FH_HIP(hipGetLastError());
}

}  // namespace fh
