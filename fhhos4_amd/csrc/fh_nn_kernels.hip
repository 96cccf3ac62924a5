// gfx950 kernels of the learned post-smoother: a message-passing network over
// the face graph of the condensed operator (one node per block row, edges =
// off-diagonal blocks), hidden width 32.
//
// Layout: node features are row-major n x 32 FP32 arrays (one 128-byte row
// per face), so the 32 lanes of a node read or write one row with one
// coalesced access.  A workgroup of 256 threads = 8 nodes x 32 channels; every
// lane computes one output channel.  The message-passing layer keeps its
// lane's weight rows in registers and strides over node batches staged in
// LDS; the smaller kernels read the weights from the cache.  Weights are FP32
// for training and BF16 (FP32 accumulate) for inference.
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

// Encoder: h = relu(We x + be), x of width din <= kMaxDin.  Lane c keeps row c
// of We in registers; the grid strides over nodes.
constexpr int kMaxDin = 12;
constexpr int kEncodeMaxWGs = 8192;

template <class TW>
__global__ __launch_bounds__(256) void k_nn_encode(int n, int din, const float* __restrict__ x,
                                                    const TW* __restrict__ We, const TW* __restrict__ be,
                                                    float* __restrict__ h) {
    __shared__ float sWe[H * kMaxDin + H];
    for (int e = threadIdx.x; e < H * din; e += blockDim.x) sWe[e] = wload(We + e);
    if (threadIdx.x < H) sWe[H * kMaxDin + threadIdx.x] = wload(be + threadIdx.x);
    __syncthreads();
    const int c = threadIdx.x & (H - 1);
    float wr[kMaxDin];
#pragma unroll
    for (int k = 0; k < kMaxDin; k++) wr[k] = k < din ? sWe[c * din + k] : 0.f;
    const float bias = sWe[H * kMaxDin + c];
    for (int64_t i = (int64_t)blockIdx.x * kNodesPerWG + (threadIdx.x >> 5); i < n;
         i += (int64_t)gridDim.x * kNodesPerWG) {
        float acc = bias;
        const float* xi = x + i * din;
#pragma unroll
        for (int k = 0; k < kMaxDin; k++)
            if (k < din) acc += wr[k] * xi[k];
        h[i * H + c] = acc > 0.f ? acc : 0.f;
    }
}

// Message-passing layer: m_i = sum_j w_ij h_j ; h'_i = relu(S h_i + N m_i + c).
// m is stored for the backward pass when m_out != nullptr.
//
// Lane c of a 32-lane half-wave owns output channel c and keeps row c of S
// and of N in registers for the whole launch (staged once per workgroup, not
// once per node).  The grid strides over batches of kLayerBatch nodes: every
// thread gathers kLayerBatch / 8 nodes' (h_i, m_i) rows into LDS, then
// computes their channel c from LDS broadcasts (16-byte reads; the row pitch
// of 2H + 4 floats puts the two half-waves' rows in different banks).  The
// batch loop bound depends on blockIdx only, so both barriers are uniform.
constexpr int kLayerBatch = 32;            // nodes per workgroup iteration
constexpr int kLayerPer = kLayerBatch / 8; // nodes per thread per iteration
constexpr int kLayerPitch = 2 * H + 4;
constexpr int kLayerMaxWGs = 768;  // 3 workgroups of 4 waves per CU (164 VGPRs: 3 waves per SIMD)

template <class TW>
__global__ __launch_bounds__(256) void k_nn_layer(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const float* __restrict__ w, const float* __restrict__ hin,
                                                   const TW* __restrict__ S, const TW* __restrict__ N,
                                                   const TW* __restrict__ cb, float* __restrict__ hout,
                                                   float* __restrict__ m_out) {
    __shared__ __align__(16) float sx[kLayerBatch][kLayerPitch];
    __shared__ float sW[2 * H * H];  // S^T then N^T: sW[k H + c] = S[c][k]
    // Weights reach the registers through LDS: one coalesced read per
    // workgroup (every lane of the grid reading the same few lines from L2
    // would serialise on their channels).
    for (int e = threadIdx.x; e < H * H; e += blockDim.x) {
        const int row = e / H, k = e - row * H;
        sW[k * H + row] = wload(S + e);
        sW[H * H + k * H + row] = wload(N + e);
    }
    __syncthreads();
    const int c = threadIdx.x & (H - 1), q = threadIdx.x >> 5;
    float ws[H], wn[H];
#pragma unroll
    for (int k = 0; k < H; k++) {
        ws[k] = sW[k * H + c];
        wn[k] = sW[H * H + k * H + c];
    }
    const float bias = wload(cb + c);
    for (int64_t base = (int64_t)blockIdx.x * kLayerBatch; base < n; base += (int64_t)gridDim.x * kLayerBatch) {
        // gather.  Lane c of the half-wave fetches edge c of each of its
        // kLayerPer rows (one coalesced load for the row's columns and
        // weights); edge e is then broadcast from lane e, and the neighbour
        // rows of four edge positions x kLayerPer nodes are loaded before
        // they are summed (edge order kept per node).  Rows longer than 32
        // edges read the rest directly.
        int beg[kLayerPer], deg[kLayerPer], dmax = 0;
        float m[kLayerPer], hs[kLayerPer];
#pragma unroll
        for (int r = 0; r < kLayerPer; r++) {
            const int64_t i = base + r * 8 + q;
            beg[r] = 0;
            deg[r] = 0;
            hs[r] = 0.f;
            m[r] = 0.f;
            if (i < n) {
                beg[r] = rp[i];
                deg[r] = rp[i + 1] - beg[r];
                hs[r] = hin[i * H + c];
            }
            dmax = deg[r] > dmax ? deg[r] : dmax;
        }
        dmax = max(dmax, __shfl_xor(dmax, 32));  // wave-uniform trip count
        int colv[kLayerPer];
        float wv[kLayerPer];
#pragma unroll
        for (int r = 0; r < kLayerPer; r++) {
            colv[r] = c < deg[r] ? ci[beg[r] + c] : 0;
            wv[r] = c < deg[r] ? w[beg[r] + c] : 0.f;
        }
        for (int e0 = 0; e0 < dmax; e0 += 4) {
            float v[4][kLayerPer], wt[4][kLayerPer];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = e0 + u;
#pragma unroll
                for (int r = 0; r < kLayerPer; r++) {
                    int col;
                    float we;
                    if (e < 32) {
                        col = __shfl(colv[r], e, 32);
                        we = __shfl(wv[r], e, 32);
                    } else {
                        col = e < deg[r] ? ci[beg[r] + e] : 0;
                        we = e < deg[r] ? w[beg[r] + e] : 0.f;
                    }
                    wt[u][r] = we;
                    v[u][r] = e < deg[r] ? hin[(int64_t)col * H + c] : 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int r = 0; r < kLayerPer; r++)
                    if (e0 + u < deg[r]) m[r] += wt[u][r] * v[u][r];
        }
#pragma unroll
        for (int r = 0; r < kLayerPer; r++) {
            const int slot = r * 8 + q;
            const int64_t i = base + slot;
            if (m_out && i < n) m_out[i * H + c] = m[r];
            sx[slot][c] = hs[r];
            sx[slot][H + c] = m[r];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kLayerPer; r++) {
            const int slot = r * 8 + q;
            const int64_t i = base + slot;
            const float4* x4 = reinterpret_cast<const float4*>(&sx[slot][0]);
            float acc = bias;
#pragma unroll
            for (int k4 = 0; k4 < H / 4; k4++) {
                const float4 hv = x4[k4], mv = x4[H / 4 + k4];
                acc += ws[4 * k4 + 0] * hv.x + wn[4 * k4 + 0] * mv.x;
                acc += ws[4 * k4 + 1] * hv.y + wn[4 * k4 + 1] * mv.y;
                acc += ws[4 * k4 + 2] * hv.z + wn[4 * k4 + 2] * mv.z;
                acc += ws[4 * k4 + 3] * hv.w + wn[4 * k4 + 3] * mv.w;
            }
            if (i < n) hout[i * H + c] = acc > 0.f ? acc : 0.f;
        }
        __syncthreads();
    }
}

// Inference layer (BF16 weights): the same layer with the dense part on the
// matrix cores.  A wave owns tiles of 32 consecutive nodes.
//  1. The tile's row pointers come in one load, its edge list (contiguous in
//     the CSR arrays) in a few coalesced loads, into the wave's LDS.
//  2. Each half-wave walks the edges of its 16 nodes in order, kMChunk rows
//     in flight (lane c = channel c), keeps a running sum per node and writes
//     (h_i, m_i) to the tile in LDS as BF16.  Tiles with more than kMEdges
//     edges walk their nodes one by one instead.
//  3. Four v_mfma_f32_32x32x16_bf16 multiply the 32 x 64 tile by the 64 x 32
//     matrix [S N]^T (FP32 accumulate).  A lane supplies row lane%32 of the
//     tile and column lane%32 of the weights with the same k elements, and
//     receives column lane%32 of the product: rows 8 (r / 4) + 4 (lane / 32)
//     + r % 4 of its 16 accumulators.
// Activations are rounded to BF16 (round to nearest even) for the product;
// the neighbour sum m_i is accumulated in FP32 in edge order.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int kMTile = 32;
constexpr int kMWaves = 4;
constexpr int kMBatch = kMTile * kMWaves;  // nodes per workgroup iteration
constexpr int kMPitch = 2 * H + 8;         // BF16 elements per tile row (144 bytes)
constexpr int kMEdges = 384;               // staged edges per tile
constexpr int kMChunk = 16;                // neighbour rows in flight per half-wave
constexpr int kMMaxWGs = 1024;             // 4 workgroups per CU

__device__ __forceinline__ __bf16 to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __builtin_bit_cast(__bf16, (uint16_t)(u >> 16));
}

__global__ __launch_bounds__(256) void k_nn_layer_mfma(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                        const float* __restrict__ w, const float* __restrict__ hin,
                                                        const bf16w* __restrict__ S, const bf16w* __restrict__ N,
                                                        const bf16w* __restrict__ cb, float* __restrict__ hout) {
    __shared__ __align__(16) __bf16 tile[kMWaves][kMTile][kMPitch];
    __shared__ uint16_t sB[2 * H * H];  // sB[k H + j] = [S N]^T[k][j]
    __shared__ int s_col[kMWaves][kMEdges];
    __shared__ float s_w[kMWaves][kMEdges];
    __shared__ uint8_t s_node[kMWaves][kMEdges];
    __shared__ int s_row[kMWaves][kMTile + 1];
    for (int e = threadIdx.x; e < H * H; e += blockDim.x) {
        const int row = e / H, k = e - row * H;
        sB[k * H + row] = S[e].u;
        sB[(H + k) * H + row] = N[e].u;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, half = lane >> 5, c = lane & 31;
    bf16x8_t bop[4];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int u = 0; u < 8; u++) bop[t][u] = __builtin_bit_cast(__bf16, sB[(16 * t + 8 * half + u) * H + c]);
    const float bias = wload(cb + c);
    // XCD-aware batch order: workgroups are dealt to the 8 XCDs round robin,
    // so XCD x (blockIdx % 8) walks the x-th contiguous slab of batches and
    // the neighbour rows a tile gathers were mostly fetched into the same
    // XCD's L2 by the tiles next to it.
    const int64_t nbat = ((int64_t)n + kMBatch - 1) / kMBatch;
    int64_t first = blockIdx.x, end = nbat, step = gridDim.x;
    if (gridDim.x % 8 == 0) {
        const int64_t slab = (nbat + 7) / 8, xcd = blockIdx.x % 8;
        first = xcd * slab + blockIdx.x / 8;
        end = min(nbat, (xcd + 1) * slab);
        step = gridDim.x / 8;
    }
    for (int64_t b = first; b < end; b += step) {
        const int64_t tb = b * kMBatch + (int64_t)wv * kMTile;
        const int nt = (int)max<int64_t>(0, min<int64_t>(kMTile, (int64_t)n - tb));  // nodes of this tile
        // 1. row pointers and edge list
        const int rpv = lane <= nt && nt > 0 ? rp[tb + lane] : 0;
        const int e_base = __shfl(rpv, 0);
        const int ne = __shfl(rpv, nt) - e_base;
        const bool staged = ne <= kMEdges;
        if (lane <= nt) s_row[wv][lane] = rpv - e_base;
        if (staged) {
            for (int e = lane; e < ne; e += 64) {
                s_col[wv][e] = ci[e_base + e];
                s_w[wv][e] = w[e_base + e];
            }
            const int rnext = __shfl(rpv, lane + 1 < 64 ? lane + 1 : 63);
            if (lane < nt)
                for (int e = rpv - e_base; e < rnext - e_base; e++) s_node[wv][e] = (uint8_t)lane;
        }
        __syncthreads();
        // 2. (h_i, m_i) rows of the tile
        const int r0 = min(16 * half, nt), r1 = min(16 * half + 16, nt);
        {
            float hs[16];
#pragma unroll
            for (int r = 0; r < 16; r++) hs[r] = r0 + r < r1 ? hin[(tb + r0 + r) * H + c] : 0.f;
#pragma unroll
            for (int r = 0; r < 16; r++)
                if (r0 + r < r1) {
                    tile[wv][r0 + r][c] = to_bf16(hs[r]);
                    tile[wv][r0 + r][H + c] = to_bf16(0.f);
                }
        }
        if (staged) {
            const int es = r0 < r1 ? s_row[wv][r0] : 0, ee = r0 < r1 ? s_row[wv][r1] : 0;
            int cur = r0;
            float run = 0.f;
            for (int e0 = es; e0 < ee; e0 += kMChunk) {
                float v[kMChunk];
#pragma unroll
                for (int u = 0; u < kMChunk; u++) {
                    const int e = e0 + u;
                    v[u] = e < ee ? hin[(int64_t)s_col[wv][e] * H + c] : 0.f;
                }
#pragma unroll
                for (int u = 0; u < kMChunk; u++) {
                    const int e = e0 + u;
                    if (e < ee) {
                        const int nd = s_node[wv][e];
                        if (nd != cur) {
                            tile[wv][cur][H + c] = to_bf16(run);
                            run = 0.f;
                            cur = nd;
                        }
                        run += s_w[wv][e] * v[u];
                    }
                }
            }
            if (ee > es) tile[wv][cur][H + c] = to_bf16(run);
        } else {
            for (int r = r0; r < r1; r++) {
                const int64_t i = tb + r;
                float m = 0.f;
                for (int p = rp[i]; p < rp[i + 1]; p++) m += w[p] * hin[(int64_t)ci[p] * H + c];
                tile[wv][r][H + c] = to_bf16(m);
            }
        }
        __syncthreads();
        // 3. dense part on the matrix cores (rows >= nt hold stale values and are not stored)
        floatx16 acc;
#pragma unroll
        for (int r = 0; r < 16; r++) acc[r] = 0.f;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&tile[wv][c][16 * t + 8 * half]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bop[t], acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int row = 8 * (r / 4) + 4 * half + (r % 4);
            const float y = acc[r] + bias;
            if (row < nt) hout[(tb + row) * H + c] = y > 0.f ? y : 0.f;
        }
        __syncthreads();
    }
}

// Output: delta_i = sigma (u_i + Wo h_i + bo), thread per unknown (nf <= kMaxNf).
constexpr int kMaxNf = 10;
template <class TW>
__global__ void k_nn_out(int n, int nf, const float* __restrict__ h, const float* __restrict__ u,
                         const TW* __restrict__ Wo, const TW* __restrict__ bo, const double* s, int slot_sigma,
                         double* __restrict__ delta) {
    __shared__ float sWo[kMaxNf * H + kMaxNf];
    for (int e = threadIdx.x; e < nf * H; e += blockDim.x) sWo[e] = wload(Wo + e);
    if (threadIdx.x < nf) sWo[kMaxNf * H + threadIdx.x] = wload(bo + threadIdx.x);
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * nf) return;
    const int i = (int)(t / nf), a = (int)(t - (int64_t)i * nf);
    const float* hi = h + (int64_t)i * H;
    float y = sWo[kMaxNf * H + a];
#pragma unroll 8
    for (int k = 0; k < H; k++) y += sWo[a * H + k] * hi[k];
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
                            int* __restrict__ counters, const double* __restrict__ r,
                            const double* __restrict__ Adelta, double* __restrict__ Az) {
    const double s1 = s[slot_s1], s2 = s[slot_s2];
    const bool use = unit_step ? (s2 - 2.0 * s1 < 0) : (s1 > 0 && s2 > 0);
    const double alpha = unit_step ? 1.0 : s1 / s2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) counters[use ? 0 : 1] += 1;
    if (t >= count) return;
    if (use) {
        z[t] += alpha * delta[t];
        // A z = A z0 + alpha A delta with A z0 = r - r1 (no product)
        if (Az) Az[t] = (r[t] - r1[t]) + alpha * Adelta[t];
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
    FH_CHECK(din >= 1 && din <= kMaxDin, -3, "encoder input width out of range");
    if (n <= 0) return;
    const int g = (int)std::min<int64_t>(((int64_t)n + kNodesPerWG - 1) / kNodesPerWG, kEncodeMaxWGs);
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
    const int g = (int)std::min<int64_t>(((int64_t)n + kLayerBatch - 1) / kLayerBatch, kLayerMaxWGs);
    if (n <= 0) return;
    if (bf16) {
        FH_CHECK(m_out == nullptr, -3, "the BF16 layer is inference only (no stored aggregate)");
        int gm = (int)std::min<int64_t>(((int64_t)n + kMBatch - 1) / kMBatch, kMMaxWGs);
        if (gm >= 64) gm -= gm % 8;  // a multiple of the XCD count selects the slab order
        hipLaunchKernelGGL(k_nn_layer_mfma, dim3(gm), dim3(256), 0, c.stream, n, rp, ci, w, hin, (const bf16w*)S,
                           (const bf16w*)N, (const bf16w*)cb, hout);
    } else
        hipLaunchKernelGGL(k_nn_layer<float>, dim3(g), dim3(256), 0, c.stream, n, rp, ci, w, hin, (const float*)S,
                           (const float*)N, (const float*)cb, hout, m_out);
    FH_HIP(hipGetLastError());
}
void nn_out(Ctx& c, int n, int nf, const float* h, const float* u, const void* Wo, const void* bo, bool bf16,
            int slot_sigma, double* delta) {
    FH_CHECK(nf >= 1 && nf <= kMaxNf, -3, "face block size out of range for the output layer");
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
               const double* invD, double omega, bool unit_step, double* z, int* counters, const double* r,
               const double* Adelta, double* Az) {
    hipLaunchKernelGGL(k_nn_select, dim3(g1(count)), dim3(256), 0, c.stream, count, nf, c.d_scalars, slot_s1, slot_s2,
                       delta, r1, invD, omega, (int)unit_step, z, counters, r, Adelta, Az);
    FH_HIP(hipGetLastError());
}

}  // namespace fh
