// gfx950 kernels of the condensed-HHO solve path.
//
// All kernels are written for CDNA wave64.  The block-row kernel is the
// workhorse.  Values are stored block row after block row, so the blocks of
// one row are one contiguous span of HBM.  For even n_f (2, 4, 6, 10), 16
// lanes stream a row's span in contiguous 16-byte chunks (k_bcsr_grp, four
// rows per wavefront).  For odd n_f, one wavefront takes a row with
// lane = (block slot, row in block) (k_bcsr).  Partial sums are reduced with
// cross-lane shuffles.  Workgroup ids are remapped so that each XCD walks a
// contiguous range of block rows (better reuse of x in that XCD's L2).
//
// Full-matrix products on operators whose rows fit one wavefront (the fine
// levels: 7 blocks per row on tetrahedra) use a sliced copy of the values
// instead (k_sliced): lane = (row, block slot, column pair), so every lane
// gathers its two x entries once and every load instruction of the
// wavefront reads one contiguous run of values.  Row sums are reduced
// through LDS.  Operators with uneven rows (the prolongations) get the
// sliced copy in two parts, short and long rows, each with a row map;
// operators with rows wider than a wavefront (the Galerkin coarse levels)
// get a chunked copy with one wavefront per row (k_wide).  The smoothing
// modes include a fused Chebyshev step (OP_CHEB).
#include "fh_kernels.h"

#include <atomic>
#include <mutex>

#include <algorithm>
#include <type_traits>
#include <cstdlib>

namespace fh {
namespace {

constexpr int kWG = 256;          // 4 wavefronts per workgroup
constexpr int kDotBlocks = 1024;  // fixed reduction grid: deterministic dots

struct bf16_t {
    uint16_t u;
};

__device__ __forceinline__ double ldv(const double* p) { return *p; }
__device__ __forceinline__ float ldv(const float* p) { return *p; }
__device__ __forceinline__ float ldv(const bf16_t* p) { return __uint_as_float(((uint32_t)p->u) << 16); }

__device__ __forceinline__ void stv(double* p, double v) { *p = v; }
__device__ __forceinline__ void stv(float* p, double v) { *p = (float)v; }
__device__ __forceinline__ void stv(bf16_t* p, double v) {
    // round to nearest even
    float f = (float)v;
    uint32_t u = __float_as_uint(f);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    p->u = (uint16_t)(u >> 16);
}

template <class TV> struct Acc { using T = double; };
template <> struct Acc<float> { using T = float; };

// Workgroup id -> logical id such that XCD x (= bid % 8) owns a contiguous
// range of logical ids.
__device__ __forceinline__ int xcd_remap(int bid, int nwg) {
    int xcd = bid & 7, idx = bid >> 3;
    int per = nwg >> 3, rem = nwg & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// Smoothing output of row entry e with z = D^-1 (b - A x): y = x + omega z
// (OP_SMOOTH), or the Chebyshev step d = alpha d + beta z, y = x + d with d
// in aux (OP_CHEB; d is not read when alpha = 0).
template <int MODE, class TV, class AT>
__device__ __forceinline__ void smooth_out(int64_t e, AT z, const TV* x, TV* y, TV* aux, double alpha, double beta,
                                           double omega) {
    if constexpr (MODE == OP_CHEB) {
        const AT dv = (alpha != 0.0 ? (AT)alpha * (AT)aux[e] : (AT)0) + (AT)beta * z;
        aux[e] = (TV)dv;
        y[e] = (TV)((AT)x[e] + dv);
    } else {
        y[e] = (TV)((AT)x[e] + (AT)omega * z);
    }
}

template <int NF, class TA, class TV, class TD, int MODE>
__global__ __launch_bounds__(kWG) void k_bcsr(int nbr, const int* __restrict__ rp, const int* __restrict__ ci,
                                               const TA* __restrict__ val, const TV* x, const TV* __restrict__ b,
                                               TV* y, double alpha, double beta, double omega,
                                               const TD* __restrict__ invD, const int* __restrict__ rows,
                                               int nrows, TV* __restrict__ aux) {
    using AT = typename Acc<TV>::T;
    constexpr int SLOTS = 64 / NF;
    constexpr int ACT = SLOTS * NF;
    const int lane = threadIdx.x & 63;
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const int w = wg * (kWG / 64) + (threadIdx.x >> 6);
    const int count = rows ? nrows : nbr;
    if (w >= count) return;  // wave-uniform
    const int br = rows ? rows[w] : w;
    const int slot = lane / NF;
    const int i = lane - slot * NF;
    AT acc = 0;
    if (lane < ACT) {
        const int s = rp[br], e = rp[br + 1];
        for (int p = s + slot; p < e; p += SLOTS) {
            const int c = ci[p];
            const TA* a = val + ((int64_t)p * NF + i) * NF;
            const TV* xs = x + (int64_t)c * NF;
#pragma unroll
            for (int j = 0; j < NF; j++) acc += (AT)ldv(a + j) * (AT)xs[j];
        }
    }
#pragma unroll
    for (int off = NF; off < ACT; off <<= 1) {
        AT o = __shfl_down(acc, off, 64);
        acc += (lane + off < 64) ? o : (AT)0;
    }
    const int64_t row = (int64_t)br * NF + i;
    if (MODE == OP_SPMV) {
        if (lane < NF) {
            AT v = (AT)alpha * acc;
            if (beta != 0.0) v += (AT)beta * (AT)y[row];
            y[row] = (TV)v;
        }
    } else if (MODE == OP_RESID) {
        if (lane < NF) y[row] = (TV)((AT)b[row] - acc);
    } else {
        AT r = (lane < NF) ? (AT)b[row] - acc : (AT)0;
        AT z = 0;
        const TD* D = invD + ((int64_t)br * NF + (lane < NF ? i : 0)) * NF;
#pragma unroll
        for (int j = 0; j < NF; j++) {
            AT rj = __shfl(r, j, 64);
            if (lane < NF) z += (AT)ldv(D + j) * rj;
        }
        if (lane < NF) smooth_out<MODE>(row, z, x, y, aux, alpha, beta, omega);
    }
}

// Scalar rows (n_f = 1): a group of G lanes per row (64 / G rows per
// wavefront), the lanes of a group striding over the row's entries, the
// partial sums reduced with xor-shuffles inside the group.  For the
// aggregation levels, whose rows hold 10-60 entries: one wavefront per row
// (k_bcsr<1>) leaves most lanes idle and launches a wavefront per row.
template <int G, class TA, class TV, class TD, int MODE>
__global__ __launch_bounds__(kWG) void k_csr_grp(int nbr, const int* __restrict__ rp, const int* __restrict__ ci,
                                                  const TA* __restrict__ val, const TV* x, const TV* __restrict__ b,
                                                  TV* y, double alpha, double beta, double omega,
                                                  const TD* __restrict__ invD, const int* __restrict__ rows,
                                                  int nrows, TV* __restrict__ aux) {
    using AT = typename Acc<TV>::T;
    const int lane = threadIdx.x & 63;
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const int w = wg * (kWG / 64) + (threadIdx.x >> 6);
    const int count = rows ? nrows : nbr;
    if (w * (64 / G) >= count) return;  // wave-uniform
    const int g = lane / G, gl = lane - g * G;
    const int idx = w * (64 / G) + g;
    AT acc = 0;
    int br = 0;
    if (idx < count) {
        br = rows ? rows[idx] : idx;
        const int s = rp[br], e = rp[br + 1];
        for (int p = s + gl; p < e; p += G) acc += (AT)ldv(val + p) * (AT)x[ci[p]];
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (idx >= count || gl != 0) return;
    const int64_t row = br;
    if (MODE == OP_SPMV) {
        AT v = (AT)alpha * acc;
        if (beta != 0.0) v += (AT)beta * (AT)y[row];
        y[row] = (TV)v;
    } else if (MODE == OP_RESID) {
        y[row] = (TV)((AT)b[row] - acc);
    } else {
        const AT z = (AT)ldv(invD + row) * ((AT)b[row] - acc);
        smooth_out<MODE>(row, z, x, y, aux, alpha, beta, omega);
    }
}

// Two consecutive values (same block row of the block: NF is even).
__device__ __forceinline__ void ld2(const double* p, double& a, double& b) {
    const double2 v = *(const double2*)p;
    a = v.x; b = v.y;
}
__device__ __forceinline__ void ld2(const float* p, float& a, float& b) {
    const float2 v = *(const float2*)p;
    a = v.x; b = v.y;
}
__device__ __forceinline__ void ld2(const bf16_t* p, float& a, float& b) {
    const uint32_t u = *(const uint32_t*)p;
    a = __uint_as_float(u << 16);
    b = __uint_as_float(u & 0xffff0000u);
}

// Block-row kernel for even NF with 16 lanes per block row (4 rows per
// wavefront).  The blocks of a row are one contiguous span of NF*NF*nblk
// values; lane g of the group streams the 2-value chunks g, g+16, g+32, ...
// of that span, so every load instruction of the wavefront reads four
// contiguous 16-lane segments.  A chunk (q) lies in block p = q / (NF*NF/2),
// row i, columns j and j+1 of that block; the lane keeps one partial sum per
// block row i and the group reduces them with xor-shuffles.
template <int NF, class TA, class TV, class TD, int MODE>
__global__ __launch_bounds__(kWG) void k_bcsr_grp(int nbr, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const TA* __restrict__ val, const TV* x,
                                                   const TV* __restrict__ b, TV* y, double alpha, double beta,
                                                   double omega, const TD* __restrict__ invD,
                                                   const int* __restrict__ rows, int nrows, TV* __restrict__ aux) {
    static_assert(NF % 2 == 0, "even block size");
    using AT = typename Acc<TV>::T;
    constexpr int G = 16, U = 4;
    constexpr int PAIRS = NF * NF / 2;
    const int wg = xcd_remap(blockIdx.x, gridDim.x);
    const int t = wg * kWG + (int)threadIdx.x;
    const int w = t / G, g = t & (G - 1);
    const int count = rows ? nrows : nbr;
    if ((t & ~63) / G >= count) return;  // whole wavefront past the end (wave-uniform)
    const bool active = w < count;
    const int br = active ? (rows ? rows[w] : w) : 0;
    const int s = active ? rp[br] : 0;
    const int npairs = active ? (rp[br + 1] - s) * PAIRS : 0;
    const TA* vb = val + (int64_t)s * (2 * PAIRS);
    AT acc[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) acc[k] = 0;
    for (int q0 = 0; q0 < npairs; q0 += G * U) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int q = q0 + u * G + g;
            const bool ok = q < npairs;
            const int qq = ok ? q : npairs - 1;
            const int p = qq / PAIRS;
            const int r = (qq - p * PAIRS) * 2;
            const int i = r / NF, j = r - i * NF;
            typename Acc<TV>::T a0, a1;
            {
                decltype(ldv(vb)) v0, v1;
                ld2(vb + (int64_t)qq * 2, v0, v1);
                a0 = (AT)v0; a1 = (AT)v1;
            }
            const int c = ci[s + p];
            TV x0, x1;
            ld2(x + (int64_t)c * NF + j, x0, x1);
            const AT prod = ok ? a0 * (AT)x0 + a1 * (AT)x1 : (AT)0;
#pragma unroll
            for (int k = 0; k < NF; k++) acc[k] += (i == k) ? prod : (AT)0;
        }
    }
#pragma unroll
    for (int k = 0; k < NF; k++) {
#pragma unroll
        for (int m = G / 2; m > 0; m >>= 1) acc[k] += __shfl_xor(acc[k], m, 64);
    }
    if (!active || g >= NF) return;
    AT mine = acc[0];
#pragma unroll
    for (int k = 1; k < NF; k++) mine = (g == k) ? acc[k] : mine;
    const int64_t row = (int64_t)br * NF + g;
    if (MODE == OP_SPMV) {
        AT v = (AT)alpha * mine;
        if (beta != 0.0) v += (AT)beta * (AT)y[row];
        y[row] = (TV)v;
    } else if (MODE == OP_RESID) {
        y[row] = (TV)((AT)b[row] - mine);
    } else {
        const TD* D = invD + row * NF;
        const TV* bb = b + (int64_t)br * NF;
        AT z = 0;
#pragma unroll
        for (int k = 0; k < NF; k++) z += (AT)ldv(D + k) * ((AT)bb[k] - acc[k]);
        smooth_out<MODE>(row, z, x, y, aux, alpha, beta, omega);
    }
}

// Non-temporal loads of the streamed matrix values (read once per product;
// keeping them out of the caches leaves L2 to the gathered x entries).
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ldnt(const double* p, double& a) { a = __builtin_nontemporal_load(p); }
__device__ __forceinline__ void ldnt(const float* p, float& a) { a = __builtin_nontemporal_load(p); }
__device__ __forceinline__ void ldnt(const bf16_t* p, float& a) {
    a = __uint_as_float(((uint32_t)__builtin_nontemporal_load((const uint16_t*)p)) << 16);
}
__device__ __forceinline__ void ldnt2(const double* p, double& a, double& b) {
    const nt_d2 v = __builtin_nontemporal_load((const nt_d2*)p);
    a = v.x; b = v.y;
}
__device__ __forceinline__ void ldnt2(const float* p, float& a, float& b) {
    const nt_f2 v = __builtin_nontemporal_load((const nt_f2*)p);
    a = v.x; b = v.y;
}
__device__ __forceinline__ void ldnt2(const bf16_t* p, float& a, float& b) {
    const uint32_t u = __builtin_nontemporal_load((const uint32_t*)p);
    a = __uint_as_float(u << 16);
    b = __uint_as_float(u & 0xffff0000u);
}

// Four BF16 values (8 bytes): block rows 2h and 2h + 1, two columns each.
__device__ __forceinline__ void ldnt4(const bf16_t* p, float& a00, float& a01, float& a10, float& a11) {
    const uint64_t u = __builtin_nontemporal_load((const uint64_t*)p);
    const uint32_t lo = (uint32_t)u, hi = (uint32_t)(u >> 32);
    a00 = __uint_as_float(lo << 16);
    a01 = __uint_as_float(lo & 0xffff0000u);
    a10 = __uint_as_float(hi << 16);
    a11 = __uint_as_float(hi & 0xffff0000u);
}
template <class T> __device__ __forceinline__ void ldnt4(const T*, float&, float&, float&, float&) {}

// Non-temporal raw buffer loads (the resource's base and row offset in
// scalar registers, the lane's byte offset in one VGPR).  Word 3 of a gfx9
// buffer resource: 32-bit data format, no swizzle.
constexpr int kBufferWord3 = 0x00020000;
constexpr int kBufferNt = 2;  // cache policy bits of the load: non-temporal
template <class T> constexpr bool kRawSliceLoads = std::is_same<T, double>::value || std::is_same<T, float>::value;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void ldbuf2(__amdgpu_buffer_rsrc_t r, int voff, int soff, double& a, double& b) {
    const u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, kBufferNt);
    a = __hiloint2double((int)u.y, (int)u.x);
    b = __hiloint2double((int)u.w, (int)u.z);
}
__device__ __forceinline__ void ldbuf2(__amdgpu_buffer_rsrc_t r, int voff, int soff, float& a, float& b) {
    const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kBufferNt);
    a = __uint_as_float(u.x);
    b = __uint_as_float(u.y);
}
__device__ __forceinline__ void ldbuf1(__amdgpu_buffer_rsrc_t r, int voff, int soff, double& a) {
    const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kBufferNt);
    a = __hiloint2double((int)u.y, (int)u.x);
}
__device__ __forceinline__ void ldbuf1(__amdgpu_buffer_rsrc_t r, int voff, int soff, float& a) {
    a = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, kBufferNt));
}
template <class T> __device__ __forceinline__ void ldbuf2(__amdgpu_buffer_rsrc_t, int, int, float&, float&) {}
template <class T> __device__ __forceinline__ void ldbuf1(__amdgpu_buffer_rsrc_t, int, int, float&) {}

constexpr int kSlicedWaves = kWG / 64;  // slices per workgroup

// Orders this wavefront's LDS accesses: the lanes of a wavefront run in
// lockstep and the LDS serves one wavefront's requests in order, so only
// the compiler has to be kept from moving accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block-row operation on the sliced layout (all rows), one slice per
// wavefront.  Lane (r, b, jp) multiplies column group jp of block slot b of
// row r by its x entries, for the NF rows i of the block; the NF partial
// sums per lane go to LDS and 64 / (RW * NF) lanes per (r, i) add the
// W * JP sums of its row.  Every wavefront works on its own LDS rows, so no
// workgroup barrier is taken.  The loads are issued in one burst: column
// indices, values, the entries of b, x, d (and y) of the slice's rows and,
// for the smoothing modes, its inverse diagonal blocks; only the gather of
// x waits, for the column indices.  The inverse blocks are spread over the
// lanes (lane (r, i, m) holds entries CW m .. CW m + CW - 1 of row i of
// block r) when RW * NF * JP <= 64, and D^-1 times the residual is then
// summed with shuffles.  MAP: the rows of this part are rowmap[0 .. nbr)
// (else 0 .. nbr).
template <int NF, class TA, class TV, class TD, int MODE, bool MAP>
__global__ __launch_bounds__(kWG) void k_sliced(int nbr, int nslices, int W, int RW, const int* __restrict__ cols,
                                                const TA* __restrict__ sv, const TV* x, const TV* b, TV* y,
                                                double alpha, double beta, double omega,
                                                const TD* __restrict__ invD, TV* __restrict__ aux,
                                                const int* __restrict__ rowmap) {
    using AT = typename Acc<TV>::T;
    constexpr int CW = NF % 2 == 0 ? 2 : 1;
    constexpr int JP = NF / CW;
    constexpr bool SMOOTHING = MODE == OP_CHEB || MODE == OP_SMOOTH;
    const int LPR = W * JP, ACT = RW * LPR;
    __shared__ AT red[kSlicedWaves][NF][64];
    // the wavefront index as a uniform value: slice offsets stay scalar
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s = xcd_remap(blockIdx.x, gridDim.x) * kSlicedWaves + wave;
    if (s >= nslices) return;  // wavefront-uniform; no workgroup barrier follows

    // product lane (rv, bs, jp)
    const int rv = lane / LPR, rl = lane - rv * LPR, bs = rl / JP, jp = rl - bs * JP;
    const bool act = lane < ACT;
    const int c = act ? cols[(int64_t)s * RW * W + rv * W + bs] : -1;
    // loaded value type: FP64 values load as doubles, FP32 and BF16 ones as
    // floats (widened to the accumulation type at use)
    using LT = std::conditional_t<std::is_same<TA, double>::value, double, float>;
    LT a[NF][CW];
    if (act) {
        // The slice's value loads: its base in a buffer resource (scalar
        // registers), the row offset scalar, the lane offset one VGPR.
        // Padding slots hold zero values.
        const TA* sbase = sv + (int64_t)s * NF * ACT * CW;
        if constexpr (kRawSliceLoads<TA>) {
            const auto rs =
                __builtin_amdgcn_make_buffer_rsrc((void*)sbase, 0, (int)(NF * ACT * CW * sizeof(TA)), kBufferWord3);
#pragma unroll
            for (int k = 0; k < NF; k++) {
                if constexpr (CW == 2)
                    ldbuf2(rs, lane * CW * (int)sizeof(TA), k * ACT * CW * (int)sizeof(TA), a[k][0], a[k][1]);
                else
                    ldbuf1(rs, lane * (int)sizeof(TA), k * ACT * (int)sizeof(TA), a[k][0]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < NF; k++) {
                if constexpr (CW == 2)
                    ldnt2(sbase + (int64_t)k * ACT * 2 + lane * 2, a[k][0], a[k][1]);
                else
                    ldnt(sbase + (int64_t)k * ACT + lane, a[k][0]);
            }
        }
    }

    // output lane (r, i, part 0) of row r, entry i; parts 1 .. NP - 1 (NP a
    // power of two) help to sum
    const int NP = 1 << (31 - __clz(64 / (RW * NF)));
    const int r = lane / (NF * NP), rem = lane - r * (NF * NP), i = rem % NF, part = rem / NF;
    const int row = s * RW + r;
    const bool out = r < RW && part == 0 && row < nbr;  // nbr: rows of this part
    const int64_t e = (int64_t)(MAP && out ? rowmap[row] : row) * NF + i;
    AT bv = 0, xv = 0, dv = 0, yv = 0;
    if (out) {
        if (MODE != OP_SPMV) bv = (AT)b[e];
        if (SMOOTHING) xv = (AT)x[e];
        if (MODE == OP_CHEB && alpha != 0.0) dv = (AT)aux[e];
        if (MODE == OP_SPMV && beta != 0.0) yv = (AT)y[e];
    }
    // inverse-block lane (dr, di, dm)
    const bool dist = SMOOTHING && RW * NF * JP <= 64;
    const bool dlane = dist && lane < RW * NF * JP;
    const int dr = dlane ? lane / (NF * JP) : 0;
    const int di = dlane ? (lane - dr * (NF * JP)) / JP : 0;
    const int dm = dlane ? lane - dr * (NF * JP) - di * JP : 0;
    AT d0 = 0, d1 = 0;
    const int drow = s * RW + dr;
    if (dlane && drow < nbr) {
        const TD* D = invD + ((int64_t)(MAP ? rowmap[drow] : drow) * NF + di) * NF + CW * dm;
        d0 = (AT)ldv(D);
        if constexpr (CW == 2) d1 = (AT)ldv(D + 1);
    }

    AT acc[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) acc[k] = 0;
    if (act) {
        // x of padding slots is taken as zero without a load
        if constexpr (CW == 2) {
            TV x0 = 0, x1 = 0;
            if (c >= 0) ld2(x + (int64_t)c * NF + 2 * jp, x0, x1);
#pragma unroll
            for (int k = 0; k < NF; k++) acc[k] = (AT)a[k][0] * (AT)x0 + (AT)a[k][1] * (AT)x1;
        } else {
            const AT x0 = c >= 0 ? (AT)x[(int64_t)c * NF + jp] : (AT)0;
#pragma unroll
            for (int k = 0; k < NF; k++) acc[k] = (AT)a[k][0] * x0;
        }
    }
    // Row k of the wavefront's LDS holds the 64 partial sums of entry k,
    // rotated by k so that the lanes summing different entries read
    // different banks.
#pragma unroll
    for (int k = 0; k < NF; k++) red[wave][k][(lane + k) & 63] = acc[k];
    wave_lds_sync();
    // Row sums: NP lanes per (row r, entry i) each add a run of the LPR
    // partial sums, then a shuffle tree over the parts leaves the row sum in
    // part 0 (the other parts end with partial or foreign sums, not used).
    const int QC = (LPR + NP - 1) / NP;
    AT sum = 0;
    if (r < RW) {
        const int q1 = min(LPR, (part + 1) * QC);
        for (int q = part * QC; q < q1; q++) sum += red[wave][i][(r * LPR + q + i) & 63];
    }
    for (int off = NP >> 1; off > 0; off >>= 1) sum += __shfl(sum, min(lane + NF * off, 63), 64);
    if (MODE == OP_SPMV) {
        if (out) {
            AT v = (AT)alpha * sum;
            if (beta != 0.0) v += (AT)beta * yv;
            y[e] = (TV)v;
        }
    } else if (MODE == OP_RESID) {
        if (out) y[e] = (TV)(bv - sum);
    } else {
        // z = D^-1 (b - A x) of the row's NF residuals
        const AT res = bv - sum;
        AT z = 0;
        if (dist) {
            const int src = min(dr * NF * NP + CW * dm, 63);
            AT zp = __shfl(res, src, 64) * d0;
            if constexpr (CW == 2) zp += __shfl(res, min(src + 1, 63), 64) * d1;
            z = zp;
            for (int p = 1; p < JP; p++) {
                const AT o = __shfl(zp, min(lane + p, 63), 64);
                if (dm == 0) z += o;
            }
            z = __shfl(z, min(r * NF * JP + i * JP, 63), 64);
        } else {
            // the residuals go through the wavefront's LDS rows, whose
            // partial sums have all been read
            AT* rs = &red[wave][0][0];
            wave_lds_sync();
            if (out) rs[r * NF + i] = res;
            wave_lds_sync();
            if (out) {
                const TD* D = invD + e * NF;
#pragma unroll
                for (int k = 0; k < NF; k++) z += (AT)ldv(D + k) * rs[r * NF + k];
            }
        }
        if (out) {
            if constexpr (MODE == OP_CHEB) {
                const AT dn = (alpha != 0.0 ? (AT)alpha * dv : (AT)0) + (AT)beta * z;
                aux[e] = (TV)dn;
                y[e] = (TV)(xv + dn);
            } else {
                y[e] = (TV)(xv + (AT)omega * z);
            }
        }
    }
}

// Residual of the first entry of every block row, rc[row] = b[row NF] -
// (A x)[row NF], on the sliced layout (one part, no row map): the residual
// followed by the p-restriction to one mode per face (LevelForHHO.h:586-620
// truncates b - A x to the lowest modes) without forming the rest of the
// residual.  Only entry 0 of every slice's values is read (one contiguous
// run per wavefront, 1/NF of the operator's bytes); the x gather is that of
// the full product.  A wavefront takes kHeadSlices consecutive slices and
// issues the loads of all of them before the first use (column indices,
// then values and b, then the x gathers), since one slice alone is a short,
// latency-bound chain.  Row sums as in k_sliced: NP lanes per row each add
// a run of the LPR partial sums, then a shuffle tree.
constexpr int kHeadSlices = 4;
template <int NF, class TA, class TV>
__global__ __launch_bounds__(kWG) void k_sliced_resid_head(int nbr, int nslices, int W, int RW,
                                                           const int* __restrict__ cols, const TA* __restrict__ sv,
                                                           const TV* x, const TV* __restrict__ b,
                                                           TV* __restrict__ rc) {
    using AT = typename Acc<TV>::T;
    using LT = std::conditional_t<std::is_same<TA, double>::value, double, float>;  // loaded value type
    constexpr int CW = NF % 2 == 0 ? 2 : 1;
    constexpr int JP = NF / CW;
    constexpr int H = kHeadSlices;
    const int LPR = W * JP, ACT = RW * LPR;
    __shared__ AT red[kSlicedWaves][64];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s0 = (xcd_remap(blockIdx.x, gridDim.x) * kSlicedWaves + wave) * H;
    if (s0 >= nslices) return;  // wavefront-uniform; no workgroup barrier follows
    const int rv = lane / LPR, rl = lane - rv * LPR, bs = rl / JP, jp = rl - bs * JP;
    const bool act = lane < ACT;
    const int NP = 1 << (31 - __clz(64 / RW));
    const int r = lane / NP, part = lane - r * NP;
    int c[H];
    LT a0[H], a1[H];
    AT bv[H];
#pragma unroll
    for (int h = 0; h < H; h++) {
        const int s = s0 + h;
        c[h] = act && s < nslices ? cols[(int64_t)s * RW * W + rv * W + bs] : -1;
    }
#pragma unroll
    for (int h = 0; h < H; h++) {
        const int s = s0 + h;
        a0[h] = 0;
        a1[h] = 0;
        if (act && s < nslices) {
            const TA* sbase = sv + (int64_t)s * NF * ACT * CW;  // entry 0 of the slice: its first ACT * CW values
            if constexpr (CW == 2) ldnt2(sbase + lane * 2, a0[h], a1[h]);
            else ldnt(sbase + lane, a0[h]);
        }
        const int row = s * RW + r;
        bv[h] = r < RW && part == 0 && row < nbr ? (AT)b[(int64_t)row * NF] : (AT)0;
    }
    AT acc[H];
#pragma unroll
    for (int h = 0; h < H; h++) {
        acc[h] = 0;
        if (c[h] >= 0) {
            if constexpr (CW == 2) {
                TV x0, x1;
                ld2(x + (int64_t)c[h] * NF + 2 * jp, x0, x1);
                acc[h] = (AT)a0[h] * (AT)x0 + (AT)a1[h] * (AT)x1;
            } else {
                acc[h] = (AT)a0[h] * (AT)x[(int64_t)c[h] * NF + jp];
            }
        }
    }
    const int QC = (LPR + NP - 1) / NP;
#pragma unroll
    for (int h = 0; h < H; h++) {
        red[wave][lane] = acc[h];
        wave_lds_sync();
        AT sum = 0;
        if (r < RW) {
            const int q1 = min(LPR, (part + 1) * QC);
            for (int q = part * QC; q < q1; q++) sum += red[wave][r * LPR + q];
        }
        wave_lds_sync();  // every lane has read its partial sums before the next slice's are stored
        for (int off = NP >> 1; off > 0; off >>= 1) sum += __shfl(sum, min(lane + off, 63), 64);
        const int row = (s0 + h) * RW + r;
        if (r < RW && part == 0 && row < nbr) rc[row] = (TV)(bv[h] - sum);
    }
}

// Block-row operation on the chunked layout of wide rows (Galerkin coarse
// operators, restrictions: 40-100 blocks per row).  One wavefront per row;
// lane (b, jp) takes column group jp of block slot b of each chunk of W
// slots, so each value load of the wavefront reads one contiguous run; two
// chunks are in flight at a time.  The NF row sums are reduced across the
// wavefront with xor-shuffles, after which every lane holds all of them.
template <int NF, class TA, class TV, class TD, int MODE>
__global__ __launch_bounds__(kWG) void k_wide(int nbr, int W, const int* __restrict__ cp,
                                              const int* __restrict__ cols, const TA* __restrict__ sv, const TV* x,
                                              const TV* __restrict__ b, TV* y, double alpha, double beta,
                                              double omega, const TD* __restrict__ invD,
                                              TV* __restrict__ aux) {
    using AT = typename Acc<TV>::T;
    constexpr int CW = NF % 2 == 0 ? 2 : 1;
    constexpr int JP = NF / CW;
    const int ACT = W * JP;
    const int lane = threadIdx.x & 63;
    const int row = xcd_remap(blockIdx.x, gridDim.x) * kSlicedWaves + (int)(threadIdx.x >> 6);
    if (row >= nbr) return;  // wave-uniform; no barrier below
    const int bs = lane / JP, jp = lane - bs * JP;
    const bool act = lane < ACT;
    AT acc[NF];
#pragma unroll
    for (int i = 0; i < NF; i++) acc[i] = 0;
    const int q0 = cp[row], q1 = cp[row + 1];
    // BF16 storage with even NF packs block rows 2h, 2h + 1 of the lane's
    // column pair into one 8-byte load: [chunk][h < NF/2][lane][4]
    constexpr bool PK = std::is_same<TA, bf16_t>::value && CW == 2;
    for (int q = q0; q < q1; q += 2) {
        AT a[2][NF][CW];
        int c[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int qq = q + u;
            const bool ok = act && qq < q1;
            c[u] = ok ? cols[(int64_t)qq * W + bs] : -1;
            const TA* base = sv + ((int64_t)qq * NF * ACT + lane) * CW;
#pragma unroll
            for (int i = 0; i < NF; i++)
#pragma unroll
                for (int k = 0; k < CW; k++) a[u][i][k] = 0;
            if (ok) {
                if constexpr (PK) {
                    const TA* b4 = sv + ((int64_t)qq * NF * ACT + (int64_t)lane * 2) * CW;
#pragma unroll
                    for (int h = 0; h < NF / 2; h++)
                        ldnt4(b4 + (int64_t)h * ACT * 4, a[u][2 * h][0], a[u][2 * h][1], a[u][2 * h + 1][0],
                              a[u][2 * h + 1][1]);
                } else {
#pragma unroll
                    for (int i = 0; i < NF; i++) {
                        if constexpr (CW == 2) ldnt2(base + (int64_t)i * ACT * 2, a[u][i][0], a[u][i][1]);
                        else ldnt(base + (int64_t)i * ACT, a[u][i][0]);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            if constexpr (CW == 2) {
                TV x0 = 0, x1 = 0;
                if (c[u] >= 0) ld2(x + (int64_t)c[u] * NF + 2 * jp, x0, x1);
#pragma unroll
                for (int i = 0; i < NF; i++) acc[i] += a[u][i][0] * (AT)x0 + a[u][i][1] * (AT)x1;
            } else {
                const AT x0 = c[u] >= 0 ? (AT)x[(int64_t)c[u] * NF + jp] : (AT)0;
#pragma unroll
                for (int i = 0; i < NF; i++) acc[i] += a[u][i][0] * x0;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NF; i++)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc[i] += __shfl_xor(acc[i], off, 64);
    if (lane >= NF) return;
    const int64_t e = (int64_t)row * NF + lane;
    AT sum = 0;
#pragma unroll
    for (int i = 0; i < NF; i++)
        if (lane == i) sum = acc[i];
    if (MODE == OP_SPMV) {
        AT v = (AT)alpha * sum;
        if (beta != 0.0) v += (AT)beta * (AT)y[e];
        y[e] = (TV)v;
    } else if (MODE == OP_RESID) {
        y[e] = (TV)((AT)b[e] - sum);
    } else {
        // y = x + omega D^-1 (b - A x); every lane holds the row's NF sums
        const TD* D = invD + e * NF;
        const TV* bb = b + (int64_t)row * NF;
        AT z = 0;
#pragma unroll
        for (int k = 0; k < NF; k++) z += (AT)ldv(D + k) * ((AT)bb[k] - acc[k]);
        smooth_out<MODE>(e, z, x, y, aux, alpha, beta, omega);
    }
}

// Block-CSR values -> chunked values of wide rows (one thread per chunk lane).
template <int NF, class TA>
__global__ void k_to_wide(int nchunks, int W, const int* __restrict__ chunk_row, const int* __restrict__ cp,
                          const int* __restrict__ rp, const TA* __restrict__ val, TA* __restrict__ sv) {
    constexpr int CW = NF % 2 == 0 ? 2 : 1;
    constexpr int JP = NF / CW;
    const int ACT = W * JP;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nchunks * ACT) return;
    const int q = (int)(t / ACT), lane = (int)(t - (int64_t)q * ACT);
    const int bs = lane / JP, jp = lane - bs * JP;
    const int row = chunk_row[q];
    const int slot = (q - cp[row]) * W + bs;
    const bool ok = slot < rp[row + 1] - rp[row];
    const int64_t blk = ok ? (int64_t)rp[row] + slot : 0;
    constexpr bool PK = std::is_same<TA, bf16_t>::value && CW == 2;
    for (int i = 0; i < NF; i++)
        for (int u = 0; u < CW; u++) {
            TA v{};
            if (ok) v = val[(blk * NF + i) * NF + CW * jp + u];
            if constexpr (PK)  // [chunk][i / 2][lane][(i % 2) * 2 + u]
                sv[(((int64_t)q * (NF / 2) + i / 2) * ACT + lane) * 4 + (i % 2) * 2 + u] = v;
            else
                sv[(((int64_t)q * NF + i) * ACT + lane) * CW + u] = v;
        }
}

// Block-CSR values -> sliced values (one thread per slice lane).
template <int NF, class TA>
__global__ void k_to_sliced(int nbr, int nslices, int W, int RW, const int* __restrict__ rp,
                            const TA* __restrict__ val, TA* __restrict__ sv, const int* __restrict__ rowmap) {
    constexpr int CW = NF % 2 == 0 ? 2 : 1;
    constexpr int JP = NF / CW;
    const int LPR = W * JP, ACT = RW * LPR;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nslices * ACT) return;
    const int s = (int)(t / ACT), lane = (int)(t - (int64_t)s * ACT);
    const int r = lane / LPR, rem = lane - r * LPR, bs = rem / JP, jp = rem - bs * JP;
    const int prow = s * RW + r;
    const int row = prow < nbr ? (rowmap ? rowmap[prow] : prow) : 0;
    const bool ok = prow < nbr && bs < rp[row + 1] - rp[row];
    const int64_t blk = ok ? (int64_t)rp[row] + bs : 0;
    for (int i = 0; i < NF; i++)
        for (int u = 0; u < CW; u++) {
            TA v{};
            if (ok) v = val[(blk * NF + i) * NF + CW * jp + u];
            sv[(((int64_t)s * NF + i) * ACT + lane) * CW + u] = v;
        }
}

// SELL-64 layout of scalar operators (n_f = 1: the aggregation levels).
// Slice s holds rows 64 s .. 64 s + 63 with w_s = sptr[s + 1] - sptr[s]
// slots (its longest row); its entries are stored slot-major from
// 64 sptr[s], lane = row, so every load instruction of a wavefront reads 64
// consecutive values and every lane has w_s independent loads to issue.
// Padding entries hold the value 0 and the lane's own row as column.
constexpr int kSellC = 64;
constexpr int kSellUnroll = 4;
template <class TA, class TV, class TD, int MODE>
__global__ __launch_bounds__(kWG) void k_sell(int nbr, int nslices, const int* __restrict__ sptr,
                                              const int* __restrict__ cols, const TA* __restrict__ sv, const TV* x,
                                              const TV* b, TV* y, double alpha, double beta, double omega,
                                              const TD* __restrict__ invD, TV* __restrict__ aux) {
    using AT = typename Acc<TV>::T;
    using LT = std::conditional_t<std::is_same<TA, double>::value, double, float>;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s = xcd_remap(blockIdx.x, gridDim.x) * (kWG / 64) + wave;
    if (s >= nslices) return;  // wavefront-uniform
    const int p0 = sptr[s], w = sptr[s + 1] - p0;
    const int row = s * kSellC + lane;
    const int* cs = cols + (int64_t)p0 * kSellC + lane;
    const TA* vs = sv + (int64_t)p0 * kSellC + lane;
    // epilogue operands first: their loads overlap the product's
    AT bv = 0, xv = 0, dv = 0, yv = 0, dinv = 0;
    const bool out = row < nbr;
    if (out) {
        if (MODE != OP_SPMV) bv = (AT)b[row];
        if (MODE == OP_CHEB || MODE == OP_SMOOTH) { xv = (AT)x[row]; dinv = (AT)ldv(invD + row); }
        if (MODE == OP_CHEB && alpha != 0.0) dv = (AT)aux[row];
        if (MODE == OP_SPMV && beta != 0.0) yv = (AT)y[row];
    }
    AT acc = 0;
    int j = 0;
    for (; j + kSellUnroll <= w; j += kSellUnroll) {
        int c[kSellUnroll];
        LT a[kSellUnroll];
#pragma unroll
        for (int u = 0; u < kSellUnroll; u++) c[u] = cs[u * kSellC];
#pragma unroll
        for (int u = 0; u < kSellUnroll; u++) ldnt(vs + u * kSellC, a[u]);
#pragma unroll
        for (int u = 0; u < kSellUnroll; u++) acc += (AT)a[u] * (AT)x[c[u]];
        cs += kSellUnroll * kSellC;
        vs += kSellUnroll * kSellC;
    }
    for (; j < w; j++) {
        LT a;
        const int c = cs[0];
        ldnt(vs, a);
        acc += (AT)a * (AT)x[c];
        cs += kSellC;
        vs += kSellC;
    }
    if (!out) return;
    if constexpr (MODE == OP_SPMV) {
        AT v = (AT)alpha * acc;
        if (beta != 0.0) v += (AT)beta * yv;
        y[row] = (TV)v;
    } else if constexpr (MODE == OP_RESID) {
        y[row] = (TV)(bv - acc);
    } else if constexpr (MODE == OP_CHEB) {
        const AT dn = (alpha != 0.0 ? (AT)alpha * dv : (AT)0) + (AT)beta * dinv * (bv - acc);
        aux[row] = (TV)dn;
        y[row] = (TV)(xv + dn);
    } else {
        y[row] = (TV)(xv + (AT)omega * dinv * (bv - acc));
    }
}

// Block-CSR (n_f = 1) -> SELL-64 values and columns, one thread per
// (slice, slot, lane).
template <class TA>
__global__ void k_to_sell(int nbr, int nslices, const int* __restrict__ sptr, const int* __restrict__ rp,
                          const int* __restrict__ ci, const TA* __restrict__ val, int* __restrict__ cols,
                          TA* __restrict__ sv, int64_t total) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    // slice of entry t: the last s with 64 sptr[s] <= t (binary search)
    int lo = 0, hi = nslices;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)sptr[mid] * kSellC <= t) lo = mid;
        else hi = mid;
    }
    const int s = lo;
    const int64_t off = t - (int64_t)sptr[s] * kSellC;
    const int j = (int)(off / kSellC), lane = (int)(off - (int64_t)j * kSellC);
    const int row = s * kSellC + lane;
    int c = row < nbr ? row : 0;
    TA v{};
    if (row < nbr && j < rp[row + 1] - rp[row]) {
        c = ci[rp[row] + j];
        v = val[rp[row] + j];
    }
    cols[t] = c;
    sv[t] = v;
}

template <class TV, class TD>
__global__ void k_bjac_zero(int count, int nf, const TV* __restrict__ b, TV* __restrict__ y, double omega,
                            const TD* __restrict__ invD, const int* __restrict__ rows, TV* __restrict__ y2,
                            int vec2) {
    using AT = typename Acc<TV>::T;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)count * nf) return;
    int w = (int)(t / nf), i = (int)(t - (int64_t)w * nf);
    int br = rows ? rows[w] : w;
    const TD* D = invD + ((int64_t)br * nf + i) * nf;
    const TV* bb = b + (int64_t)br * nf;
    AT z = 0;
    if (vec2) {
        // even block size and 2-value-aligned b and invD: 2-value loads
        for (int j = 0; j < nf; j += 2) {
            TD d0, d1;
            TV b0, b1;
            ld2(D + j, d0, d1);
            ld2(bb + j, b0, b1);
            z += (AT)d0 * (AT)b0;
            z += (AT)d1 * (AT)b1;
        }
    } else {
        for (int j = 0; j < nf; j++) z += (AT)ldv(D + j) * (AT)bb[j];
    }
    const TV v = (TV)((AT)omega * z);
    y[(int64_t)br * nf + i] = v;
    if (y2) y2[(int64_t)br * nf + i] = v;  // first Chebyshev step: iterate and direction
}

// y = omega D^-1 b over all block rows, LDS-tiled: a workgroup takes
// R = 256 / NF consecutive block rows.  Their inverse blocks (R NF^2 values,
// one contiguous span) and their b entries are streamed into LDS with
// consecutive lanes on consecutive values, so every load instruction reads
// one contiguous run; then thread t of the tile forms output entry t from
// LDS and the outputs are written contiguously.  The inverses are read once
// (non-temporal).
template <int NF, class TV, class TD>
__global__ __launch_bounds__(kWG) void k_bjac_zero_tile(int nbr, const TV* __restrict__ b, TV* __restrict__ y,
                                                        double omega, const TD* __restrict__ invD,
                                                        TV* __restrict__ y2) {
    using AT = typename Acc<TV>::T;
    constexpr int R = kWG / NF;
    constexpr int NV = R * NF;  // output entries of a full tile
    __shared__ __align__(16) TD sD[R * NF * NF];
    __shared__ AT sb[NV];
    const int t = threadIdx.x;
    const int r0 = blockIdx.x * R;
    const int rows = min(R, nbr - r0);
    const int nv = rows * NF;
    const int ne = nv * NF;
    const TD* Dg = invD + (int64_t)r0 * NF * NF;
    if constexpr (NF % 2 == 0 && sizeof(TD) == 8) {
        // two values per load (a tile starts at an even value offset; the
        // launcher checks the base alignment)
        typedef double d2_t __attribute__((ext_vector_type(2)));
        const d2_t* D2 = (const d2_t*)Dg;
        d2_t* S2 = (d2_t*)sD;
        for (int e = t; e < ne / 2; e += kWG) S2[e] = __builtin_nontemporal_load(D2 + e);
    } else {
        for (int e = t; e < ne; e += kWG) sD[e] = __builtin_nontemporal_load(Dg + e);
    }
    if (t < nv) sb[t] = (AT)b[(int64_t)r0 * NF + t];
    __syncthreads();
    if (t >= nv) return;
    const int r = t / NF;
    const TD* Dr = sD + t * NF;  // row i = t - r NF of block r: (r NF + i) NF = t NF
    const AT* br = sb + r * NF;
    AT z = 0;
#pragma unroll
    for (int j = 0; j < NF; j++) z += (AT)Dr[j] * br[j];
    const TV v = (TV)((AT)omega * z);
    const int64_t o = (int64_t)r0 * NF + t;
    y[o] = v;
    if (y2) y2[o] = v;  // first Chebyshev step: iterate and direction
}

template <class TD>
__global__ void k_block_inverse(int nbr, int nf, const double* __restrict__ val, const int* __restrict__ diag,
                                TD* __restrict__ invD, int* flag) {
    int br = blockIdx.x * blockDim.x + threadIdx.x;
    if (br >= nbr) return;
    double a[100], inv[100];
    const double* D = val + (int64_t)diag[br] * nf * nf;
    for (int k = 0; k < nf * nf; k++) a[k] = D[k];
    for (int r = 0; r < nf; r++)
        for (int c = 0; c < nf; c++) inv[r * nf + c] = r == c ? 1.0 : 0.0;
    for (int c = 0; c < nf; c++) {
        int p = c;
        for (int r = c + 1; r < nf; r++)
            if (fabs(a[r * nf + c]) > fabs(a[p * nf + c])) p = r;
        if (a[p * nf + c] == 0.0) {
            atomicCAS(flag, 0, br + 1);
            return;
        }
        if (p != c)
            for (int j = 0; j < nf; j++) {
                double t = a[p * nf + j]; a[p * nf + j] = a[c * nf + j]; a[c * nf + j] = t;
                t = inv[p * nf + j]; inv[p * nf + j] = inv[c * nf + j]; inv[c * nf + j] = t;
            }
        double d = 1.0 / a[c * nf + c];
        for (int j = 0; j < nf; j++) { a[c * nf + j] *= d; inv[c * nf + j] *= d; }
        for (int r = 0; r < nf; r++) {
            if (r == c) continue;
            double f = a[r * nf + c];
            if (f == 0.0) continue;
            for (int j = 0; j < nf; j++) { a[r * nf + j] -= f * a[c * nf + j]; inv[r * nf + j] -= f * inv[c * nf + j]; }
        }
    }
    TD* out = invD + (int64_t)br * nf * nf;
    for (int k = 0; k < nf * nf; k++) stv(out + k, inv[k]);
}

template <class TS, class TDd>
__global__ void k_convert(int64_t n, const TS* __restrict__ s, TDd* __restrict__ d) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        stv(d + t, (double)ldv(s + t));
}

template <class T>
__global__ void k_fill(int64_t n, T* x, double v) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        stv(x + t, v);
}

__device__ __forceinline__ double scal(const double* s, double h, int ia, int ib) {
    double v = h;
    if (ia >= 0) v *= s[ia];
    if (ib >= 0) v /= s[ib];
    return v;
}

template <class T>
__global__ void k_axpby(int64_t n, const T* __restrict__ x, T* __restrict__ y, const double* s, double ha, int ia,
                        int ib, double hb, int ic, int id) {
    using AT = typename Acc<T>::T;
    const AT a = (AT)scal(s, ha, ia, ib);
    const AT bb = (AT)scal(s, hb, ic, id);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        AT v = a * (AT)x[t];
        if (bb != (AT)0) v += bb * (AT)y[t];
        y[t] = (T)v;
    }
}

// FCG's orthogonalisation of z against up to kOrthVecs kept directions in
// one pass: out = z + sum_k a_k v_k with a_k = -s[num_k] / s[den_k], added in
// k order with the arithmetic of successive k_axpby updates.  out may alias z.
constexpr int kOrthVecs = 4;
struct OrthArgs {
    const double* v[kOrthVecs];
    int num[kOrthVecs], den[kOrthVecs];
};
template <int NV>
__global__ void k_fcg_orth(int64_t n, const double* z, OrthArgs p, const double* s, double* out) {
    double a[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) a[k] = scal(s, -1.0, p.num[k], p.den[k]);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        double y = z[t];
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double v = a[k] * p.v[k][t];
            v += 1.0 * y;
            y = v;
        }
        out[t] = y;
    }
}

// Chebyshev smoothing step: d = a d + c (t - x); x = x + d, where
// t = x + D^-1 (b - A x) was produced by the block-row kernel.
template <class T>
__global__ void k_cg_update(int64_t n, T* __restrict__ x, const T* __restrict__ d, T* __restrict__ r,
                            const T* __restrict__ q, const double* s, int inum, int iden, double sign) {
    using AT = typename Acc<T>::T;
    const double al = sign * s[inum] / s[iden];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        x[t] = (T)((AT)x[t] + (AT)al * (AT)d[t]);
        r[t] = (T)((AT)r[t] - (AT)al * (AT)q[t]);
    }
}

template <int ND>
__device__ __forceinline__ void block_reduce(double (&v)[3], double* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < ND; k++) {
        double a = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
        if (lane == 0) sh[k * 16 + wid] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
#pragma unroll
        for (int k = 0; k < ND; k++) {
            double a = 0;
            for (int w = 0; w < nw; w++) a += sh[k * 16 + w];
            v[k] = a;
        }
    }
}

// SX: every product has the same x (read once; the products and sums are
// those of separate x reads)
template <class T, int ND, bool SX>
__global__ __launch_bounds__(kWG) void k_dots(int64_t n, const T* x0, const T* y0, const T* x1, const T* y1,
                                               const T* x2, const T* y2, double* partials) {
    __shared__ double sh[3 * 16];
    double v[3] = {0, 0, 0};
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const double xv = (double)x0[t];
        v[0] += xv * (double)y0[t];
        if (ND > 1) v[1] += (SX ? xv : (double)x1[t]) * (double)y1[t];
        if (ND > 2) v[2] += (SX ? xv : (double)x2[t]) * (double)y2[t];
    }
    block_reduce<ND>(v, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < ND; k++) partials[k * kDotBlocks + blockIdx.x] = v[k];
}

// The CG / FCG update fused with |r|^2 of the updated residual, on the
// reduction grid of k_dots (the same per-thread order and tree, so the sum
// is the one a separate dot product of r with itself gives).
__global__ __launch_bounds__(kWG) void k_cg_update_rr(int64_t n, double* __restrict__ x, const double* __restrict__ d,
                                                      double* __restrict__ r, const double* __restrict__ q,
                                                      const double* s, int inum, int iden, double sign,
                                                      double* partials) {
    __shared__ double sh[3 * 16];
    const double al = sign * s[inum] / s[iden];
    double v[3] = {0, 0, 0};
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        x[t] = x[t] + al * d[t];
        const double rn = r[t] - al * q[t];
        r[t] = rn;
        v[0] += rn * rn;
    }
    block_reduce<1>(v, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

template <int ND>
__global__ __launch_bounds__(1024) void k_dots_final(const double* partials, double* s, int s0, int s1, int s2) {
    __shared__ double sh[3 * 16];
    double v[3] = {0, 0, 0};
    const int t = threadIdx.x;
    for (int k = 0; k < ND; k++) v[k] = partials[k * kDotBlocks + t];
    block_reduce<ND>(v, sh);
    if (t == 0) {
        s[s0] = v[0];
        if (ND > 1) s[s1] = v[1];
        if (ND > 2) s[s2] = v[2];
    }
}

__global__ void k_scalar_op(double* s, int op, int dst, int a, int b) {
    double r = 0;
    switch (op) {
        case 0: r = s[a] / s[b]; break;
        case 1: r = sqrt(s[a]); break;
        case 2: r = s[a] * s[b]; break;
        case 3: r = s[a] - s[b]; break;
        default: r = s[a]; break;
    }
    s[dst] = r;
}

// Coefficients of the K-cycle's FCG step (Multigrid::FCGForKCycle,
// Multigrid.h:611-656) from the device slots s[k .. k+6] = rho1, alpha1,
// |r|^2, |r_new|^2, gamma, beta, alpha2: when |r_new| <= t |r| the first
// correction alone is kept (ca = alpha1 / rho1, cb = 0), else
// rho2 = beta - gamma^2 / rho1, ca = alpha1 / rho1 - gamma alpha2 / (rho1 rho2),
// cb = alpha2 / rho2.  ca and cb replace s[k+2] and s[k+3].
__global__ void k_kcycle_coeffs(double* s, int k, double t) {
    const double rho1 = s[k], alpha1 = s[k + 1], rold = sqrt(s[k + 2]), rnew = sqrt(s[k + 3]);
    double ca, cb;
    if (rnew <= t * rold) {
        ca = alpha1 / rho1;
        cb = 0.0;
    } else {
        const double gamma = s[k + 4], beta = s[k + 5], alpha2 = s[k + 6];
        const double rho2 = __dsub_rn(beta, __ddiv_rn(__dmul_rn(gamma, gamma), rho1));
        ca = __dsub_rn(__ddiv_rn(alpha1, rho1), __ddiv_rn(__dmul_rn(gamma, alpha2), __dmul_rn(rho1, rho2)));
        cb = alpha2 / rho2;
    }
    s[k + 2] = ca;
    s[k + 3] = cb;
}

template <class T>
__global__ void k_agg_restrict(int nagg, int nf, const int* __restrict__ mptr, const int* __restrict__ members,
                               const T* __restrict__ r, T* __restrict__ rc) {
    using AT = typename Acc<T>::T;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nagg * nf) return;
    int I = (int)(t / nf), a = (int)(t - (int64_t)I * nf);
    AT s = 0;
    for (int q = mptr[I]; q < mptr[I + 1]; q++) s += (AT)r[(int64_t)members[q] * nf + a];
    rc[t] = (T)s;
}

template <class T>
__global__ void k_agg_prolong_add(int nb, int nf, const int* __restrict__ agg, const T* __restrict__ ec,
                                  T* __restrict__ x) {
    using AT = typename Acc<T>::T;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nb * nf) return;
    int k = (int)(t / nf), a = (int)(t - (int64_t)k * nf);
    x[t] = (T)((AT)x[t] + (AT)ec[(int64_t)agg[k] * nf + a]);
}

template <class T>
__global__ void k_p_restrict(int nb, int nf, int nfc, const T* __restrict__ r, T* __restrict__ rc) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nb * nfc) return;
    int i = (int)(t / nfc), a = (int)(t - (int64_t)i * nfc);
    rc[t] = r[(int64_t)i * nf + a];
}

template <class T>
__global__ void k_p_prolong_add(int nb, int nf, int nfc, const T* __restrict__ ec, T* __restrict__ x) {
    using AT = typename Acc<T>::T;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nb * nfc) return;
    int i = (int)(t / nfc), a = (int)(t - (int64_t)i * nfc);
    int64_t j = (int64_t)i * nf + a;
    x[j] = (T)((AT)x[j] + (AT)ec[t]);
}

// One wavefront per row of a dense symmetric matrix.
template <class TM, class TV>
__global__ __launch_bounds__(kWG) void k_dense_symv(int n, const TM* __restrict__ M, const TV* __restrict__ x,
                                                     TV* __restrict__ y) {
    using AT = typename Acc<TV>::T;
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (kWG / 64) + (threadIdx.x >> 6);
    if (row >= n) return;
    const TM* m = M + (int64_t)row * n;
    AT s = 0;
    for (int j = lane; j < n; j += 64) s += (AT)ldv(m + j) * (AT)x[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) y[row] = (TV)s;
}

inline int grid_for(int64_t n, int bs = 256, int cap = 65536) {
    int64_t g = (n + bs - 1) / bs;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

// ---------------------------------------------------------------- dispatch
// Block-row kernel variant for even block sizes: 1 (default) 16 lanes per
// row with contiguous 2-value chunks, 0 one wavefront per row.  Read once
// from FHGPU_BCSR.
int bcsr_variant() {
    static const int v = [] {
        const char* e = std::getenv("FHGPU_BCSR");
        return e ? std::atoi(e) : 1;
    }();
    return v;
}

template <int NF, class TA, class TV, class TD>
void bcsr_dispatch_mode(Ctx& c, const Bcsr& A, BcsrMode mode, const void* x, const void* b, void* y, double alpha,
                        double beta, double omega, const void* invD, const int* rows, int nrows, void* aux) {
    int count = rows ? nrows : A.nbr;
    if (count <= 0) return;
    auto rp = A.rowptr.as<int>();
    auto ci = A.colidx.as<int>();
    auto val = A.vals.as<TA>();
    auto xx = (const TV*)x;
    auto bb = (const TV*)b;
    auto yy = (TV*)y;
    auto D = (const TD*)invD;
    auto ax = (TV*)aux;
    if constexpr (NF == 1) {
        if (!rows && A.sliced_state == 1 && A.sliced->kind == 2) {
            const Sliced& S = *A.sliced;
            const int nwg = (S.nslices + kWG / 64 - 1) / (kWG / 64);
            auto sp = S.chunk_ptr.as<int>();
            auto sc = S.cols.as<int>();
            auto sv = S.vals.as<TA>();
            switch (mode) {
                case OP_SPMV:
                    hipLaunchKernelGGL((k_sell<TA, TV, TD, OP_SPMV>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                                       S.nslices, sp, sc, sv, xx, bb, yy, alpha, beta, omega, D, ax);
                    break;
                case OP_RESID:
                    hipLaunchKernelGGL((k_sell<TA, TV, TD, OP_RESID>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                                       S.nslices, sp, sc, sv, xx, bb, yy, alpha, beta, omega, D, ax);
                    break;
                case OP_CHEB:
                    hipLaunchKernelGGL((k_sell<TA, TV, TD, OP_CHEB>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                                       S.nslices, sp, sc, sv, xx, bb, yy, alpha, beta, omega, D, ax);
                    break;
                default:
                    hipLaunchKernelGGL((k_sell<TA, TV, TD, OP_SMOOTH>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                                       S.nslices, sp, sc, sv, xx, bb, yy, alpha, beta, omega, D, ax);
                    break;
            }
            FH_HIP(hipGetLastError());
            return;
        }
    }
    if (!rows && A.sliced_state == 1 && A.sliced->kind == 1) {
        const Sliced& S = *A.sliced;
        const bool aligned = S.cw == 1 || ((uintptr_t)x % (2 * sizeof(TV))) == 0;
        if (aligned) {
            const int nwg = (A.nbr + kSlicedWaves - 1) / kSlicedWaves;
            auto cp = S.chunk_ptr.as<int>();
            auto sc = S.cols.as<int>();
            auto sv = S.vals.as<TA>();
            switch (mode) {
                case OP_SPMV:
                    hipLaunchKernelGGL((k_wide<NF, TA, TV, TD, OP_SPMV>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                                       S.W, cp, sc, sv, xx, bb, yy, alpha, beta, omega, D, ax);
                    break;
                case OP_RESID:
                    hipLaunchKernelGGL((k_wide<NF, TA, TV, TD, OP_RESID>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                                       S.W, cp, sc, sv, xx, bb, yy, alpha, beta, omega, D, ax);
                    break;
                case OP_CHEB:
                    hipLaunchKernelGGL((k_wide<NF, TA, TV, TD, OP_CHEB>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                                       S.W, cp, sc, sv, xx, bb, yy, alpha, beta, omega, D, ax);
                    break;
                default:
                    hipLaunchKernelGGL((k_wide<NF, TA, TV, TD, OP_SMOOTH>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                                       S.W, cp, sc, sv, xx, bb, yy, alpha, beta, omega, D, ax);
                    break;
            }
            FH_HIP(hipGetLastError());
            return;
        }
    }
    if (!rows && A.sliced_state == 1 && A.sliced->kind == 0 &&
        (A.sliced->cw == 1 || ((uintptr_t)x % (2 * sizeof(TV))) == 0)) {
        for (const Sliced* P = A.sliced.get(); P; P = P->next.get()) {
            const Sliced& S = *P;
            auto sc = S.cols.as<int>();
            auto sv = S.vals.as<TA>();
            auto rm = S.rowmap.as<int>();
            const int nrp = S.nrows;
            auto kern = rm ? (mode == OP_SPMV    ? k_sliced<NF, TA, TV, TD, OP_SPMV, true>
                              : mode == OP_RESID ? k_sliced<NF, TA, TV, TD, OP_RESID, true>
                              : mode == OP_CHEB  ? k_sliced<NF, TA, TV, TD, OP_CHEB, true>
                                                 : k_sliced<NF, TA, TV, TD, OP_SMOOTH, true>)
                           : (mode == OP_SPMV    ? k_sliced<NF, TA, TV, TD, OP_SPMV, false>
                              : mode == OP_RESID ? k_sliced<NF, TA, TV, TD, OP_RESID, false>
                              : mode == OP_CHEB  ? k_sliced<NF, TA, TV, TD, OP_CHEB, false>
                                                 : k_sliced<NF, TA, TV, TD, OP_SMOOTH, false>);
            const int nwg = (S.nslices + kSlicedWaves - 1) / kSlicedWaves;
            hipLaunchKernelGGL(kern, dim3(nwg), dim3(kWG), 0, c.stream, nrp, S.nslices, S.W, S.RW, sc, sv, xx, bb, yy,
                               alpha, beta, omega, D, ax, rm);
            FH_HIP(hipGetLastError());
        }
        return;
    }
    if constexpr (NF % 2 == 0) {
        const bool aligned = ((uintptr_t)x % (2 * sizeof(TV))) == 0;
        if (bcsr_variant() == 1 && aligned) {
            const int nwg = (int)(((int64_t)count * 16 + kWG - 1) / kWG);
            switch (mode) {
                case OP_SPMV:
                    hipLaunchKernelGGL((k_bcsr_grp<NF, TA, TV, TD, OP_SPMV>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                                       rp, ci, val, xx, bb, yy, alpha, beta, omega, D, rows, nrows, ax);
                    break;
                case OP_RESID:
                    hipLaunchKernelGGL((k_bcsr_grp<NF, TA, TV, TD, OP_RESID>), dim3(nwg), dim3(kWG), 0, c.stream,
                                       A.nbr, rp, ci, val, xx, bb, yy, alpha, beta, omega, D, rows, nrows, ax);
                    break;
                case OP_CHEB:
                    hipLaunchKernelGGL((k_bcsr_grp<NF, TA, TV, TD, OP_CHEB>), dim3(nwg), dim3(kWG), 0, c.stream,
                                       A.nbr, rp, ci, val, xx, bb, yy, alpha, beta, omega, D, rows, nrows, ax);
                    break;
                default:
                    hipLaunchKernelGGL((k_bcsr_grp<NF, TA, TV, TD, OP_SMOOTH>), dim3(nwg), dim3(kWG), 0, c.stream,
                                       A.nbr, rp, ci, val, xx, bb, yy, alpha, beta, omega, D, rows, nrows, ax);
                    break;
            }
            FH_HIP(hipGetLastError());
            return;
        }
    }
    if constexpr (NF == 1) {
        // rows of scalar levels: G lanes per row from the mean row length
        const double mean = A.nbr > 0 ? (double)A.nnzb / A.nbr : 0.0;
        const int G = mean <= 6 ? 4 : (mean <= 12 ? 8 : (mean <= 24 ? 16 : 32));
        const int nwg = (int)(((int64_t)count * G + kWG - 1) / kWG);
#define FH_CSR(GG, MD)                                                                                              \
        hipLaunchKernelGGL((k_csr_grp<GG, TA, TV, TD, MD>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr, rp, ci, val, \
                           xx, bb, yy, alpha, beta, omega, D, rows, nrows, ax)
#define FH_CSR_G(MD)                          \
        switch (G) {                           \
            case 4: FH_CSR(4, MD); break;      \
            case 8: FH_CSR(8, MD); break;      \
            case 16: FH_CSR(16, MD); break;    \
            default: FH_CSR(32, MD); break;    \
        }
        switch (mode) {
            case OP_SPMV: FH_CSR_G(OP_SPMV); break;
            case OP_RESID: FH_CSR_G(OP_RESID); break;
            case OP_CHEB: FH_CSR_G(OP_CHEB); break;
            default: FH_CSR_G(OP_SMOOTH); break;
        }
#undef FH_CSR_G
#undef FH_CSR
        FH_HIP(hipGetLastError());
        return;
    }
    const int nwg = (count + 3) / 4;
    switch (mode) {
        case OP_SPMV:
            hipLaunchKernelGGL((k_bcsr<NF, TA, TV, TD, OP_SPMV>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr, rp, ci, val,
                               xx, bb, yy, alpha, beta, omega, D, rows, nrows, ax);
            break;
        case OP_RESID:
            hipLaunchKernelGGL((k_bcsr<NF, TA, TV, TD, OP_RESID>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr, rp, ci,
                               val, xx, bb, yy, alpha, beta, omega, D, rows, nrows, ax);
            break;
        case OP_CHEB:
            hipLaunchKernelGGL((k_bcsr<NF, TA, TV, TD, OP_CHEB>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr, rp, ci,
                               val, xx, bb, yy, alpha, beta, omega, D, rows, nrows, ax);
            break;
        default:
            hipLaunchKernelGGL((k_bcsr<NF, TA, TV, TD, OP_SMOOTH>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr, rp, ci,
                               val, xx, bb, yy, alpha, beta, omega, D, rows, nrows, ax);
            break;
    }
    FH_HIP(hipGetLastError());
}

template <int NF>
void bcsr_dispatch_types(Ctx& c, const Bcsr& A, BcsrMode mode, Prec vp, const void* x, const void* b, void* y,
                         double alpha, double beta, double omega, const void* invD, Prec dp, const int* rows,
                         int nrows, void* aux) {
    if (vp == Prec::F64) {
        FH_CHECK(A.prec == Prec::F64 && ((mode != OP_SMOOTH && mode != OP_CHEB) || dp == Prec::F64), -3,
                 "FP64 vectors need FP64 matrix and FP64 block inverses");
        bcsr_dispatch_mode<NF, double, double, double>(c, A, mode, x, b, y, alpha, beta, omega, invD, rows, nrows, aux);
    } else if (vp == Prec::F32) {
        FH_CHECK((mode != OP_SMOOTH && mode != OP_CHEB) || dp == Prec::F32, -3, "FP32 smoothing needs FP32 block inverses");
        if (A.prec == Prec::F32)
            bcsr_dispatch_mode<NF, float, float, float>(c, A, mode, x, b, y, alpha, beta, omega, invD, rows, nrows, aux);
        else if (A.prec == Prec::BF16)
            bcsr_dispatch_mode<NF, bf16_t, float, float>(c, A, mode, x, b, y, alpha, beta, omega, invD, rows, nrows, aux);
        else
            throw Error(-3, "FP32 vectors need an FP32 or BF16 matrix");
    } else
        throw Error(-3, "vectors must be FP64 or FP32");
}

// Rows per wavefront and blocks per row slot of the sliced layout; false
// when a row of A does not fit one wavefront (or the block size has no
// kernel).
bool sliced_geometry(const Bcsr& A, int& W, int& RW, int& cw) {
    const int nf = A.bs_r;
    if (A.bs_r != A.bs_c || A.nbr <= 0 || A.h_rowptr.size() != (size_t)A.nbr + 1 ||
        A.h_colidx.size() != (size_t)A.nnzb || A.h_rowptr[A.nbr] != A.nnzb)
        return false;
    if (nf != 1 && nf != 2 && nf != 3 && nf != 4 && nf != 6 && nf != 10) return false;
    cw = nf % 2 == 0 ? 2 : 1;
    W = 0;
    for (int i = 0; i < A.nbr; i++) W = std::max(W, A.h_rowptr[i + 1] - A.h_rowptr[i]);
    if (W == 0 || W * (nf / cw) > 64) return false;
    // every lane of a row holds a slot, and lane (r, i) < RW * nf writes
    // output entry i of row r
    RW = std::min(64 / (W * (nf / cw)), 64 / nf);
    return true;
}

// Chunk width of the wide-row layout (block slots per chunk = 64 / column
// groups per block) and the number of chunks; false without host arrays or
// a kernel for the block size.
bool wide_geometry(const Bcsr& A, int& W, int& cw, int64_t& nchunks) {
    const int nf = A.bs_r;
    if (A.bs_r != A.bs_c || A.nbr <= 0 || A.h_rowptr.size() != (size_t)A.nbr + 1 ||
        A.h_colidx.size() != (size_t)A.nnzb || A.h_rowptr[A.nbr] != A.nnzb)
        return false;
    if (nf != 1 && nf != 2 && nf != 3 && nf != 4 && nf != 6 && nf != 10) return false;
    cw = nf % 2 == 0 ? 2 : 1;
    W = 64 / (nf / cw);
    nchunks = 0;
    for (int i = 0; i < A.nbr; i++) nchunks += (A.h_rowptr[i + 1] - A.h_rowptr[i] + W - 1) / W;
    return nchunks > 0 && nchunks < INT32_MAX / 64;
}

template <int NF, class TA>
void wide_convert(Ctx& c, const Bcsr& A, Sliced& S, const int* chunk_row) {
    const int64_t n = (int64_t)S.nslices * S.act(NF);
    if (n == 0) return;
    hipLaunchKernelGGL((k_to_wide<NF, TA>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, S.nslices, S.W,
                       chunk_row, S.chunk_ptr.as<int>(), A.rowptr.as<int>(), A.vals.as<TA>(), S.vals.as<TA>());
    FH_HIP(hipGetLastError());
}

template <int NF>
void wide_convert_prec(Ctx& c, const Bcsr& A, Sliced& S, const int* chunk_row) {
    if (A.prec == Prec::F64) wide_convert<NF, double>(c, A, S, chunk_row);
    else if (A.prec == Prec::F32) wide_convert<NF, float>(c, A, S, chunk_row);
    else wide_convert<NF, bf16_t>(c, A, S, chunk_row);
}

// Slot counts of the SELL-64 layout of a scalar operator: sptr (nslices + 1
// entries, in units of 64 entries); returns the total slots.
int64_t sell_geometry(const Bcsr& A, std::vector<int>& sptr) {
    const int nslices = (A.nbr + kSellC - 1) / kSellC;
    sptr.assign(nslices + 1, 0);
    for (int sl = 0; sl < nslices; sl++) {
        int w = 0;
        for (int i = sl * kSellC; i < std::min(A.nbr, (sl + 1) * kSellC); i++)
            w = std::max(w, A.h_rowptr[i + 1] - A.h_rowptr[i]);
        sptr[sl + 1] = sptr[sl] + w;
    }
    return (int64_t)sptr[nslices] * kSellC;
}

template <class TA>
void sell_convert(Ctx& c, const Bcsr& A, Sliced& S, int64_t total) {
    if (total == 0) return;
    hipLaunchKernelGGL((k_to_sell<TA>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c.stream, A.nbr,
                       S.nslices, S.chunk_ptr.as<int>(), A.rowptr.as<int>(), A.colidx.as<int>(), A.vals.as<TA>(),
                       S.cols.as<int>(), S.vals.as<TA>(), total);
    FH_HIP(hipGetLastError());
}

void sell_build(Ctx& c, const Bcsr& A, const std::vector<int>& sptr, int64_t total) {
    auto S = std::make_unique<Sliced>();
    S->kind = 2;
    S->W = 0; S->RW = kSellC; S->cw = 1;
    S->nslices = (int)sptr.size() - 1;
    S->nrows = A.nbr;
    FH_CHECK(A.bs_r == 1 && A.bs_c == 1 && A.nbr == A.nbc && total == (int64_t)sptr.back() * kSellC, -10,
             "SELL layout: square scalar operators");
    S->chunk_ptr.alloc(sizeof(int) * sptr.size());
    FH_HIP(hipMemcpyAsync(S->chunk_ptr.p, sptr.data(), sizeof(int) * sptr.size(), hipMemcpyHostToDevice, c.stream));
    S->cols.alloc(sizeof(int) * std::max<int64_t>(1, total));
    S->vals.alloc(prec_bytes(A.prec) * std::max<int64_t>(1, total));
    if (A.prec == Prec::F64) sell_convert<double>(c, A, *S, total);
    else if (A.prec == Prec::F32) sell_convert<float>(c, A, *S, total);
    else sell_convert<bf16_t>(c, A, *S, total);
    c.sync();  // the host slot table goes out of scope with the caller
    A.sliced = std::move(S);
    A.sliced_state = 1;
}

void wide_build(Ctx& c, const Bcsr& A, int W, int cw, int64_t nchunks) {
    auto S = std::make_unique<Sliced>();
    const int nf = A.bs_r;
    S->kind = 1;
    S->W = W; S->RW = 1; S->cw = cw;
    S->nslices = (int)nchunks;
    std::vector<int> cp(A.nbr + 1, 0), crow((size_t)nchunks), hc((size_t)nchunks * W, -1);
    for (int i = 0; i < A.nbr; i++) {
        const int len = A.h_rowptr[i + 1] - A.h_rowptr[i];
        cp[i + 1] = cp[i] + (len + W - 1) / W;
        for (int q = cp[i]; q < cp[i + 1]; q++) crow[q] = i;
        for (int k = 0; k < len; k++) hc[(size_t)cp[i] * W + k] = A.h_colidx[A.h_rowptr[i] + k];
    }
    FH_CHECK(cp[A.nbr] == nchunks && S->act(nf) <= 64 && nf <= 64, -10, "wide layout sizes");
    S->chunk_ptr.alloc(sizeof(int) * cp.size());
    S->cols.alloc(sizeof(int) * hc.size());
    DevBuf d_crow(sizeof(int) * crow.size());
    FH_HIP(hipMemcpyAsync(S->chunk_ptr.p, cp.data(), sizeof(int) * cp.size(), hipMemcpyHostToDevice, c.stream));
    FH_HIP(hipMemcpyAsync(S->cols.p, hc.data(), sizeof(int) * hc.size(), hipMemcpyHostToDevice, c.stream));
    FH_HIP(hipMemcpyAsync(d_crow.p, crow.data(), sizeof(int) * crow.size(), hipMemcpyHostToDevice, c.stream));
    // layout [chunk][i < nf][lane < W * JP][cw]
    const size_t nvals = (size_t)nchunks * nf * S->act(nf) * cw;
    S->vals.alloc(nvals * prec_bytes(A.prec));
    FH_CHECK(S->vals.bytes == nvals * prec_bytes(A.prec), -10, "wide layout sizes");
    const int* cr = d_crow.as<int>();
    switch (nf) {
        case 1: wide_convert_prec<1>(c, A, *S, cr); break;
        case 2: wide_convert_prec<2>(c, A, *S, cr); break;
        case 3: wide_convert_prec<3>(c, A, *S, cr); break;
        case 4: wide_convert_prec<4>(c, A, *S, cr); break;
        case 6: wide_convert_prec<6>(c, A, *S, cr); break;
        default: wide_convert_prec<10>(c, A, *S, cr); break;
    }
    c.sync();  // host arrays and the chunk-row map released on return
    A.sliced = std::move(S);
    A.sliced_state = 1;
}

template <int NF, class TA>
void sliced_convert(Ctx& c, const Bcsr& A, Sliced& S) {
    const int64_t n = (int64_t)S.nslices * S.act(NF);
    if (n == 0) return;
    hipLaunchKernelGGL((k_to_sliced<NF, TA>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, S.nrows,
                       S.nslices, S.W, S.RW, A.rowptr.as<int>(), A.vals.as<TA>(), S.vals.as<TA>(),
                       S.rowmap.as<int>());
    FH_HIP(hipGetLastError());
}

template <int NF>
void sliced_convert_prec(Ctx& c, const Bcsr& A, Sliced& S) {
    if (A.prec == Prec::F64) sliced_convert<NF, double>(c, A, S);
    else if (A.prec == Prec::F32) sliced_convert<NF, float>(c, A, S);
    else sliced_convert<NF, bf16_t>(c, A, S);
}

// One part of the sliced layout over `rows` (ascending operator rows; null:
// all rows, no row map), W = its widest row.
std::unique_ptr<Sliced> sliced_part(Ctx& c, const Bcsr& A, int W, int RW, int cw, const std::vector<int>* rows) {
    auto S = std::make_unique<Sliced>();
    const int nf = A.bs_r;
    S->W = W; S->RW = RW; S->cw = cw;
    S->nrows = rows ? (int)rows->size() : A.nbr;
    S->nslices = (S->nrows + RW - 1) / RW;
    std::vector<int> hc((size_t)S->nslices * RW * W, -1);
    for (int k = 0; k < S->nrows; k++) {
        const int i = rows ? (*rows)[k] : k;
        FH_CHECK(A.h_rowptr[i + 1] - A.h_rowptr[i] <= W, -10, "sliced part narrower than a row");
        for (int p = A.h_rowptr[i]; p < A.h_rowptr[i + 1]; p++)
            hc[(size_t)(k / RW) * RW * W + (size_t)(k % RW) * W + (p - A.h_rowptr[i])] = A.h_colidx[p];
    }
    S->cols.alloc(sizeof(int) * hc.size());
    FH_HIP(hipMemcpyAsync(S->cols.p, hc.data(), sizeof(int) * hc.size(), hipMemcpyHostToDevice, c.stream));
    if (rows) {
        S->rowmap.alloc(sizeof(int) * std::max<size_t>(1, rows->size()));
        FH_HIP(hipMemcpyAsync(S->rowmap.p, rows->data(), sizeof(int) * rows->size(), hipMemcpyHostToDevice, c.stream));
    }
    // layout [slice][i < nf][lane < ACT][cw]: k_to_sliced writes and
    // k_sliced reads indices below nslices * nf * ACT * cw
    const size_t nvals = (size_t)S->nslices * nf * S->act(nf) * cw;
    S->vals.alloc(nvals * prec_bytes(A.prec));
    FH_CHECK(S->vals.bytes == nvals * prec_bytes(A.prec) && hc.size() == (size_t)S->nslices * RW * W &&
                 RW * nf <= 64 && S->act(nf) <= 64,
             -10, "sliced layout sizes");
    switch (nf) {
        case 1: sliced_convert_prec<1>(c, A, *S); break;
        case 2: sliced_convert_prec<2>(c, A, *S); break;
        case 3: sliced_convert_prec<3>(c, A, *S); break;
        case 4: sliced_convert_prec<4>(c, A, *S); break;
        case 6: sliced_convert_prec<6>(c, A, *S); break;
        default: sliced_convert_prec<10>(c, A, *S); break;
    }
    c.sync();  // host arrays released on return
    return S;
}

// Rows per wavefront for rows of at most W blocks.
int sliced_rw(int nf, int cw, int W) { return std::min(64 / (W * (nf / cw)), 64 / nf); }

void sliced_build(Ctx& c, const Bcsr& A, int W, int RW, int cw) {
    A.sliced = sliced_part(c, A, W, RW, cw, nullptr);
    A.sliced_state = 1;
}

// Rows split by length into two classes (length <= t and > t), each part as
// wide as its widest row.  Returns the padded slot count of the best split
// (t in *t_best), or -1 when no split has two nonempty classes.
int64_t sliced_split(const Bcsr& A, int cw, int& t_best) {
    const int nf = A.bs_r;
    int W = 0;
    for (int i = 0; i < A.nbr; i++) W = std::max(W, A.h_rowptr[i + 1] - A.h_rowptr[i]);
    std::vector<int64_t> cnt(W + 1, 0);
    for (int i = 0; i < A.nbr; i++) cnt[A.h_rowptr[i + 1] - A.h_rowptr[i]]++;
    int64_t best = -1;
    for (int t = 1; t < W; t++) {
        int64_t n1 = 0, n2 = 0;
        int w1 = 0;
        for (int l = 0; l <= t; l++)
            if (cnt[l]) { n1 += cnt[l]; w1 = std::max(w1, l); }
        for (int l = t + 1; l <= W; l++) n2 += cnt[l];
        if (n1 == 0 || n2 == 0 || w1 == 0) continue;
        const int rw1 = sliced_rw(nf, cw, w1), rw2 = sliced_rw(nf, cw, W);
        const int64_t slots = (n1 + rw1 - 1) / rw1 * rw1 * w1 + (n2 + rw2 - 1) / rw2 * rw2 * W;
        if (best < 0 || slots < best) { best = slots; t_best = t; }
    }
    return best;
}

// FHGPU_SLICED=0 keeps every product on the block-CSR kernels.
// Counts changes of any matrix's execution layout: a captured graph whose
// kernels read a layout copy is replayed only while the count is unchanged.
std::atomic<uint64_t> g_layout_epoch{0};

// FHGPU_SELL=0 keeps scalar operators on the other layouts.
bool sell_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("FHGPU_SELL");
        return e ? std::atoi(e) != 0 : true;
    }();
    return on;
}

bool sliced_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("FHGPU_SLICED");
        return e ? std::atoi(e) != 0 : true;
    }();
    return on;
}

// First full-matrix product on A: use the sliced layout when the rows fit a
// wavefront, the matrix is large enough for the layout to pay and padding
// adds at most 25% to the values.  Not decided while the stream is being
// captured into a graph (no allocation there); the block-CSR kernels run
// meanwhile.
void sliced_decide(Ctx& c, const Bcsr& A) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FH_HIP(hipStreamIsCapturing(c.stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return;
    g_layout_epoch.fetch_add(1);
    int W, RW, cw;
    if (!sliced_enabled() || A.nbr < 4096) {
        A.sliced_state = -1;
        return;
    }
    // scalar square operators (the aggregation levels): SELL-64 when its
    // padding stays within half the entries
    if (A.bs_r == 1 && A.bs_c == 1 && A.nbr == A.nbc && A.h_rowptr.size() == (size_t)A.nbr + 1 && sell_enabled()) {
        std::vector<int> sptr;
        const int64_t total = sell_geometry(A, sptr);
        if ((double)total <= 1.5 * (double)A.nnzb) {
            sell_build(c, A, sptr, total);
            return;
        }
    }
    if (sliced_geometry(A, W, RW, cw)) {
        const int64_t nslices = (A.nbr + RW - 1) / RW;
        if ((double)nslices * RW * W <= 1.25 * (double)A.nnzb) {
            sliced_build(c, A, W, RW, cw);
            return;
        }
        // uneven rows (the prolongations: 1-7 blocks): two parts, short and
        // long rows, each as wide as its widest row, on operators large
        // enough for the second launch to pay
        constexpr int kSplitMinRows = 32768;
        int t = 0;
        const int64_t slots = A.nbr >= kSplitMinRows ? sliced_split(A, cw, t) : -1;
        if (slots < 0 || (double)slots > 1.25 * (double)A.nnzb) {
            A.sliced_state = -1;
            return;
        }
        std::vector<int> shortr, longr;
        int w1 = 0;
        for (int i = 0; i < A.nbr; i++) {
            const int len = A.h_rowptr[i + 1] - A.h_rowptr[i];
            if (len <= t) { shortr.push_back(i); w1 = std::max(w1, len); }
            else longr.push_back(i);
        }
        auto first = sliced_part(c, A, w1, sliced_rw(A.bs_r, cw, w1), cw, &shortr);
        first->next = sliced_part(c, A, W, RW, cw, &longr);
        A.sliced = std::move(first);
        A.sliced_state = 1;
        return;
    }
    // rows wider than a wavefront: the chunked layout, under the same bound
    // on padding.  BF16 storage with odd n_f stays on the block-CSR kernel
    // (2-byte value loads either way).
    int64_t nchunks;
    if ((A.prec != Prec::BF16 || A.bs_r % 2 == 0) && wide_geometry(A, W, cw, nchunks) &&
        (double)nchunks * W <= 1.25 * (double)A.nnzb)
        wide_build(c, A, W, cw, nchunks);
    else
        A.sliced_state = -1;
}

}  // namespace

bool set_sliced(Ctx& c, const Bcsr& A, bool on) {
    if (!on) {
        if (A.sliced) g_layout_epoch.fetch_add(1);  // captured graphs may hold the dropped copy
        A.sliced.reset();
        A.sliced_state = -1;
        return false;
    }
    if (A.sliced_state == 1) return true;
    int W, RW, cw;
    g_layout_epoch.fetch_add(1);
    if (sliced_geometry(A, W, RW, cw)) {
        sliced_build(c, A, W, RW, cw);
        return true;
    }
    int64_t nchunks;
    if (wide_geometry(A, W, cw, nchunks)) {
        wide_build(c, A, W, cw, nchunks);
        return true;
    }
    A.sliced_state = -1;
    return false;
}

uint64_t layout_epoch() { return g_layout_epoch.load(); }
void layout_epoch_bump() { g_layout_epoch.fetch_add(1); }

// ------------------------------------------------ allocation-free scratch
// Device memory of the allocation-free entry points: one array per GPU the
// library runs on, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): the address of each device's copy is looked up once, at context
// creation with that device current, and only that address is used.
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_scratch_init(int device) {
    FH_CHECK(device >= 0 && device < kMaxDevices, -2, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    int cur = -1;
    FH_HIP(hipGetDevice(&cur));
    FH_CHECK(cur == device, -3, "the scratch address must be looked up with its device current");
    void* ptr = nullptr;
    FH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    FH_CHECK(ptr != nullptr, -3, "no device address for the scratch array");
    g_scratch_ptr[device] = ptr;
}

void* noalloc_scratch(int device) {
    FH_CHECK(device >= 0 && device < kMaxDevices, -2, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    FH_CHECK(g_scratch_ptr[device] != nullptr, -3, "scratch array not set up on this device (create a context first)");
    return g_scratch_ptr[device];
}

static void launch_bcsr_impl(Ctx& c, const Bcsr& A, BcsrMode mode, Prec vprec, const void* x, const void* b, void* y,
                             double alpha, double beta, double omega, const void* invD, Prec dprec, const int* rows,
                             int nrows, void* aux);

void launch_bcsr(Ctx& c, const Bcsr& A, BcsrMode mode, Prec vprec, const void* x, const void* b, void* y,
                 double alpha, double beta, double omega, const void* invD, Prec dprec, const int* rows,
                 int nrows) {
    // full products and residuals on the element layout when A has an
    // active one; otherwise on the sliced copy or the block rows
    const bool elem = !rows && A.elem_on && (mode == OP_SPMV || mode == OP_RESID);
    if (!elem && !rows && A.sliced_state == 0 && !c.no_alloc) sliced_decide(c, A);
    const bool prof = c.prof_matrix == (const void*)&A && c.prof_mode == (int)mode && !rows && !c.no_alloc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof) {
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        FH_HIP(hipEventRecord(e0, c.stream));
    }
    if (!(elem && launch_elem(c, A, mode, vprec, x, b, y, alpha, beta, nullptr, dprec, nullptr))) {
        if (elem && A.sliced_state == 0 && !c.no_alloc) sliced_decide(c, A);
        launch_bcsr_impl(c, A, mode, vprec, x, b, y, alpha, beta, omega, invD, dprec, rows, nrows, nullptr);
    }
    if (prof) {
        FH_HIP(hipEventRecord(e1, c.stream));
        c.prof_events.push_back({e0, e1});
    }
}

void launch_bcsr_cheb(Ctx& c, const Bcsr& A, Prec vprec, const void* x, const void* b, void* xnew, void* d,
                      double a, double cc, const void* invD, Prec dprec) {
    FH_CHECK(xnew != x && d != nullptr && d != x && d != xnew, -3, "Chebyshev step: x, xnew and d must differ");
    const bool elem = A.elem_on;
    if (!elem && A.sliced_state == 0 && !c.no_alloc) sliced_decide(c, A);
    const bool prof = c.prof_matrix == (const void*)&A && c.prof_mode == (int)OP_CHEB && !c.no_alloc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof) {
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        FH_HIP(hipEventRecord(e0, c.stream));
    }
    if (!(elem && launch_elem(c, A, OP_CHEB, vprec, x, b, xnew, a, cc, invD, dprec, d))) {
        if (elem && A.sliced_state == 0 && !c.no_alloc) sliced_decide(c, A);
        launch_bcsr_impl(c, A, OP_CHEB, vprec, x, b, xnew, a, cc, 1.0, invD, dprec, nullptr, 0, d);
    }
    if (prof) {
        FH_HIP(hipEventRecord(e1, c.stream));
        c.prof_events.push_back({e0, e1});
    }
}

// y = A x with FP32-stored values and FP64 vectors (FP64 accumulation), on
// the one-part sliced copy (n_f >= 2) or the SELL-64 copy (n_f = 1).
template <int NF>
static bool spmv_f32vals_nf(Ctx& c, const Bcsr& A, const void* x, void* y) {
    const Sliced& S = *A.sliced;
    if constexpr (NF == 1) {
        if (S.kind != 2) return false;
        const int nwg = (S.nslices + kWG / 64 - 1) / (kWG / 64);
        hipLaunchKernelGGL((k_sell<float, double, double, OP_SPMV>), dim3(nwg), dim3(kWG), 0, c.stream, A.nbr,
                           S.nslices, S.chunk_ptr.as<int>(), S.cols.as<int>(), S.vals.as<float>(), (const double*)x,
                           (const double*)nullptr, (double*)y, 1.0, 0.0, 0.0, (const double*)nullptr,
                           (double*)nullptr);
    } else {
        if (S.kind != 0 || S.next || S.rowmap.p) return false;
        if (S.cw == 2 && (uintptr_t)x % 16 != 0) return false;
        const int nwg = (S.nslices + kSlicedWaves - 1) / kSlicedWaves;
        hipLaunchKernelGGL((k_sliced<NF, float, double, double, OP_SPMV, false>), dim3(nwg), dim3(kWG), 0, c.stream,
                           S.nrows, S.nslices, S.W, S.RW, S.cols.as<int>(), S.vals.as<float>(), (const double*)x,
                           (const double*)nullptr, (double*)y, 1.0, 0.0, 0.0, (const double*)nullptr,
                           (double*)nullptr, (const int*)nullptr);
    }
    FH_HIP(hipGetLastError());
    return true;
}

bool launch_spmv_f32vals(Ctx& c, const Bcsr& A, const void* x, void* y) {
    if (A.prec != Prec::F32 || A.bs_r != A.bs_c) return false;
    const bool elem = A.elem_on && A.elem->prec == Prec::F32;
    if (!elem && A.sliced_state == 0 && !c.no_alloc) sliced_decide(c, A);
    if (!elem && A.sliced_state != 1) return false;
    const bool prof = c.prof_matrix == (const void*)&A && c.prof_mode == (int)OP_SPMV && !c.no_alloc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof) {
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        FH_HIP(hipEventRecord(e0, c.stream));
    }
    bool done = elem && launch_elem(c, A, OP_SPMV, Prec::F64, x, nullptr, y, 1.0, 0.0, nullptr, Prec::F64, nullptr);
    if (!done && A.sliced_state == 0 && !c.no_alloc) sliced_decide(c, A);
    if (!done && A.sliced_state == 1) switch (A.bs_r) {
        case 1: done = spmv_f32vals_nf<1>(c, A, x, y); break;
        case 2: done = spmv_f32vals_nf<2>(c, A, x, y); break;
        case 3: done = spmv_f32vals_nf<3>(c, A, x, y); break;
        case 4: done = spmv_f32vals_nf<4>(c, A, x, y); break;
        case 6: done = spmv_f32vals_nf<6>(c, A, x, y); break;
        case 10: done = spmv_f32vals_nf<10>(c, A, x, y); break;
        default: done = false; break;
    }
    if (prof) {
        FH_HIP(hipEventRecord(e1, c.stream));
        if (done) c.prof_events.push_back({e0, e1});
        else {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
    }
    return done;
}

template <int NF, class TA, class TV>
static void resid_head_launch(Ctx& c, const Sliced& S, int nbr, const void* x, const void* b, void* rc) {
    const int waves = (S.nslices + kHeadSlices - 1) / kHeadSlices;
    const int nwg = (waves + kSlicedWaves - 1) / kSlicedWaves;
    hipLaunchKernelGGL((k_sliced_resid_head<NF, TA, TV>), dim3(nwg), dim3(kWG), 0, c.stream, nbr, S.nslices, S.W,
                       S.RW, S.cols.as<int>(), S.vals.as<TA>(), (const TV*)x, (const TV*)b, (TV*)rc);
    FH_HIP(hipGetLastError());
}

template <int NF>
static bool resid_head_nf(Ctx& c, const Bcsr& A, Prec vp, const void* x, const void* b, void* rc) {
    const Sliced& S = *A.sliced;
    if (vp == Prec::F64 && A.prec == Prec::F64)
        resid_head_launch<NF, double, double>(c, S, A.nbr, x, b, rc);
    else if (vp == Prec::F32 && A.prec == Prec::F32)
        resid_head_launch<NF, float, float>(c, S, A.nbr, x, b, rc);
    else if (vp == Prec::F32 && A.prec == Prec::BF16)
        resid_head_launch<NF, bf16_t, float>(c, S, A.nbr, x, b, rc);
    else
        return false;
    return true;
}

bool launch_resid_head(Ctx& c, const Bcsr& A, Prec vp, const void* x, const void* b, void* rc) {
    if (A.sliced_state == 0 && !c.no_alloc) sliced_decide(c, A);
    if (A.sliced_state != 1 || A.sliced->kind != 0 || A.sliced->next || A.sliced->rowmap.p || A.bs_r != A.bs_c)
        return false;
    if (A.sliced->cw == 2 && (uintptr_t)x % (2 * prec_bytes(vp)) != 0) return false;
    const bool prof = c.prof_matrix == (const void*)&A && c.prof_mode == (int)OP_RESID_HEAD && !c.no_alloc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof) {
        FH_HIP(hipEventCreate(&e0));
        FH_HIP(hipEventCreate(&e1));
        FH_HIP(hipEventRecord(e0, c.stream));
    }
    bool done;
    switch (A.bs_r) {
        case 1: done = resid_head_nf<1>(c, A, vp, x, b, rc); break;
        case 2: done = resid_head_nf<2>(c, A, vp, x, b, rc); break;
        case 3: done = resid_head_nf<3>(c, A, vp, x, b, rc); break;
        case 4: done = resid_head_nf<4>(c, A, vp, x, b, rc); break;
        case 6: done = resid_head_nf<6>(c, A, vp, x, b, rc); break;
        case 10: done = resid_head_nf<10>(c, A, vp, x, b, rc); break;
        default: done = false; break;
    }
    if (prof) {
        FH_HIP(hipEventRecord(e1, c.stream));
        if (done) c.prof_events.push_back({e0, e1});
        else {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
    }
    return done;
}

static void launch_bcsr_impl(Ctx& c, const Bcsr& A, BcsrMode mode, Prec vprec, const void* x, const void* b, void* y,
                             double alpha, double beta, double omega, const void* invD, Prec dprec, const int* rows,
                             int nrows, void* aux) {
    FH_CHECK(A.bs_r == A.bs_c, -3, "block-row kernel needs square blocks");
    switch (A.bs_r) {
        case 1: bcsr_dispatch_types<1>(c, A, mode, vprec, x, b, y, alpha, beta, omega, invD, dprec, rows, nrows, aux); break;
        case 2: bcsr_dispatch_types<2>(c, A, mode, vprec, x, b, y, alpha, beta, omega, invD, dprec, rows, nrows, aux); break;
        case 3: bcsr_dispatch_types<3>(c, A, mode, vprec, x, b, y, alpha, beta, omega, invD, dprec, rows, nrows, aux); break;
        case 4: bcsr_dispatch_types<4>(c, A, mode, vprec, x, b, y, alpha, beta, omega, invD, dprec, rows, nrows, aux); break;
        case 6: bcsr_dispatch_types<6>(c, A, mode, vprec, x, b, y, alpha, beta, omega, invD, dprec, rows, nrows, aux); break;
        case 10: bcsr_dispatch_types<10>(c, A, mode, vprec, x, b, y, alpha, beta, omega, invD, dprec, rows, nrows, aux); break;
        default: throw Error(-3, "unsupported block size " + std::to_string(A.bs_r) + " (1,2,3,4,6,10)");
    }
}

template <int NF>
static bool bjac_zero_tile_nf(Ctx& c, int nbr, Prec vp, const void* b, void* y, double omega, const void* invD,
                              Prec dp, void* y2) {
    constexpr int R = kWG / NF;
    const int g = (nbr + R - 1) / R;
    if (vp == Prec::F64 && dp == Prec::F64)
        hipLaunchKernelGGL((k_bjac_zero_tile<NF, double, double>), dim3(g), dim3(kWG), 0, c.stream, nbr,
                           (const double*)b, (double*)y, omega, (const double*)invD, (double*)y2);
    else if (vp == Prec::F32 && dp == Prec::F32)
        hipLaunchKernelGGL((k_bjac_zero_tile<NF, float, float>), dim3(g), dim3(kWG), 0, c.stream, nbr,
                           (const float*)b, (float*)y, omega, (const float*)invD, (float*)y2);
    else
        return false;
    FH_HIP(hipGetLastError());
    return true;
}

// The LDS-tiled zero-guess pass over all rows, for the block sizes of the
// block-row kernels; false if it does not apply.
static bool bjac_zero_tile(Ctx& c, int nbr, int nf, Prec vp, const void* b, void* y, double omega,
                           const void* invD, Prec dp, void* y2) {
    switch (nf) {
        case 1: return bjac_zero_tile_nf<1>(c, nbr, vp, b, y, omega, invD, dp, y2);
        case 2: return bjac_zero_tile_nf<2>(c, nbr, vp, b, y, omega, invD, dp, y2);
        case 3: return bjac_zero_tile_nf<3>(c, nbr, vp, b, y, omega, invD, dp, y2);
        case 4: return bjac_zero_tile_nf<4>(c, nbr, vp, b, y, omega, invD, dp, y2);
        case 6: return bjac_zero_tile_nf<6>(c, nbr, vp, b, y, omega, invD, dp, y2);
        case 10: return bjac_zero_tile_nf<10>(c, nbr, vp, b, y, omega, invD, dp, y2);
        default: return false;
    }
}

void launch_bjac_zero(Ctx& c, int nbr, int nf, Prec vp, const void* b, void* y, double omega, const void* invD,
                      Prec dp, const int* rows, int nrows, void* y2) {
    int count = rows ? nrows : nbr;
    if (count <= 0) return;
    if (!rows && (uintptr_t)invD % 16 == 0 && bjac_zero_tile(c, nbr, nf, vp, b, y, omega, invD, dp, y2)) return;
    int64_t nt = (int64_t)count * nf;
    int g = (int)((nt + 255) / 256);
    const size_t vb = 2 * prec_bytes(vp), db = 2 * prec_bytes(dp);
    const int vec2 = nf % 2 == 0 && (uintptr_t)b % vb == 0 && (uintptr_t)invD % db == 0;
    if (vp == Prec::F64 && dp == Prec::F64)
        hipLaunchKernelGGL((k_bjac_zero<double, double>), dim3(g), dim3(256), 0, c.stream, count, nf, (const double*)b,
                           (double*)y, omega, (const double*)invD, rows, (double*)y2, vec2);
    else if (vp == Prec::F32 && dp == Prec::F32)
        hipLaunchKernelGGL((k_bjac_zero<float, float>), dim3(g), dim3(256), 0, c.stream, count, nf, (const float*)b,
                           (float*)y, omega, (const float*)invD, rows, (float*)y2, vec2);
    else
        throw Error(-3, "unsupported precision pair for block Jacobi");
    FH_HIP(hipGetLastError());
}

void launch_block_inverse(Ctx& c, const Bcsr& A, const int* diag, void* invD, Prec dp, int* d_flag) {
    FH_CHECK(A.prec == Prec::F64, -3, "block inverse reads FP64 values");
    FH_CHECK(A.bs_r <= 10, -3, "block size above 10");
    int g = (A.nbr + 127) / 128;
    if (A.nbr == 0) return;
    if (dp == Prec::F64)
        hipLaunchKernelGGL((k_block_inverse<double>), dim3(g), dim3(128), 0, c.stream, A.nbr, A.bs_r,
                           A.vals.as<double>(), diag, (double*)invD, d_flag);
    else if (dp == Prec::F32)
        hipLaunchKernelGGL((k_block_inverse<float>), dim3(g), dim3(128), 0, c.stream, A.nbr, A.bs_r,
                           A.vals.as<double>(), diag, (float*)invD, d_flag);
    else
        throw Error(-3, "block inverses are stored in FP64 or FP32");
    FH_HIP(hipGetLastError());
}

void launch_convert(Ctx& c, int64_t n, const void* src, Prec sp, void* dst, Prec dp) {
    if (n <= 0) return;
    int g = grid_for(n);
#define FH_CONV(TS, TD)                                                                                  \
    hipLaunchKernelGGL((k_convert<TS, TD>), dim3(g), dim3(256), 0, c.stream, n, (const TS*)src, (TD*)dst); \
    break
    if (sp == Prec::F64) {
        switch (dp) {
            case Prec::F64: FH_CONV(double, double);
            case Prec::F32: FH_CONV(double, float);
            case Prec::BF16: FH_CONV(double, bf16_t);
        }
    } else if (sp == Prec::F32) {
        switch (dp) {
            case Prec::F64: FH_CONV(float, double);
            case Prec::F32: FH_CONV(float, float);
            case Prec::BF16: FH_CONV(float, bf16_t);
        }
    } else {
        switch (dp) {
            case Prec::F64: FH_CONV(bf16_t, double);
            case Prec::F32: FH_CONV(bf16_t, float);
            case Prec::BF16: FH_CONV(bf16_t, bf16_t);
        }
    }
#undef FH_CONV
    FH_HIP(hipGetLastError());
}

void launch_fill(Ctx& c, int64_t n, void* x, Prec p, double v) {
    if (n <= 0) return;
    int g = grid_for(n);
    if (p == Prec::F64) hipLaunchKernelGGL((k_fill<double>), dim3(g), dim3(256), 0, c.stream, n, (double*)x, v);
    else if (p == Prec::F32) hipLaunchKernelGGL((k_fill<float>), dim3(g), dim3(256), 0, c.stream, n, (float*)x, v);
    else hipLaunchKernelGGL((k_fill<bf16_t>), dim3(g), dim3(256), 0, c.stream, n, (bf16_t*)x, v);
    FH_HIP(hipGetLastError());
}

void launch_axpby(Ctx& c, int64_t n, Prec p, const void* x, void* y, double ha, int ia, int ib, double hb, int ic,
                  int id) {
    if (n <= 0) return;
    int g = grid_for(n);
    if (p == Prec::F64)
        hipLaunchKernelGGL((k_axpby<double>), dim3(g), dim3(256), 0, c.stream, n, (const double*)x, (double*)y,
                           c.d_scalars, ha, ia, ib, hb, ic, id);
    else
        hipLaunchKernelGGL((k_axpby<float>), dim3(g), dim3(256), 0, c.stream, n, (const float*)x, (float*)y,
                           c.d_scalars, ha, ia, ib, hb, ic, id);
    FH_HIP(hipGetLastError());
}

void launch_fcg_orth(Ctx& c, int64_t n, const double* z, int nv, const double* const* vs, const int* num,
                     const int* den, double* out) {
    if (n <= 0) return;
    const int g = grid_for(n);
    if (nv == 0) {
        if (out != z) FH_HIP(hipMemcpyAsync(out, z, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
        return;
    }
    for (int k0 = 0; k0 < nv; k0 += kOrthVecs) {
        const int cnt = std::min(kOrthVecs, nv - k0);
        OrthArgs a{};
        for (int k = 0; k < cnt; k++) {
            a.v[k] = vs[k0 + k];
            a.num[k] = num[k0 + k];
            a.den[k] = den[k0 + k];
        }
        const double* src = k0 == 0 ? z : out;
        switch (cnt) {
            case 1: hipLaunchKernelGGL((k_fcg_orth<1>), dim3(g), dim3(256), 0, c.stream, n, src, a, c.d_scalars, out); break;
            case 2: hipLaunchKernelGGL((k_fcg_orth<2>), dim3(g), dim3(256), 0, c.stream, n, src, a, c.d_scalars, out); break;
            case 3: hipLaunchKernelGGL((k_fcg_orth<3>), dim3(g), dim3(256), 0, c.stream, n, src, a, c.d_scalars, out); break;
            default: hipLaunchKernelGGL((k_fcg_orth<4>), dim3(g), dim3(256), 0, c.stream, n, src, a, c.d_scalars, out); break;
        }
        FH_HIP(hipGetLastError());
    }
}

void launch_cg_update_rr(Ctx& c, int64_t n, double* x, const double* d, double* r, const double* q, int inum,
                         int iden, double sign, int slot_rr) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cg_update_rr, dim3(kDotBlocks), dim3(kWG), 0, c.stream, n, x, d, r, q, c.d_scalars, inum, iden,
                       sign, c.d_partials);
    hipLaunchKernelGGL((k_dots_final<1>), dim3(1), dim3(kDotBlocks), 0, c.stream, c.d_partials, c.d_scalars, slot_rr, 0,
                       0);
    FH_HIP(hipGetLastError());
}

void launch_cg_update(Ctx& c, int64_t n, Prec p, void* x, const void* d, void* r, const void* q, int inum, int iden,
                      double sign) {
    if (n <= 0) return;
    int g = grid_for(n);
    if (p == Prec::F64)
        hipLaunchKernelGGL((k_cg_update<double>), dim3(g), dim3(256), 0, c.stream, n, (double*)x, (const double*)d,
                           (double*)r, (const double*)q, c.d_scalars, inum, iden, sign);
    else
        hipLaunchKernelGGL((k_cg_update<float>), dim3(g), dim3(256), 0, c.stream, n, (float*)x, (const float*)d,
                           (float*)r, (const float*)q, c.d_scalars, inum, iden, sign);
    FH_HIP(hipGetLastError());
}

void launch_dots(Ctx& c, int64_t n, Prec p, int nd, const void* const* xs, const void* const* ys, const int* slots) {
    FH_CHECK(nd >= 1 && nd <= 3, -3, "1 to 3 dot products per launch");
    const void* X[3] = {xs[0], nd > 1 ? xs[1] : xs[0], nd > 2 ? xs[2] : xs[0]};
    const void* Y[3] = {ys[0], nd > 1 ? ys[1] : ys[0], nd > 2 ? ys[2] : ys[0]};
    const bool sx = X[1] == X[0] && X[2] == X[0];
#define FH_DOTS(T, ND)                                                                                              \
    do {                                                                                                            \
        if (sx)                                                                                                     \
            hipLaunchKernelGGL((k_dots<T, ND, true>), dim3(kDotBlocks), dim3(kWG), 0, c.stream, n, (const T*)X[0],   \
                               (const T*)Y[0], (const T*)X[1], (const T*)Y[1], (const T*)X[2], (const T*)Y[2],       \
                               c.d_partials);                                                                       \
        else                                                                                                        \
            hipLaunchKernelGGL((k_dots<T, ND, false>), dim3(kDotBlocks), dim3(kWG), 0, c.stream, n, (const T*)X[0],  \
                               (const T*)Y[0], (const T*)X[1], (const T*)Y[1], (const T*)X[2], (const T*)Y[2],       \
                               c.d_partials);                                                                       \
    } while (0)
    if (p == Prec::F64) {
        if (nd == 1) FH_DOTS(double, 1); else if (nd == 2) FH_DOTS(double, 2); else FH_DOTS(double, 3);
    } else {
        if (nd == 1) FH_DOTS(float, 1); else if (nd == 2) FH_DOTS(float, 2); else FH_DOTS(float, 3);
    }
#undef FH_DOTS
    int s0 = slots[0], s1 = nd > 1 ? slots[1] : 0, s2 = nd > 2 ? slots[2] : 0;
    if (nd == 1) hipLaunchKernelGGL((k_dots_final<1>), dim3(1), dim3(kDotBlocks), 0, c.stream, c.d_partials, c.d_scalars, s0, s1, s2);
    else if (nd == 2) hipLaunchKernelGGL((k_dots_final<2>), dim3(1), dim3(kDotBlocks), 0, c.stream, c.d_partials, c.d_scalars, s0, s1, s2);
    else hipLaunchKernelGGL((k_dots_final<3>), dim3(1), dim3(kDotBlocks), 0, c.stream, c.d_partials, c.d_scalars, s0, s1, s2);
    FH_HIP(hipGetLastError());
}

void launch_scalar_op(Ctx& c, int op, int dst, int a, int b) {
    hipLaunchKernelGGL(k_scalar_op, dim3(1), dim3(1), 0, c.stream, c.d_scalars, op, dst, a, b);
    FH_HIP(hipGetLastError());
}

void launch_kcycle_coeffs(Ctx& c, int first, double t) {
    hipLaunchKernelGGL(k_kcycle_coeffs, dim3(1), dim3(1), 0, c.stream, c.d_scalars, first, t);
    FH_HIP(hipGetLastError());
}

void launch_agg_restrict(Ctx& c, int nagg, int nf, Prec p, const int* mptr, const int* members, const void* r,
                         void* rc) {
    int64_t nt = (int64_t)nagg * nf;
    if (nt <= 0) return;
    int g = (int)((nt + 255) / 256);
    if (p == Prec::F64)
        hipLaunchKernelGGL((k_agg_restrict<double>), dim3(g), dim3(256), 0, c.stream, nagg, nf, mptr, members,
                           (const double*)r, (double*)rc);
    else
        hipLaunchKernelGGL((k_agg_restrict<float>), dim3(g), dim3(256), 0, c.stream, nagg, nf, mptr, members,
                           (const float*)r, (float*)rc);
    FH_HIP(hipGetLastError());
}

void launch_agg_prolong_add(Ctx& c, int nb, int nf, Prec p, const int* agg, const void* ec, void* x) {
    int64_t nt = (int64_t)nb * nf;
    if (nt <= 0) return;
    int g = (int)((nt + 255) / 256);
    if (p == Prec::F64)
        hipLaunchKernelGGL((k_agg_prolong_add<double>), dim3(g), dim3(256), 0, c.stream, nb, nf, agg,
                           (const double*)ec, (double*)x);
    else
        hipLaunchKernelGGL((k_agg_prolong_add<float>), dim3(g), dim3(256), 0, c.stream, nb, nf, agg,
                           (const float*)ec, (float*)x);
    FH_HIP(hipGetLastError());
}

void launch_p_restrict(Ctx& c, int nb, int nf, int nfc, Prec p, const void* r, void* rc) {
    int64_t nt = (int64_t)nb * nfc;
    if (nt <= 0) return;
    int g = (int)((nt + 255) / 256);
    if (p == Prec::F64)
        hipLaunchKernelGGL((k_p_restrict<double>), dim3(g), dim3(256), 0, c.stream, nb, nf, nfc, (const double*)r,
                           (double*)rc);
    else
        hipLaunchKernelGGL((k_p_restrict<float>), dim3(g), dim3(256), 0, c.stream, nb, nf, nfc, (const float*)r,
                           (float*)rc);
    FH_HIP(hipGetLastError());
}

void launch_p_prolong_add(Ctx& c, int nb, int nf, int nfc, Prec p, const void* ec, void* x) {
    int64_t nt = (int64_t)nb * nfc;
    if (nt <= 0) return;
    int g = (int)((nt + 255) / 256);
    if (p == Prec::F64)
        hipLaunchKernelGGL((k_p_prolong_add<double>), dim3(g), dim3(256), 0, c.stream, nb, nf, nfc,
                           (const double*)ec, (double*)x);
    else
        hipLaunchKernelGGL((k_p_prolong_add<float>), dim3(g), dim3(256), 0, c.stream, nb, nf, nfc, (const float*)ec,
                           (float*)x);
    FH_HIP(hipGetLastError());
}

// Leading nfc x nfc sub-block of every nf x nf block (FP64 values): the
// Galerkin operator R A P of the p-injection P = [I_nfc; 0] per face.
__global__ void k_mode_submatrix(int64_t nnzb, int nf, int nfc, const double* __restrict__ A, double* __restrict__ S) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = nnzb * nfc * nfc;
    if (e >= n) return;
    const int64_t p = e / (nfc * nfc);
    const int r = (int)(e - p * nfc * nfc);
    const int a = r / nfc, b = r - a * nfc;
    S[e] = A[p * nf * nf + a * nf + b];
}

void launch_mode_submatrix(Ctx& c, int64_t nnzb, int nf, int nfc, const double* A, double* S) {
    const int64_t n = nnzb * nfc * nfc;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mode_submatrix, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, nnzb, nf, nfc, A, S);
    FH_HIP(hipGetLastError());
}

void launch_dense_symv(Ctx& c, int n, const void* M, Prec mp, Prec vp, const void* x, void* y) {
    if (n <= 0) return;
    int g = (n + 3) / 4;
    if (mp == Prec::F64 && vp == Prec::F64)
        hipLaunchKernelGGL((k_dense_symv<double, double>), dim3(g), dim3(kWG), 0, c.stream, n, (const double*)M,
                           (const double*)x, (double*)y);
    else if (mp == Prec::F64 && vp == Prec::F32)
        hipLaunchKernelGGL((k_dense_symv<double, float>), dim3(g), dim3(kWG), 0, c.stream, n, (const double*)M,
                           (const float*)x, (float*)y);
    else
        throw Error(-3, "coarse matrix must be FP64");
    FH_HIP(hipGetLastError());
}

}  // namespace fh
