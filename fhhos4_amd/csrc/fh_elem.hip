// Element execution layout of a statically condensed HHO operator: products,
// residuals and Chebyshev steps applied cell by cell from the condensed cell
// matrices S_c kept by the assembly (ElemLayout in fh_common.h).
//
// The face system is A = sum_c P_c^T S_c P_c (Diffusion_HHO.h:633-637 adds
// every cell's condensed matrix into the global one).  A block row of A
// stores a face-pair block once per direction; a cell matrix is symmetric,
// so it is stored once per pair of its unknowns: about 56% of the block
// rows' bytes per product (1860 values per hexahedral cell at k = 3 against
// the 3200 of its 32 block-row slots).
//
// Storage by cyclic diagonals: diagonal t (t = 0 .. m/2) of a cell holds
// S[l][(l + t) mod m] at position l, l = 0 .. m-1.  Every unordered pair
// {i, j} lies on exactly one stored diagonal (t = (j - i) mod m or
// (i - j) mod m, whichever is <= m/2; diagonal m/2 holds both orientations),
// and every stored diagonal is a full run of m values.  Diagonals 2p + 1
// and 2p + 2 are interleaved, so that lane l reads both with one two-value
// load, and the wavefront reads one contiguous run per pair.
//
// 64 / m cells per wavefront (one at k = 2 and 3, three at k = 1, ten at
// k = 0), lane g m + l owning local unknown l of the g-th cell (face l / nf,
// entry l % nf).  For diagonal t lane l reads S[l][l+t] (one contiguous run
// across the wavefront) and adds S[l][l+t] x[l+t], x rotated by t within the
// lane group, and, for 0 < t < m/2, sends S[l][l+t] x[l] to lane l + t (the
// transposed entry): two wavefront shuffles per diagonal.  No LDS, no
// atomics: the cells are split into two colours such that every unknown
// face has one cell of each, the first colour's launch writes the face's
// entries and the second's adds to them, so every entry is the same
// two-term sum on every run.
#include "fh_kernels.h"

#include <cstdlib>
#include <cstring>

namespace fh {
namespace {

constexpr int kElemWaves = 4;  // wavefronts per workgroup
constexpr int kElemFaces = 6;  // faces per cell (hexahedra)

template <class TV> struct ElemAcc { using T = double; };
template <> struct ElemAcc<float> { using T = float; };

template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

typedef double elem_d2 __attribute__((ext_vector_type(2)));
typedef float elem_f2 __attribute__((ext_vector_type(2)));
template <class TA> struct ElemPair { using T = elem_d2; };
template <> struct ElemPair<float> { using T = elem_f2; };

// MODE: OP_SPMV  y = alpha A x + beta y
//       OP_RESID y = b - A x
//       OP_CHEB  d = alpha d + beta D^-1 (b - A x), y = x + d   (d in aux)
// PASS 0 (first colour) writes the unknowns of its cells; PASS 1 (second
// colour) completes them.  In OP_CHEB the first pass leaves b - (S_c x) of
// the first cell in y, which the second pass reads back.
template <int NF, class TA, class TV, class TD, int MODE, int PASS>
__global__ __launch_bounds__(kElemWaves * 64) void k_elem(int ncell, int64_t cell0, const int* __restrict__ frow,
                                                          const TA* __restrict__ vals, int stride, const TV* x,
                                                          const TV* __restrict__ b, TV* y, double alpha,
                                                          double beta, const TD* __restrict__ invD,
                                                          TV* __restrict__ aux) {
    using AT = typename ElemAcc<TV>::T;
    constexpr int M = kElemFaces * NF;
    static_assert(M <= 64 && M % 2 == 0, "a cell's unknowns must fit one wavefront");
    constexpr int G = 64 / M;  // cells per wavefront (lane groups of m)
    constexpr int H = M / 2;   // last stored diagonal
    const int lane = threadIdx.x & 63;
    const int q = lane / M, l = lane - q * M;  // group, local unknown
    // workgroup -> cells: every XCD takes a contiguous range of cells, so
    // faces that share a cache line are fetched into one L2
    int bid = blockIdx.x;
    {
        const int nwg = gridDim.x, xcd = bid & 7, idx = bid >> 3, per = nwg >> 3, rem = nwg & 7;
        bid = xcd * per + (xcd < rem ? xcd : rem) + idx;
    }
    const int64_t w = (int64_t)bid * kElemWaves + (threadIdx.x >> 6);
    if (w * G >= ncell) return;  // wavefront-uniform; no barrier follows
    const bool has = q < G && w * G + q < ncell;
    const int64_t cell = cell0 + (has ? w * G + q : 0);
    const int f = l / NF, i = l - f * NF;
    const int row = has ? frow[cell * kElemFaces + f] : -1;
    const int64_t e = (int64_t)row * NF + i;
    const AT xl = row >= 0 ? (AT)x[e] : (AT)0;
    // the Chebyshev step's row of D^-1, loaded before the product
    TD dinv[MODE == OP_CHEB && PASS == 1 ? NF : 1];
    if constexpr (MODE == OP_CHEB && PASS == 1) {
        const TD* D = invD + (row >= 0 ? e : 0) * NF;
#pragma unroll
        for (int k = 0; k < NF; k++) dinv[k] = D[k];
    }
    // lanes without a cell read inside cell 0's values (l < m) and drop them
    const TA* v = vals + cell * stride;
    using PT = typename ElemPair<TA>::T;
    const PT* vp = (const PT*)(v + M) + l;  // pair p of lane l at vp[p m]
    const int base = (q < G ? q : 0) * M;  // first lane of the group
    AT acc = has ? (AT)ld_stream(v + l) * xl : (AT)0;  // diagonal 0
    // Pairs of diagonals (2p + 1, 2p + 2), p < P, in an even number NCH of
    // chunks of C pairs, two chunks per iteration of a loop that is not
    // unrolled: the loads of one chunk are in flight while the other is
    // used, and the compiler cannot hoist every load and shuffle of the cell
    // (which would leave one wavefront per SIMD).  Loads are unconditional
    // (a slot past the last pair re-reads it and is dropped), so that the
    // loop body is one block.
    constexpr int P = (H + 1) / 2;
    constexpr int NCH = P > 10 ? 8 : 2;
    constexpr int C = (P + NCH - 1) / NCH;
    PT ua[C], ub[C];
    auto load_chunk = [&](int ch, PT* u) {
#pragma unroll
        for (int j = 0; j < C; j++) {
            const int p = ch * C + j;
            u[j] = ld_stream(vp + (p < P ? p : P - 1) * M);
        }
    };
    auto use_diag = [&](int t, AT uj) {
        int src = l + t;
        src = base + (src < M ? src : src - M);
        acc += uj * __shfl(xl, src, 64);
        if (t < H) {
            int dst = l - t;
            dst = base + (dst >= 0 ? dst : dst + M);
            acc += __shfl(uj * xl, dst, 64);
        }
    };
    auto use_chunk = [&](int ch, const PT* u) {
#pragma unroll
        for (int j = 0; j < C; j++) {
            const int p = ch * C + j;
            if (p >= P) break;  // wavefront-uniform
            use_diag(2 * p + 1, has ? (AT)u[j].x : (AT)0);
            if (2 * p + 2 <= H) use_diag(2 * p + 2, has ? (AT)u[j].y : (AT)0);
        }
    };
    load_chunk(0, ua);
#pragma unroll 1
    for (int ch = 0; ch < NCH; ch += 2) {
        load_chunk(ch + 1, ub);
        use_chunk(ch, ua);
        load_chunk(ch + 2 < NCH ? ch + 2 : NCH - 1, ua);  // the last one re-reads cached values
        use_chunk(ch + 1, ub);
    }
    if constexpr (MODE == OP_SPMV) {
        if (row >= 0) {
            if constexpr (PASS == 0) {
                AT o = (AT)alpha * acc;
                if (beta != 0.0) o += (AT)beta * (AT)y[e];
                y[e] = (TV)o;
            } else {
                y[e] = (TV)((AT)y[e] + (AT)alpha * acc);
            }
        }
    } else if constexpr (MODE == OP_RESID) {
        if (row >= 0) y[e] = (TV)((PASS == 0 ? (AT)b[e] : (AT)y[e]) - acc);
    } else {
        if constexpr (PASS == 0) {
            if (row >= 0) y[e] = (TV)((AT)b[e] - acc);
        } else {
            // z = D^-1 (b - A x) over the face's nf residuals, held by the
            // face's lanes; every lane takes part in the shuffles
            const AT res = row >= 0 ? (AT)y[e] - acc : (AT)0;
            TD dk[NF];
#pragma unroll
            for (int k = 0; k < NF; k++) dk[k] = row >= 0 ? dinv[k] : (TD)0;
            AT z = 0;
#pragma unroll
            for (int k = 0; k < NF; k++) z += (AT)dk[k] * __shfl(res, base + f * NF + k, 64);
            if (row >= 0) {
                const AT dn = (alpha != 0.0 ? (AT)alpha * (AT)aux[e] : (AT)0) + (AT)beta * z;
                aux[e] = (TV)dn;
                y[e] = (TV)(xl + dn);
            }
        }
    }
}

template <class TI, class TO>
__global__ void k_elem_convert(int64_t n, const TI* __restrict__ in, TO* __restrict__ out) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = (TO)in[t];
}

template <int NF, class TA, class TV, class TD, int MODE>
void elem_two_pass(Ctx& c, const ElemLayout& E, const void* x, const void* b, void* y, double alpha, double beta,
                   const void* invD, void* aux) {
    const int* rows = E.rows->as<int>();
    const TA* vals = E.vals.as<TA>();
    int64_t first = 0;
    for (int q = 0; q < 2; q++) {
        const int n = E.ncol[q];
        if (n > 0) {
            constexpr int G = 64 / (kElemFaces * NF);  // cells per wavefront
            const int64_t waves = (n + G - 1) / G;
            const dim3 grid((unsigned)((waves + kElemWaves - 1) / kElemWaves)), block(kElemWaves * 64);
            if (q == 0)
                hipLaunchKernelGGL((k_elem<NF, TA, TV, TD, MODE, 0>), grid, block, 0, c.stream, n, first, rows, vals,
                                   E.stride, (const TV*)x, (const TV*)b, (TV*)y, alpha, beta, (const TD*)invD,
                                   (TV*)aux);
            else
                hipLaunchKernelGGL((k_elem<NF, TA, TV, TD, MODE, 1>), grid, block, 0, c.stream, n, first, rows, vals,
                                   E.stride, (const TV*)x, (const TV*)b, (TV*)y, alpha, beta, (const TD*)invD,
                                   (TV*)aux);
            FH_HIP(hipGetLastError());
        }
        first += n;
    }
}

template <int NF>
bool elem_dispatch(Ctx& c, const ElemLayout& E, BcsrMode mode, Prec vp, const void* x, const void* b, void* y,
                   double alpha, double beta, const void* invD, Prec dp, void* aux) {
    if (E.prec == Prec::F64 && vp == Prec::F64) {
        if (mode == OP_SPMV)
            elem_two_pass<NF, double, double, double, OP_SPMV>(c, E, x, b, y, alpha, beta, nullptr, nullptr);
        else if (mode == OP_RESID)
            elem_two_pass<NF, double, double, double, OP_RESID>(c, E, x, b, y, alpha, beta, nullptr, nullptr);
        else if (mode == OP_CHEB && dp == Prec::F64)
            elem_two_pass<NF, double, double, double, OP_CHEB>(c, E, x, b, y, alpha, beta, invD, aux);
        else
            return false;
        return true;
    }
    if (E.prec == Prec::F32 && vp == Prec::F32) {
        if (mode == OP_SPMV)
            elem_two_pass<NF, float, float, float, OP_SPMV>(c, E, x, b, y, alpha, beta, nullptr, nullptr);
        else if (mode == OP_RESID)
            elem_two_pass<NF, float, float, float, OP_RESID>(c, E, x, b, y, alpha, beta, nullptr, nullptr);
        else if (mode == OP_CHEB && dp == Prec::F32)
            elem_two_pass<NF, float, float, float, OP_CHEB>(c, E, x, b, y, alpha, beta, invD, aux);
        else
            return false;
        return true;
    }
    if (E.prec == Prec::F32 && vp == Prec::F64 && (mode == OP_SPMV || mode == OP_RESID)) {
        // FP32-stored values with FP64 vectors and accumulation
        if (mode == OP_SPMV)
            elem_two_pass<NF, float, double, double, OP_SPMV>(c, E, x, b, y, alpha, beta, nullptr, nullptr);
        else
            elem_two_pass<NF, float, double, double, OP_RESID>(c, E, x, b, y, alpha, beta, nullptr, nullptr);
        return true;
    }
    return false;
}

}  // namespace

bool elem_enabled() {
    static const bool on = [] {
        const char* s = std::getenv("FHGPU_ELEM");
        return !(s && std::strcmp(s, "0") == 0);
    }();
    return on;
}

bool set_elem(const Bcsr& A, bool on) {
    const bool want = on && A.elem != nullptr;
    if (want != A.elem_on) layout_epoch_bump();  // captured graphs hold the previous launches
    A.elem_on = want;
    return want;
}

bool launch_elem(Ctx& c, const Bcsr& A, BcsrMode mode, Prec vp, const void* x, const void* b, void* y, double alpha,
                 double beta, const void* invD, Prec dp, void* aux) {
    if (!A.elem || !A.elem_on || !elem_enabled()) return false;
    FH_CHECK(x != y, -3, "element-layout product: x must not alias y");
    const ElemLayout& E = *A.elem;
    FH_CHECK(E.nf == A.bs_r && E.nfc == kElemFaces && E.m == E.nfc * E.nf && E.stride == elem_stride(E.m), -3,
             "element layout does not match");
    switch (E.nf) {
        case 1: return elem_dispatch<1>(c, E, mode, vp, x, b, y, alpha, beta, invD, dp, aux);
        case 3: return elem_dispatch<3>(c, E, mode, vp, x, b, y, alpha, beta, invD, dp, aux);
        case 6: return elem_dispatch<6>(c, E, mode, vp, x, b, y, alpha, beta, invD, dp, aux);
        case 10: return elem_dispatch<10>(c, E, mode, vp, x, b, y, alpha, beta, invD, dp, aux);
        default: return false;
    }
}

std::shared_ptr<ElemLayout> elem_convert(Ctx& c, const ElemLayout& E, Prec p) {
    FH_CHECK(E.prec == Prec::F64 && p == Prec::F32, -3, "element layout: FP64 to FP32 conversion only");
    auto F = std::make_shared<ElemLayout>();
    F->nfc = E.nfc; F->nf = E.nf; F->m = E.m; F->stride = E.stride;
    F->ncol[0] = E.ncol[0]; F->ncol[1] = E.ncol[1];
    F->prec = p;
    F->rows = E.rows;
    const int64_t n = E.ncell() * E.stride;
    F->vals.alloc(sizeof(float) * std::max<int64_t>(1, n));
    hipLaunchKernelGGL((k_elem_convert<double, float>), dim3(4096), dim3(256), 0, c.stream, n, E.vals.as<double>(),
                       F->vals.as<float>());
    FH_HIP(hipGetLastError());
    return F;
}

}  // namespace fh
