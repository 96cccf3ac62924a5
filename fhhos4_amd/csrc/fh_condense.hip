// Batched static condensation of caller-supplied HHO element matrices and the
// recovery of the cell unknowns, on gfx950.
//
// Input per cell: the local matrix of the element (cell unknowns first, then
// its faces in element order; Diff_HHOElement::A, Diff_HHOElement.h:144,
// FirstDOFNumber :522-525) and the system number of each face (-1 for a
// face whose unknowns are eliminated, a Dirichlet face).  The face system is
//   A = sum over cells of (A_FF - A_TF^T A_TT^-1 A_TF),
//   b = B_F - sum over cells of A_TF^T A_TT^-1 B_T,
// the reference's A = A_ndF_ndF - A_T_ndF^T Solve_A_T_T(A_T_ndF) and
// CondensedRHS (Diffusion_HHO.h:624-637, :1026-1032) cell by cell, and the
// cell unknowns are x_T = A_TT^-1 (B_T - A_TF x_F) (SolveCellUnknowns,
// Diffusion_HHO.h:960-966).
//
// One workgroup per cell: A_TT, A_TF and A_TT^-1 [A_TF B_T] live in LDS
// (nc^2 + nc m + nc (m + 1) doubles, m = faces x face unknowns: 22 KB for
// hexahedra at k = 3); the Cholesky factorisation runs column by column
// with barriers, the triangular solves one column per thread, and the Schur
// blocks are computed entry by entry and added into the block-CSR values
// with FP64 atomics.  On a conforming mesh an off-diagonal block receives
// one cell's contribution and a diagonal block at most two, so the sums do
// not depend on the order of the additions.  The right-hand side is formed
// without atomics: every cell stores its A_TF^T A_TT^-1 B_T, and every face
// adds those of its cells in cell order.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "fh_condense.h"

namespace fh {
namespace {

constexpr int kCondWG = 256;

struct CondDev {
    int ncells, nfc, nc, nf, nl, m;
    const int* faces;     // ncells x nfc system face numbers, -1: eliminated
    const int* slot;      // ncells x nfc x nfc block index of (face j1, face j2), -1: none
    const double* local;  // ncells x nl x nl
    const double* BT;     // ncells x nc, or null (zero)
    double* vals;         // block values (nf x nf row-major per block)
    double* cstage;       // ncells x m: A_TF^T A_TT^-1 B_T
    int* err;             // bit 0: an A_TT is not positive definite
};

// In-place lower Cholesky factor of the nc x nc matrix L (row-major, LDS).
// Returns false (the same on every thread) if a pivot is not positive.
__device__ bool chol_lds(double* L, int nc) {
    for (int j = 0; j < nc; j++) {
        __syncthreads();
        const double piv = L[j * nc + j];
        if (!(piv > 0.0)) return false;  // uniform: every thread read the same pivot
        const double d = sqrt(piv);
        __syncthreads();
        for (int r = j + 1 + (int)threadIdx.x; r < nc; r += blockDim.x) L[r * nc + j] /= d;
        if (threadIdx.x == 0) L[j * nc + j] = d;
        __syncthreads();
        const int w = nc - j - 1;  // trailing block, lower triangle
        for (int t = threadIdx.x; t < w * w; t += blockDim.x) {
            const int r = j + 1 + t / w, q = j + 1 + t % w;
            if (q <= r) L[r * nc + q] -= L[r * nc + j] * L[q * nc + j];
        }
    }
    __syncthreads();
    return true;
}

// Solves L L^T y = y for column col of the nc x ld array Y (row-major).
__device__ void chol_solve_col(const double* L, int nc, double* Y, int ld, int col) {
    for (int r = 0; r < nc; r++) {
        double s = Y[r * ld + col];
        for (int q = 0; q < r; q++) s -= L[r * nc + q] * Y[q * ld + col];
        Y[r * ld + col] = s / L[r * nc + r];
    }
    for (int r = nc - 1; r >= 0; r--) {
        double s = Y[r * ld + col];
        for (int q = r + 1; q < nc; q++) s -= L[q * nc + r] * Y[q * ld + col];
        Y[r * ld + col] = s / L[r * nc + r];
    }
}

// Schur entries of a cell on MFMA: S = A_FF - A_TF^T X with X = A_TT^-1 A_TF,
// an m x m x nc product cut into 16 x 16 output tiles, one wavefront per
// tile, v_mfma_f64_16x16x4 over k = 4 cell unknowns at a time (zero-padded
// beyond nc and m).  Operand lanes: A[i][k] = T[k][p] with i = lane % 16,
// k = lane / 16; B[k][j] = X[k][q] with j = lane % 16, k = lane / 16; result
// element v of a lane is D[lane / 16 + 4 v][lane % 16] (the layout measured
// on gfx950 with one-hot operands).
typedef double d4_t __attribute__((ext_vector_type(4)));
__device__ void schur_mfma(const CondDev& a, const double* T, const double* X, const double* loc, const int* slot) {
    const int nc = a.nc, m = a.m, nl = a.nl, nf = a.nf, nfc = a.nfc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int tiles = (m + 15) / 16;
    for (int t = wave; t < tiles * tiles; t += nw) {  // wavefront-uniform trip count
        const int ti = t / tiles, tj = t - ti * tiles;
        const int p_op = 16 * ti + (lane & 15), q = 16 * tj + (lane & 15);
        d4_t acc = {0.0, 0.0, 0.0, 0.0};
        for (int kk = 0; kk < nc; kk += 4) {
            const int r = kk + (lane >> 4);
            const double av = (r < nc && p_op < m) ? T[r * m + p_op] : 0.0;
            const double bv = (r < nc && q < m) ? X[r * (m + 1) + q] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int p = 16 * ti + (lane >> 4) + 4 * v;
            const int s = p < m && q < m ? slot[(p / nf) * nfc + q / nf] : -1;
            if (s >= 0) {
                const double val = loc[(int64_t)(nc + p) * nl + nc + q] - acc[v];
                atomicAdd(a.vals + (int64_t)s * nf * nf + (p % nf) * nf + q % nf, val);
            }
        }
    }
}

template <bool MFMA>
__global__ __launch_bounds__(kCondWG) void k_condense(CondDev a) {
    extern __shared__ double lds[];
    const int nc = a.nc, m = a.m, nl = a.nl, nf = a.nf, nfc = a.nfc;
    const int cell = blockIdx.x;
    double* L = lds;
    double* T = L + nc * nc;  // A_TF, nc x m
    double* X = T + nc * m;   // A_TT^-1 [A_TF B_T], nc x (m + 1)
    const double* loc = a.local + (int64_t)cell * nl * nl;
    for (int t = threadIdx.x; t < nc * nc; t += blockDim.x) L[t] = loc[(int64_t)(t / nc) * nl + t % nc];
    for (int t = threadIdx.x; t < nc * m; t += blockDim.x) {
        const int r = t / m, q = t - r * m;
        const double v = loc[(int64_t)r * nl + nc + q];
        T[t] = v;
        X[r * (m + 1) + q] = v;
    }
    for (int r = threadIdx.x; r < nc; r += blockDim.x) X[r * (m + 1) + m] = a.BT ? a.BT[(int64_t)cell * nc + r] : 0.0;
    if (!chol_lds(L, nc)) {
        if (threadIdx.x == 0) atomicOr(a.err, 1);
        return;  // the whole workgroup leaves together
    }
    for (int col = threadIdx.x; col <= m; col += blockDim.x) chol_solve_col(L, nc, X, m + 1, col);
    __syncthreads();
    // Schur entries (p, q) of the cell's faces j1 = p / nf, j2 = q / nf
    const int* slot = a.slot + (int64_t)cell * nfc * nfc;
    if constexpr (MFMA) {
        schur_mfma(a, T, X, loc, slot);
    } else {
        for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
            const int p = t / m, q = t - p * m;
            const int s = slot[(p / nf) * nfc + q / nf];
            if (s < 0) continue;
            double v = loc[(int64_t)(nc + p) * nl + nc + q];
            for (int r = 0; r < nc; r++) v -= T[r * m + p] * X[r * (m + 1) + q];
            atomicAdd(a.vals + (int64_t)s * nf * nf + (p % nf) * nf + q % nf, v);
        }
    }
    if (a.cstage) {
        for (int p = threadIdx.x; p < m; p += blockDim.x) {
            double v = 0;
            for (int r = 0; r < nc; r++) v += T[r * m + p] * X[r * (m + 1) + m];
            a.cstage[(int64_t)cell * m + p] = v;
        }
    }
}

// b_f = B_F[f] - sum over the cells of face f (cell order) of their stored
// A_TF^T A_TT^-1 B_T entries; thread per face unknown.
__global__ void k_condense_rhs(int nfaces, int nf, int m, const int* fptr, const int* fent, const double* cstage,
                               const double* BF, double* b) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nfaces * nf) return;
    const int f = (int)(t / nf), i = (int)(t - (int64_t)f * nf);
    double v = BF ? BF[t] : 0.0;
    for (int e = fptr[f]; e < fptr[f + 1]; e++) {
        const int cj = fent[e];  // cell * (m / nf) + local face
        v -= cstage[(int64_t)cj * nf + i];
    }
    b[t] = v;
}

__global__ __launch_bounds__(kCondWG) void k_cell_recover(CondDev a, const double* xF, double* xT) {
    extern __shared__ double lds[];
    const int nc = a.nc, m = a.m, nl = a.nl, nf = a.nf;
    const int cell = blockIdx.x;
    double* L = lds;
    double* y = L + nc * nc;
    const double* loc = a.local + (int64_t)cell * nl * nl;
    const int* faces = a.faces + (int64_t)cell * a.nfc;
    for (int t = threadIdx.x; t < nc * nc; t += blockDim.x) L[t] = loc[(int64_t)(t / nc) * nl + t % nc];
    for (int r = threadIdx.x; r < nc; r += blockDim.x) {
        double v = a.BT ? a.BT[(int64_t)cell * nc + r] : 0.0;
        for (int p = 0; p < m; p++) {
            const int f = faces[p / nf];
            if (f >= 0) v -= loc[(int64_t)r * nl + nc + p] * xF[(int64_t)f * nf + p % nf];
        }
        y[r] = v;
    }
    if (!chol_lds(L, nc)) {
        if (threadIdx.x == 0) atomicOr(a.err, 1);
        return;
    }
    if (threadIdx.x == 0) chol_solve_col(L, nc, y, 1, 0);
    __syncthreads();
    for (int r = threadIdx.x; r < nc; r += blockDim.x) xT[(int64_t)cell * nc + r] = y[r];
}

// FHGPU_MFMA=0 keeps the Schur product on the vector ALUs.
bool mfma_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("FHGPU_MFMA");
        return e ? std::atoi(e) != 0 : true;
    }();
    return on;
}

size_t cond_lds_bytes(int nc, int m) { return sizeof(double) * ((size_t)nc * nc + (size_t)nc * m + (size_t)nc * (m + 1)); }

void set_lds(const void* kernel, size_t bytes) {
    int dev = 0, maxb = 0;
    FH_HIP(hipGetDevice(&dev));
    FH_HIP(hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    FH_CHECK(bytes <= (size_t)maxb, -3, "cell matrices do not fit the LDS of one workgroup");
    FH_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

void check_args(int ncells, int nfc, int nc, int nf, int nfaces, const int* cell_faces) {
    FH_CHECK(ncells > 0 && nfc > 0 && nc > 0 && nf > 0 && nfaces > 0, -2, "sizes must be positive");
    FH_CHECK(cell_faces != nullptr, -2, "null cell face array");
    const int64_t nl = nc + (int64_t)nfc * nf;
    FH_CHECK(nl * nl * ncells < ((int64_t)1 << 62), -2, "local matrices too large");
    for (int64_t t = 0; t < (int64_t)ncells * nfc; t++)
        FH_CHECK(cell_faces[t] >= -1 && cell_faces[t] < nfaces, -2, "cell face number out of range");
}

}  // namespace

std::unique_ptr<Bcsr> condense_batched(Ctx& c, int ncells, int nfc, int nc, int nf, int nfaces,
                                       const int* cell_faces, const double* local, const double* BT,
                                       const double* BF, double* b_out) {
    FH_CHECK(!c.host_only, -9, "batched condensation needs a device context");
    check_args(ncells, nfc, nc, nf, nfaces, cell_faces);
    FH_CHECK(local != nullptr, -2, "null local matrices");
    const int m = nfc * nf, nl = nc + m;
    // block pattern: every pair of system faces of one cell, sorted columns
    std::vector<std::vector<int>> adj(nfaces);
    for (int e = 0; e < ncells; e++)
        for (int j1 = 0; j1 < nfc; j1++) {
            const int f1 = cell_faces[(int64_t)e * nfc + j1];
            if (f1 < 0) continue;
            for (int j2 = 0; j2 < nfc; j2++) {
                const int f2 = cell_faces[(int64_t)e * nfc + j2];
                if (f2 >= 0) adj[f1].push_back(f2);
            }
        }
    auto A = std::make_unique<Bcsr>();
    A->nbr = A->nbc = nfaces;
    A->bs_r = A->bs_c = nf;
    A->prec = Prec::F64;
    A->h_rowptr.assign(nfaces + 1, 0);
    for (int f = 0; f < nfaces; f++) {
        auto& v = adj[f];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        FH_CHECK((int64_t)A->h_rowptr[f] + (int64_t)v.size() < INT32_MAX, -2, "too many blocks");
        A->h_rowptr[f + 1] = A->h_rowptr[f] + (int)v.size();
    }
    A->nnzb = A->h_rowptr[nfaces];
    A->h_colidx.resize(std::max<int64_t>(1, A->nnzb));
    for (int f = 0; f < nfaces; f++) std::copy(adj[f].begin(), adj[f].end(), A->h_colidx.begin() + A->h_rowptr[f]);
    // block index of every (cell, j1, j2)
    std::vector<int> slot((size_t)ncells * nfc * nfc, -1);
    for (int e = 0; e < ncells; e++)
        for (int j1 = 0; j1 < nfc; j1++) {
            const int f1 = cell_faces[(int64_t)e * nfc + j1];
            if (f1 < 0) continue;
            const int* b0 = A->h_colidx.data() + A->h_rowptr[f1];
            const int* b1 = A->h_colidx.data() + A->h_rowptr[f1 + 1];
            for (int j2 = 0; j2 < nfc; j2++) {
                const int f2 = cell_faces[(int64_t)e * nfc + j2];
                if (f2 >= 0) slot[((size_t)e * nfc + j1) * nfc + j2] = (int)(std::lower_bound(b0, b1, f2) - A->h_colidx.data());
            }
        }
    A->rowptr.alloc(sizeof(int) * A->h_rowptr.size());
    A->colidx.alloc(sizeof(int) * A->h_colidx.size());
    FH_HIP(hipMemcpyAsync(A->rowptr.p, A->h_rowptr.data(), sizeof(int) * A->h_rowptr.size(), hipMemcpyHostToDevice,
                          c.stream));
    FH_HIP(hipMemcpyAsync(A->colidx.p, A->h_colidx.data(), sizeof(int) * A->h_colidx.size(), hipMemcpyHostToDevice,
                          c.stream));
    const size_t nval = (size_t)A->nnzb * nf * nf;
    A->vals.alloc(sizeof(double) * std::max<size_t>(1, nval));
    FH_HIP(hipMemsetAsync(A->vals.p, 0, sizeof(double) * nval, c.stream));
    DevBuf dslot(sizeof(int) * slot.size()), derr(sizeof(int));
    FH_HIP(hipMemcpyAsync(dslot.p, slot.data(), sizeof(int) * slot.size(), hipMemcpyHostToDevice, c.stream));
    FH_HIP(hipMemsetAsync(derr.p, 0, sizeof(int), c.stream));
    DevBuf cstage;
    if (b_out) cstage.alloc(sizeof(double) * (size_t)ncells * m);
    CondDev d{ncells, nfc, nc, nf, nl, m, nullptr, dslot.as<int>(), local, BT, A->vals.as<double>(),
              b_out ? cstage.as<double>() : nullptr, derr.as<int>()};
    const size_t lds = cond_lds_bytes(nc, m);
    // the Schur product on MFMA where the element blocks are dense enough
    // for 16 x 16 x 4 tiles (at least 2 x 2 output tiles and 2 steps of k)
    if (m >= 32 && nc >= 8 && mfma_enabled()) {
        set_lds((const void*)k_condense<true>, lds);
        hipLaunchKernelGGL(k_condense<true>, dim3(ncells), dim3(kCondWG), lds, c.stream, d);
    } else {
        set_lds((const void*)k_condense<false>, lds);
        hipLaunchKernelGGL(k_condense<false>, dim3(ncells), dim3(kCondWG), lds, c.stream, d);
    }
    FH_HIP(hipGetLastError());
    if (b_out) {
        // (cell, local face) entries of every face, in cell order
        std::vector<int> fptr(nfaces + 1, 0), fent;
        for (int64_t t = 0; t < (int64_t)ncells * nfc; t++)
            if (cell_faces[t] >= 0) fptr[cell_faces[t] + 1]++;
        for (int f = 0; f < nfaces; f++) fptr[f + 1] += fptr[f];
        fent.resize(std::max(1, fptr[nfaces]));
        std::vector<int> fill(fptr.begin(), fptr.end() - 1);
        for (int64_t t = 0; t < (int64_t)ncells * nfc; t++)
            if (cell_faces[t] >= 0) fent[fill[cell_faces[t]]++] = (int)t;
        DevBuf dptr(sizeof(int) * fptr.size()), dent(sizeof(int) * fent.size());
        FH_HIP(hipMemcpyAsync(dptr.p, fptr.data(), sizeof(int) * fptr.size(), hipMemcpyHostToDevice, c.stream));
        FH_HIP(hipMemcpyAsync(dent.p, fent.data(), sizeof(int) * fent.size(), hipMemcpyHostToDevice, c.stream));
        const int64_t n = (int64_t)nfaces * nf;
        hipLaunchKernelGGL(k_condense_rhs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, nfaces, nf, m,
                           dptr.as<int>(), dent.as<int>(), cstage.as<double>(), BF, b_out);
        FH_HIP(hipGetLastError());
        c.sync();  // the staging buffers are freed on return
    }
    int err = 0;
    FH_HIP(hipMemcpyAsync(&err, derr.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    FH_CHECK(err == 0, -4, "a cell matrix A_TT is not positive definite");
    return A;
}

void cell_recover_batched(Ctx& c, int ncells, int nfc, int nc, int nf, int nfaces, const int* cell_faces,
                          const double* local, const double* BT, const double* xF, double* xT) {
    FH_CHECK(!c.host_only, -9, "cell recovery needs a device context");
    check_args(ncells, nfc, nc, nf, nfaces, cell_faces);
    FH_CHECK(local && xF && xT, -2, "null local matrices or vectors");
    const int m = nfc * nf, nl = nc + m;
    DevBuf dfaces(sizeof(int) * (size_t)ncells * nfc), derr(sizeof(int));
    FH_HIP(hipMemcpyAsync(dfaces.p, cell_faces, sizeof(int) * (size_t)ncells * nfc, hipMemcpyHostToDevice, c.stream));
    FH_HIP(hipMemsetAsync(derr.p, 0, sizeof(int), c.stream));
    CondDev d{ncells, nfc, nc, nf, nl, m, dfaces.as<int>(), nullptr, local, BT, nullptr, nullptr, derr.as<int>()};
    const size_t lds = sizeof(double) * ((size_t)nc * nc + nc);
    set_lds((const void*)k_cell_recover, lds);
    hipLaunchKernelGGL(k_cell_recover, dim3(ncells), dim3(kCondWG), lds, c.stream, d, xF, xT);
    FH_HIP(hipGetLastError());
    int err = 0;
    FH_HIP(hipMemcpyAsync(&err, derr.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    FH_CHECK(err == 0, -4, "a cell matrix A_TT is not positive definite");
}

}  // namespace fh
