// Batched HHO local operators and static condensation on gfx950.
//
// One workgroup of one wavefront per cell.  The cell's dense matrices live in
// LDS (about 40 KB for tetrahedra with k = 2): reconstruction stiffness S,
// the right-hand sides B of the reconstruction problem, the reconstruction P
// (solved with a Cholesky factorisation of S without its constant row), the
// local operator A = P^T S P + stabilisation, and the condensation
// A_FF - A_TF^T A_TT^-1 A_TF with its right-hand side.  Every loop runs over
// output entries across the 64 lanes; the factorisations run column by
// column with a barrier per column.  The condensed blocks are added into
// the block-CSR operator with FP64 atomics: an off-diagonal block receives
// one cell's contribution and a diagonal block two, so the result does not
// depend on the order of the additions.
//
// Arithmetic per cell (tetrahedra, k = 2): about 1e5 FMAs on 5e3 doubles of
// LDS, with ~8 KB of HBM traffic (geometry, tables from L2, scattered
// blocks).  The blocks are too small for MFMA tiles to pay off (the largest
// product, A_TF^T X, is 24 x 10 x 24): the kernel is bound by LDS latency
// and barriers, and the whole 1.57 M-cell mesh takes milliseconds.
#include "fh_hho_dev.h"

namespace fh {
namespace {

constexpr int kT = 64;

// position of entry (r, s) of a cell matrix of n unknowns stored by cyclic
// diagonals (diagonal t = (s - r) mod n, position r; diagonal 0 first, then
// diagonals 2p + 1 and 2p + 2 interleaved), or -1 when the entry is stored
// as its transpose (t > n / 2)
__device__ __forceinline__ int elem_entry(int n, int r, int s) {
    const int t = s >= r ? s - r : s - r + n;
    if (t > n / 2) return -1;
    return t == 0 ? r : n + 2 * n * ((t - 1) / 2) + 2 * r + (t - 1) % 2;
}

__device__ double source_at(const HhoDev& d, double x, double y, double z) {
    const double X[3] = {x, y, z};
    if (d.src == 2) {
        double f = 0;
        for (int m = 0; m < d.nmodes; m++) {
            const double* md = d.modes + 4 * m;
            double s = 1, k2 = 0;
            for (int a = 0; a < d.dim; a++) {
                s *= sin(md[1 + a] * M_PI * X[a]);
                k2 += md[1 + a] * md[1 + a];
            }
            f += md[0] * M_PI * M_PI * k2 * s;
        }
        return f;
    }
    const double w = d.src == 1 ? 4 * M_PI : M_PI;
    double s = 1;
    for (int a = 0; a < d.dim; a++) s *= sin(w * X[a]);
    return d.dim * w * w * s;
}

// Exact solution of the source (its Laplacian is -source_at).
__device__ double exact_at(const HhoDev& d, double x, double y, double z) {
    const double X[3] = {x, y, z};
    if (d.src == 2) {
        double u = 0;
        for (int m = 0; m < d.nmodes; m++) {
            const double* md = d.modes + 4 * m;
            double s = 1;
            for (int a = 0; a < d.dim; a++) s *= sin(md[1 + a] * M_PI * X[a]);
            u += md[0] * s;
        }
        return u;
    }
    const double w = d.src == 1 ? 4 * M_PI : M_PI;
    double s = 1;
    for (int a = 0; a < d.dim; a++) s *= sin(w * X[a]);
    return s;
}

// In-place Cholesky of the n x n matrix M (row-major, leading dimension n)
// by columns; same summation order as the host factorisation.
__device__ void lds_chol(double* M, int n, int t, int* err) {
    for (int j = 0; j < n; j++) {
        if (t == 0) {
            double dj = M[j * n + j];
            for (int k = 0; k < j; k++) dj -= M[j * n + k] * M[j * n + k];
            if (!(dj > 0)) {
                *err = 1;
                dj = 1.0;  // keep going; the launch reports the failure
            }
            M[j * n + j] = sqrt(dj);
        }
        __syncthreads();
        const double dj = M[j * n + j];
        for (int i = j + 1 + t; i < n; i += kT) {
            double s = M[i * n + j];
            for (int k = 0; k < j; k++) s -= M[i * n + k] * M[j * n + k];
            M[i * n + j] = s / dj;
        }
        __syncthreads();
    }
}

// Solves L L^T X = R for the columns of R (n rows, leading dimension ld,
// columns [0, ncol)), one column per lane.
__device__ void lds_chol_solve(const double* L, int n, double* R, int ld, int ncol, int t) {
    for (int c = t; c < ncol; c += kT) {
        for (int i = 0; i < n; i++) {
            double s = R[i * ld + c];
            for (int k = 0; k < i; k++) s -= L[i * n + k] * R[k * ld + c];
            R[i * ld + c] = s / L[i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = R[i * ld + c];
            for (int k = i + 1; k < n; k++) s -= L[k * n + i] * R[k * ld + c];
            R[i * ld + c] = s / L[i * n + i];
        }
    }
}

__global__ __launch_bounds__(kT) void k_hho_local(HhoDev d, int* err) {
    extern __shared__ double sm[];
    __shared__ int tab[8];
    const int c = blockIdx.x, t = threadIdx.x;
    if (c >= d.ncell) return;  // whole workgroup
    const int m = d.nfc, nr = d.nr, nc = d.nc, nfb = d.nfb, nh = nc + m * nfb, nF = m * nfb, n1 = nr - 1;
    const int dim = d.dim;
    double* S = sm;                   // nr x nr
    double* B = S + nr * nr;          // nr x nh, later W = S P
    double* P = B + nr * nh;          // nr x nh
    double* A = P + nr * nh;          // nh x nh
    double* PF = A + nh * nh;         // m x nfb x nr
    double* DF = PF + m * nfb * nr;   // m x nfb x nh
    double* T = DF + m * nfb * nh;    // n1 x n1 (also nc x nc)
    double* X = T + n1 * n1;          // nc x (nF + 1)
    double* FQ = X + nc * (nF + 1);   // nq
    double* FS = FQ + d.nq;           // m x 8 face scalars
    double* BT = FS + m * 8;          // nc

    // ---- geometry (every lane; cheap)
    const int* cvv = d.cv + (size_t)c * d.nvc;
    double v0[3], J[3][3], Ji[3][3];
    for (int a = 0; a < 3; a++) v0[a] = d.V[3 * (size_t)cvv[0] + a];
    const int idx[3] = {1, 2, d.kind == 0 ? 3 : 4};
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) J[r][q] = 0;
    for (int q = 0; q < dim; q++)
        for (int r = 0; r < 3; r++) J[r][q] = d.V[3 * (size_t)cvv[idx[q]] + r] - v0[r];
    if (dim == 2) J[2][2] = 1;
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double id = 1.0 / det;
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const double adet = fabs(det);
    const double* Kc = d.K + (size_t)c * 9;
    double G[3][3];
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) {
            double s = 0;
            for (int u = 0; u < 3; u++)
                for (int w = 0; w < 3; w++) s += Ji[r][u] * Kc[u * 3 + w] * Ji[q][w];
            G[r][q] = s;
        }

    // ---- S = sum_ab G_ab Shat_ab
    for (int e = t; e < nr * nr; e += kT) {
        double s = 0;
        for (int ax = 0; ax < dim; ax++)
            for (int bx = 0; bx < dim; bx++) s += G[ax][bx] * d.Shat[(size_t)(ax * 3 + bx) * nr * nr + e];
        S[e] = s;
    }
    // ---- face scalars: J^-1 K n, sqrt(|F|/|det|), |F|/|det|, (K n.n)/h_F
    if (t < m) {
        const int f = t;
        const int fid = d.cf[(size_t)c * m + f];
        const int* fvv = d.fv + (size_t)fid * d.nvf;
        double a0[3], E1[3], E2[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++) {
            a0[a] = d.V[3 * (size_t)fvv[0] + a];
            E1[a] = d.V[3 * (size_t)fvv[1] + a] - a0[a];
        }
        if (dim == 3)
            for (int a = 0; a < 3; a++) E2[a] = d.V[3 * (size_t)fvv[d.kind == 1 ? 3 : 2] + a] - a0[a];
        double sF, hF, nF_[3];
        if (dim == 3) {
            const double cr[3] = {E1[1] * E2[2] - E1[2] * E2[1], E1[2] * E2[0] - E1[0] * E2[2],
                                  E1[0] * E2[1] - E1[1] * E2[0]};
            sF = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            for (int a = 0; a < 3; a++) nF_[a] = cr[a] / sF;
            const double l1 = sqrt(E1[0] * E1[0] + E1[1] * E1[1] + E1[2] * E1[2]);
            const double l2 = sqrt(E2[0] * E2[0] + E2[1] * E2[1] + E2[2] * E2[2]);
            if (d.kind == 0) {
                double e3[3];
                for (int a = 0; a < 3; a++) e3[a] = d.V[3 * (size_t)fvv[2] + a] - d.V[3 * (size_t)fvv[1] + a];
                const double l3 = sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
                hF = fmax(fmax(l1, l2), l3);
            } else
                hF = fmax(l1, l2);
        } else {
            sF = sqrt(E1[0] * E1[0] + E1[1] * E1[1] + E1[2] * E1[2]);
            hF = sF;
            nF_[0] = E1[1] / sF;
            nF_[1] = -E1[0] / sF;
            nF_[2] = 0;
        }
        double cc[3] = {0, 0, 0}, fc[3];
        for (int l = 0; l < d.nvc; l++)
            for (int a = 0; a < 3; a++) cc[a] += d.V[3 * (size_t)cvv[l] + a] / d.nvc;
        for (int a = 0; a < 3; a++) fc[a] = a0[a];
        for (int l = 1; l < d.nvf; l++)
            for (int a = 0; a < 3; a++) fc[a] += d.V[3 * (size_t)fvv[l] + a];
        for (int a = 0; a < 3; a++) fc[a] /= d.nvf;
        if (nF_[0] * (fc[0] - cc[0]) + nF_[1] * (fc[1] - cc[1]) + nF_[2] * (fc[2] - cc[2]) < 0)
            for (int a = 0; a < 3; a++) nF_[a] = -nF_[a];
        double Kn[3];
        for (int a = 0; a < 3; a++) Kn[a] = Kc[3 * a] * nF_[0] + Kc[3 * a + 1] * nF_[1] + Kc[3 * a + 2] * nF_[2];
        for (int a = 0; a < 3; a++) FS[f * 8 + a] = Ji[a][0] * Kn[0] + Ji[a][1] * Kn[1] + Ji[a][2] * Kn[2];
        FS[f * 8 + 3] = sqrt(sF / adet);
        FS[f * 8 + 4] = sF / adet;
        FS[f * 8 + 5] = (Kn[0] * nF_[0] + Kn[1] * nF_[1] + Kn[2] * nF_[2]) / hF;
        tab[f] = d.cft[(size_t)c * m + f];
    }
    // ---- source at the cell rule
    {
        const double sq = sqrt(adet);
        for (int q = t; q < d.nq; q += kT) {
            const double* xi = d.cqp + 3 * q;
            double x[3];
            for (int a = 0; a < 3; a++) x[a] = v0[a] + J[a][0] * xi[0] + J[a][1] * xi[1] + J[a][2] * xi[2];
            FQ[q] = d.cqw[q] * source_at(d, x[0], x[1], x[2]) * sq;
        }
    }
    __syncthreads();
    // ---- B (right-hand sides of the reconstruction) and the face projections
    for (int e = t; e < nr * nh; e += kT) {
        const int i = e / nh, j = e - i * nh;
        double v;
        if (j < nc) {
            v = S[i * nr + j];
            for (int f = 0; f < m; f++) {
                const double* Ps = d.Psi + (size_t)tab[f] * 3 * nr * nc;
                const double s2 = FS[f * 8 + 4];
                for (int ax = 0; ax < dim; ax++) v -= s2 * FS[f * 8 + ax] * Ps[((size_t)ax * nr + i) * nc + j];
            }
        } else {
            const int jf = j - nc, f = jf / nfb, jj = jf - f * nfb;
            const double* Ph = d.Phi + (size_t)tab[f] * 3 * nr * nfb;
            const double s1 = FS[f * 8 + 3];
            v = 0;
            for (int ax = 0; ax < dim; ax++) v += s1 * FS[f * 8 + ax] * Ph[((size_t)ax * nr + i) * nfb + jj];
        }
        B[e] = v;
    }
    for (int e = t; e < m * nfb * nr; e += kT) {
        const int f = e / (nfb * nr), r = e - f * nfb * nr;
        PF[e] = FS[f * 8 + 3] * d.MFT[(size_t)tab[f] * nfb * nr + r];
    }
    __syncthreads();
    for (int e = t; e < n1 * n1; e += kT) {
        const int i = e / n1, j = e - i * n1;
        T[e] = S[(i + 1) * nr + j + 1];
    }
    for (int j = t; j < nh; j += kT) P[j] = j == 0 ? 1.0 : 0.0;
    for (int e = t; e < n1 * nh; e += kT) P[nh + e] = B[nh + e];
    __syncthreads();
    // ---- reconstruction: rows 1.. solve S' P' = B'
    if (n1 > 0) {
        lds_chol(T, n1, t, err);
        lds_chol_solve(T, n1, P + nh, nh, nh, t);
    }
    __syncthreads();
    // ---- W = S P (into B), then A = P^T W
    for (int e = t; e < nr * nh; e += kT) {
        const int i = e / nh, j = e - i * nh;
        double s = 0;
        for (int l = 0; l < nr; l++) s += S[i * nr + l] * P[l * nh + j];
        B[e] = s;
    }
    __syncthreads();
    for (int e = t; e < nh * nh; e += kT) {
        const int r = e / nh, j = e - r * nh;
        double s = 0;
        for (int i = 0; i < nr; i++) s += P[i * nh + r] * B[i * nh + j];
        A[e] = s;
    }
    // ---- stabilisation: DF_f = Pi_F (P - D_T) - I_F for every face
    for (int e = t; e < m * nfb * nh; e += kT) {
        const int f = e / (nfb * nh), rem = e - f * nfb * nh, j = rem / nh, r = rem - j * nh;
        const double* pf = PF + ((size_t)f * nfb + j) * nr;
        double v = 0;
        for (int i = 0; i < nr; i++) v += pf[i] * P[i * nh + r];
        if (r == nc + f * nfb + j) v -= 1.0;
        for (int i = 0; i < nc; i++) v -= pf[i] * (P[i * nh + r] - (i == r ? 1.0 : 0.0));
        DF[e] = v;
    }
    __syncthreads();
    for (int e = t; e < nh * nh; e += kT) {
        const int r = e / nh, s = e - r * nh;
        double v = A[e];
        for (int f = 0; f < m; f++) {
            const double cf_ = FS[f * 8 + 5];
            const double* D = DF + (size_t)f * nfb * nh;
            for (int j = 0; j < nfb; j++) v += (cf_ * D[j * nh + r]) * D[j * nh + s];
        }
        A[e] = v;
    }
    // ---- cell right-hand side
    for (int i = t; i < nc; i += kT) {
        double s = 0;
        for (int q = 0; q < d.nq; q++) s += FQ[q] * d.cellPsi[(size_t)q * nr + i];
        BT[i] = s;
    }
    __syncthreads();
    // ---- static condensation
    for (int e = t; e < nc * nc; e += kT) {
        const int i = e / nc, j = e - i * nc;
        T[e] = A[i * nh + j];
    }
    for (int e = t; e < nc * (nF + 1); e += kT) {
        const int i = e / (nF + 1), j = e - i * (nF + 1);
        X[e] = j < nF ? A[i * nh + nc + j] : BT[i];
    }
    __syncthreads();
    lds_chol(T, nc, t, err);
    lds_chol_solve(T, nc, X, nF + 1, nF + 1, t);
    __syncthreads();
    const int* frow = d.frow + (size_t)c * m;
    if (d.post) {
        // ---- recovery v_T = A_TT^-1 (B_T - A_TF v_F) = X[:, nF] - X[:, :nF] v_F,
        // reconstruction p = P [v_T; v_F], L2 integrals at the cell rule
        double* w = DF;       // nh: [v_T; v_F]
        double* pc = DF + nh; // nr
        for (int r = t; r < nF; r += kT) {
            const int f = r / nfb, a = r - f * nfb, row = frow[f];
            w[nc + r] = row >= 0 ? d.xsol[(size_t)row * nfb + a] : 0.0;  // homogeneous Dirichlet
        }
        __syncthreads();
        for (int i = t; i < nc; i += kT) {
            double s = X[i * (nF + 1) + nF];
            for (int j = 0; j < nF; j++) s -= X[i * (nF + 1) + j] * w[nc + j];
            w[i] = s;
        }
        __syncthreads();
        for (int i = t; i < nr; i += kT) {
            double s = 0;
            for (int j = 0; j < nh; j++) s += P[i * nh + j] * w[j];
            pc[i] = s;
        }
        __syncthreads();
        const double isq = 1.0 / sqrt(adet);
        double e2 = 0, u2 = 0;
        for (int q = t; q < d.nq; q += kT) {
            const double* xi = d.cqp + 3 * q;
            double x[3];
            for (int a = 0; a < 3; a++) x[a] = v0[a] + J[a][0] * xi[0] + J[a][1] * xi[1] + J[a][2] * xi[2];
            double ph = 0;
            for (int i = 0; i < nr; i++) ph += pc[i] * d.cellPsi[(size_t)q * nr + i];
            ph *= isq;
            const double u = exact_at(d, x[0], x[1], x[2]);
            const double wq = d.cqw[q] * adet;
            e2 += wq * (u - ph) * (u - ph);
            u2 += wq * u * u;
        }
        for (int o = kT / 2; o > 0; o >>= 1) {  // one wavefront per cell
            e2 += __shfl_xor(e2, o, kT);
            u2 += __shfl_xor(u2, o, kT);
        }
        if (t == 0) {
            d.cell_err[2 * (size_t)c] = e2;
            d.cell_err[2 * (size_t)c + 1] = u2;
        }
        return;
    }
    // ---- condensed blocks and right-hand side, added into the global system
    const int* blk = d.blk + (size_t)c * m * m;
    const int bsz = nfb * nfb;
    for (int e = t; e < nF * nF; e += kT) {
        const int r = e / nF, s = e - r * nF;
        const int f = r / nfb, a = r - f * nfb, g = s / nfb, bb = s - g * nfb;
        const int p = blk[f * m + g];
        const int ep = d.elem_vals != nullptr ? elem_entry(nF, r, s) : -1;
        if (p < 0 && ep < 0) continue;
        double v = A[(nc + r) * nh + nc + s];
        for (int i = 0; i < nc; i++) v -= A[i * nh + nc + r] * X[i * (nF + 1) + s];
        if (ep >= 0) d.elem_vals[(size_t)d.elem_pos[c] * d.elem_stride + ep] = v;
        if (p >= 0) atomicAdd(d.vals + (size_t)p * bsz + a * nfb + bb, v);
    }
    for (int r = t; r < nF; r += kT) {
        const int f = r / nfb, a = r - f * nfb;
        const int row = frow[f];
        if (row < 0) continue;
        double v = 0;
        for (int i = 0; i < nc; i++) v -= A[i * nh + nc + r] * X[i * (nF + 1) + nF];
        atomicAdd(d.b + (size_t)row * nfb + a, v);
    }
}

// ============================================================================
// General cells (bilinear quadrilaterals, trilinear hexahedra): per-cell
// quadrature through the cell and face maps.  One workgroup of kTG threads
// per cell; the cell's matrices stay in LDS (146 KB for hexahedra at k = 3,
// one workgroup per CU).  The steps and the formulas are those of the host
// path (HhoProblem::Impl::local_general): integrals staged by slices of
// quadrature points, the saddle-point reconstruction with the mean-value
// condition, P^T S P, the general "hho" stabilisation with the face and cell
// L2 projections, and the static condensation.
constexpr int kTG = 256;

struct GenLayout {
    int S, Lm, MT, BT, P, A, Mf, TF, R, T, G, total;
    int Qc, SPc, SPf, stage;
};

__host__ __device__ inline GenLayout gen_layout(const HhoDev& d) {
    GenLayout L{};
    const int m = d.nfc, nr = d.nr, nc = d.nc, nfb = d.nfb, nh = nc + m * nfb, nF = m * nfb, n1 = nr - 1;
    const int nq1 = d.ng1, nqf = d.dim == 3 ? nq1 * nq1 : nq1;
    L.Qc = nq1 * nq1;                          // one slice of the cell rule per chunk
    L.SPc = 2 + 7 * nr;                         // wd, f*wd, psi, grad psi, K grad psi
    L.SPf = 2 + 2 * nr + nfb;                   // w, w*da, psi, K grad psi . n, phi
    L.stage = L.SPc * L.Qc > L.SPf * nqf ? L.SPc * L.Qc : L.SPf * nqf;
    int o = 0;
    L.S = o; o += nr * nr;
    L.Lm = o; o += nr;
    L.MT = o; o += nc * nr;
    L.BT = o; o += nc;
    L.P = o; o += nr * nh;
    L.A = o; o += nh * nh > L.stage ? nh * nh : L.stage;
    L.Mf = o; o += m * nfb * nfb;
    L.TF = o; o += m * nfb * nr;
    L.R = o;
    {
        int r = nr * nh;                         // W = S P
        const int r2 = nc * nh + 2 * nfb * nh;   // Dt, DF, MD
        const int r3 = nc * (nF + 1);            // X
        if (r2 > r) r = r2;
        if (r3 > r) r = r3;
        o += r;
    }
    L.T = o;
    {
        int t = n1 * n1;
        if (nc * nc > t) t = nc * nc;
        if (nfb * nfb > t) t = nfb * nfb;
        o += t;
    }
    L.G = o; o += 128;                          // cell geometry, face scalars, reduction
    L.total = o;
    return L;
}

__device__ inline void leg_all(int p, double x, double* L, double* dL) {
    L[0] = 1;
    dL[0] = 0;
    if (p >= 1) { L[1] = x; dL[1] = 1; }
    for (int q = 2; q <= p; q++) {
        L[q] = ((2 * q - 1) * x * L[q - 1] - (q - 1) * L[q - 2]) / q;
        dL[q] = dL[q - 2] + (2 * q - 1) * L[q - 1];
    }
}

// Reconstruction basis at x (cell frame ctr, hw): psi[i]; grad psi_i into
// g[3 i .. 3 i + 2] when g is not null; grad psi_i . kn into kgn[i] when kgn
// is not null.
__device__ inline void gen_basis(const HhoDev& d, const double* ctr, const double* hw, const double x[3], double* psi,
                                 double* g, const double* kn = nullptr, double* kgn = nullptr) {
    const int nr = d.nr;
    const int deg = d.rexp[3 * (nr - 1)] + d.rexp[3 * (nr - 1) + 1] + d.rexp[3 * (nr - 1) + 2];  // graded: last is highest
    double L[3][6], dL[3][6];
    for (int a = 0; a < 3; a++) {
        if (a < d.dim) leg_all(deg, (x[a] - ctr[a]) / hw[a], L[a], dL[a]);
        else { L[a][0] = 1; dL[a][0] = 0; }
    }
    for (int i = 0; i < nr; i++) {
        const int e0 = d.rexp[3 * i], e1 = d.rexp[3 * i + 1], e2 = d.rexp[3 * i + 2];
        psi[i] = L[0][e0] * L[1][e1] * L[2][e2];
        if (g || kgn) {
            const double g0 = dL[0][e0] / hw[0] * L[1][e1] * L[2][e2];
            const double g1 = d.dim >= 2 ? L[0][e0] * dL[1][e1] / hw[1] * L[2][e2] : 0.0;
            const double g2 = d.dim == 3 ? L[0][e0] * L[1][e1] * dL[2][e2] / hw[2] : 0.0;
            if (g) { g[3 * i] = g0; g[3 * i + 1] = g1; g[3 * i + 2] = g2; }
            if (kgn) kgn[i] = g0 * kn[0] + g1 * kn[1] + g2 * kn[2];
        }
    }
}

// Cell map at xi: x and det J (corners CV[8][3], corner code x + 2y + 4z).
__device__ inline double gen_cell_map(int dim, const double* CV, const double xi[3], double x[3]) {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    x[0] = x[1] = x[2] = 0;
    const int ncorner = dim == 3 ? 8 : 4;
    for (int cc = 0; cc < ncorner; cc++) {
        double f[3], df[3];
        for (int a = 0; a < 3; a++) {
            const double sg = ((cc >> a) & 1) ? 1.0 : -1.0;
            f[a] = a < dim ? 0.5 * (1 + sg * xi[a]) : 1.0;
            df[a] = a < dim ? 0.5 * sg : 0.0;
        }
        const double N = f[0] * f[1] * f[2];
        const double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
        for (int r = 0; r < 3; r++) {
            x[r] += N * CV[cc * 3 + r];
            for (int q = 0; q < dim; q++) J[r][q] += dN[q] * CV[cc * 3 + r];
        }
    }
    if (dim == 2) J[2][2] = 1;
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// Face map at (s, t): x and the (unoriented) vector area element (FV: the
// face's corners a, b, c, d, or a, b in 2-D).
__device__ inline void gen_face_map(int dim, const double* FV, double s, double t, double x[3], double nv[3]) {
    if (dim == 2) {
        double xs[3];
        for (int a = 0; a < 3; a++) {
            x[a] = 0.5 * (1 - s) * FV[a] + 0.5 * (1 + s) * FV[3 + a];
            xs[a] = 0.5 * (FV[3 + a] - FV[a]);
        }
        nv[0] = xs[1];
        nv[1] = -xs[0];
        nv[2] = 0;
        return;
    }
    double xs[3], xt[3];
    for (int a = 0; a < 3; a++) {
        const double A = FV[a], B = FV[3 + a], C = FV[6 + a], D = FV[9 + a];
        x[a] = 0.25 * ((1 - s) * (1 - t) * A + (1 + s) * (1 - t) * B + (1 + s) * (1 + t) * C + (1 - s) * (1 + t) * D);
        xs[a] = 0.25 * ((1 - t) * (B - A) + (1 + t) * (C - D));
        xt[a] = 0.25 * ((1 - s) * (D - A) + (1 + s) * (C - B));
    }
    nv[0] = xs[1] * xt[2] - xs[2] * xt[1];
    nv[1] = xs[2] * xt[0] - xs[0] * xt[2];
    nv[2] = xs[0] * xt[1] - xs[1] * xt[0];
}

__device__ void gen_chol(double* M, int n, int t, int* err) {
    for (int j = 0; j < n; j++) {
        if (t == 0) {
            double dj = M[j * n + j];
            for (int k = 0; k < j; k++) dj -= M[j * n + k] * M[j * n + k];
            if (!(dj > 0)) {
                atomicOr(err, 1);
                dj = 1.0;
            }
            M[j * n + j] = sqrt(dj);
        }
        __syncthreads();
        const double dj = M[j * n + j];
        for (int i = j + 1 + t; i < n; i += kTG) {
            double s = M[i * n + j];
            for (int k = 0; k < j; k++) s -= M[i * n + k] * M[j * n + k];
            M[i * n + j] = s / dj;
        }
        __syncthreads();
    }
}

__device__ void gen_chol_solve(const double* L, int n, double* R, int ld, int ncol, int t) {
    for (int c = t; c < ncol; c += kTG) {
        for (int i = 0; i < n; i++) {
            double s = R[i * ld + c];
            for (int k = 0; k < i; k++) s -= L[i * n + k] * R[k * ld + c];
            R[i * ld + c] = s / L[i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = R[i * ld + c];
            for (int k = i + 1; k < n; k++) s -= L[k * n + i] * R[k * ld + c];
            R[i * ld + c] = s / L[i * n + i];
        }
    }
}

__global__ __launch_bounds__(kTG) void k_hho_general(HhoDev d, int* err) {
    extern __shared__ double sm[];
    const int c = blockIdx.x, t = threadIdx.x;
    if (c >= d.ncell) return;  // whole workgroup
    const GenLayout Lo = gen_layout(d);
    const int m = d.nfc, nr = d.nr, nc = d.nc, nfb = d.nfb, nh = nc + m * nfb, nF = m * nfb, n1 = nr - 1;
    const int dim = d.dim, nq1 = d.ng1;
    double* S = sm + Lo.S;
    double* Lm = sm + Lo.Lm;
    double* MT = sm + Lo.MT;
    double* BT = sm + Lo.BT;
    double* P = sm + Lo.P;    // B, then P
    double* A = sm + Lo.A;    // staging during the integrals
    double* Mf = sm + Lo.Mf;
    double* TF = sm + Lo.TF;  // trace integrals, then face projections
    double* R = sm + Lo.R;
    double* T = sm + Lo.T;
    double* Gm = sm + Lo.G;
    double* CV = Gm;          // 24: cell corners
    double* ctr = Gm + 24;    // 3
    double* hw = Gm + 27;     // 3
    double* Kc = Gm + 30;     // 9
    double* FVs = Gm + 40;    // 12: corners of the current face
    double* FSg = Gm + 52;    // m: outward sign of each face;  Gm + 60: m coefficients (K n.n)/h_F
    double* Cf = Gm + 60;
    double* red = Gm + 72;    // 2 per wavefront: reduction scratch
    const int nvc = d.nvc;

    // ---- cell geometry
    if (t < nvc * 3) CV[t] = d.V[3 * (size_t)d.cv[(size_t)c * nvc + t / 3] + t % 3];
    if (t < 9) Kc[t] = d.K[(size_t)c * 9 + t];
    __syncthreads();
    if (t == 0) {
        for (int a = 0; a < 3; a++) {
            double lo = 1e300, hi = -1e300, s = 0;
            for (int l = 0; l < nvc; l++) {
                const double v = CV[l * 3 + a];
                s += v;
                lo = fmin(lo, v);
                hi = fmax(hi, v);
            }
            ctr[a] = s / nvc;
            hw[a] = a < dim ? 0.5 * (hi - lo) : 1.0;
        }
    }
    for (int e = t; e < nr * nr; e += kTG) S[e] = 0;
    for (int e = t; e < nr; e += kTG) Lm[e] = 0;
    for (int e = t; e < nc * nr; e += kTG) MT[e] = 0;
    for (int e = t; e < nc; e += kTG) BT[e] = 0;
    for (int e = t; e < m * nfb * nfb; e += kTG) Mf[e] = 0;
    for (int e = t; e < m * nfb * nr; e += kTG) TF[e] = 0;
    __syncthreads();

    // ---- cell integrals, one slice of quadrature points at a time
    const int nqc = dim == 3 ? nq1 * nq1 * nq1 : nq1 * nq1;
    const int SPc = Lo.SPc;
    const double* gx = d.g1;
    const double* gw = d.g1 + nq1;
    for (int q0 = 0; q0 < nqc; q0 += Lo.Qc) {
        const int nqq = min(Lo.Qc, nqc - q0);
        for (int qq = t; qq < nqq; qq += kTG) {
            const int q = q0 + qq;
            const int i0 = q % nq1, i1 = (q / nq1) % nq1, i2 = dim == 3 ? q / (nq1 * nq1) : 0;
            const double xi[3] = {gx[i0], gx[i1], dim == 3 ? gx[i2] : 0.0};
            const double w = gw[i0] * gw[i1] * (dim == 3 ? gw[i2] : 1.0);
            double x[3];
            const double det = gen_cell_map(dim, CV, xi, x);
            if (!(det > 0)) atomicOr(err, 2);
            double* st = A + qq * SPc;
            st[0] = w * det;
            st[1] = w * det * source_at(d, x[0], x[1], x[2]);
            double* psi = st + 2;
            double* g = psi + nr;
            double* kg = g + 3 * nr;
            gen_basis(d, ctr, hw, x, psi, g);
            for (int i = 0; i < nr; i++)
                for (int r = 0; r < 3; r++)
                    kg[3 * i + r] = Kc[r * 3] * g[3 * i] + Kc[r * 3 + 1] * g[3 * i + 1] + Kc[r * 3 + 2] * g[3 * i + 2];
        }
        __syncthreads();
        for (int e = t; e < nr * nr; e += kTG) {
            const int i = e / nr, j = e - i * nr;
            double s = 0;
            for (int qq = 0; qq < nqq; qq++) {
                const double* st = A + qq * SPc;
                const double* g = st + 2 + nr;
                const double* kg = g + 3 * nr;
                s += st[0] * (kg[3 * i] * g[3 * j] + kg[3 * i + 1] * g[3 * j + 1] + kg[3 * i + 2] * g[3 * j + 2]);
            }
            S[e] += s;
        }
        for (int e = t; e < nc * nr; e += kTG) {
            const int i = e / nr, j = e - i * nr;
            double s = 0;
            for (int qq = 0; qq < nqq; qq++) {
                const double* st = A + qq * SPc;
                s += st[0] * st[2 + i] * st[2 + j];
            }
            MT[e] += s;
        }
        for (int i = t; i < nr + nc; i += kTG) {
            double s = 0;
            if (i < nr)
                for (int qq = 0; qq < nqq; qq++) s += A[qq * SPc] * A[qq * SPc + 2 + i];
            else
                for (int qq = 0; qq < nqq; qq++) s += A[qq * SPc + 1] * A[qq * SPc + 2 + (i - nr)];
            if (i < nr) Lm[i] += s;
            else BT[i - nr] += s;
        }
        __syncthreads();
    }
    // B = [S[:, :nc] | 0]
    for (int e = t; e < nr * nh; e += kTG) {
        const int i = e / nh, j = e - i * nh;
        P[e] = j < nc ? S[i * nr + j] : 0.0;
    }
    __syncthreads();

    // ---- face integrals
    const int nqf = dim == 3 ? nq1 * nq1 : nq1;
    const int SPf = Lo.SPf;
    const int nvf = d.nvf;
    for (int f = 0; f < m; f++) {
        const int fid = d.cf[(size_t)c * m + f];
        if (t < nvf * 3) FVs[t] = d.V[3 * (size_t)d.fv[(size_t)fid * nvf + t / 3] + t % 3];
        __syncthreads();
        if (t == 0) {
            double x0[3], nv[3];
            gen_face_map(dim, FVs, 0, 0, x0, nv);
            const double dd = nv[0] * (x0[0] - ctr[0]) + nv[1] * (x0[1] - ctr[1]) + nv[2] * (x0[2] - ctr[2]);
            const double sg = dd < 0 ? -1.0 : 1.0;
            FSg[f] = sg;
            const double l0 = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
            double nu[3] = {sg * nv[0] / l0, sg * nv[1] / l0, sg * nv[2] / l0};
            double Kn = 0;
            for (int r = 0; r < 3; r++)
                for (int q = 0; q < 3; q++) Kn += nu[r] * Kc[r * 3 + q] * nu[q];
            double hF = 0;
            if (dim == 2) {
                const double e0 = FVs[3] - FVs[0], e1 = FVs[4] - FVs[1], e2 = FVs[5] - FVs[2];
                hF = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            } else
                for (int e = 0; e < 4; e++) {
                    const int a0 = e, a1 = (e + 1) % 4;
                    const double e0 = FVs[3 * a1] - FVs[3 * a0], e1 = FVs[3 * a1 + 1] - FVs[3 * a0 + 1],
                                 e2 = FVs[3 * a1 + 2] - FVs[3 * a0 + 2];
                    hF = fmax(hF, sqrt(e0 * e0 + e1 * e1 + e2 * e2));
                }
            Cf[f] = Kn / hF;
        }
        __syncthreads();
        for (int q = t; q < nqf; q += kTG) {
            const int i0 = q % nq1, i1 = dim == 3 ? q / nq1 : 0;
            const double sv = gx[i0], tv = dim == 3 ? gx[i1] : 0.0;
            const double w = gw[i0] * (dim == 3 ? gw[i1] : 1.0);
            double x[3], nv[3];
            gen_face_map(dim, FVs, sv, tv, x, nv);
            const double sg = FSg[f];
            for (int a = 0; a < 3; a++) nv[a] *= sg;
            double* st = A + q * SPf;
            st[0] = w;
            st[1] = w * sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
            double* psi = st + 2;
            double* kgn = psi + nr;
            double* phi = kgn + nr;
            // (K grad psi) . n = grad psi . (K n), K symmetric
            const double kn[3] = {Kc[0] * nv[0] + Kc[1] * nv[1] + Kc[2] * nv[2], Kc[3] * nv[0] + Kc[4] * nv[1] + Kc[5] * nv[2],
                                  Kc[6] * nv[0] + Kc[7] * nv[1] + Kc[8] * nv[2]};
            gen_basis(d, ctr, hw, x, psi, nullptr, kn, kgn);
            double Ls[6], dLs[6], Lt[6], dLt[6];
            const int kf = dim == 3 ? (d.fexp[2 * (nfb - 1)] + d.fexp[2 * (nfb - 1) + 1]) : d.fexp[2 * (nfb - 1)];
            leg_all(kf, sv, Ls, dLs);
            leg_all(kf, tv, Lt, dLt);
            for (int j = 0; j < nfb; j++) phi[j] = Ls[d.fexp[2 * j]] * (dim == 3 ? Lt[d.fexp[2 * j + 1]] : 1.0);
        }
        __syncthreads();
        // B[i][nc + f nfb + a] += w Kgn_i phi_a ; B[i][j] -= w Kgn_i psi_j (j < nc)
        for (int e = t; e < nr * (nfb + nc); e += kTG) {
            const int i = e / (nfb + nc), jj = e - i * (nfb + nc);
            double s = 0;
            if (jj < nfb) {
                for (int q = 0; q < nqf; q++) {
                    const double* st = A + q * SPf;
                    s += st[0] * st[2 + nr + i] * st[2 + 2 * nr + jj];
                }
                P[i * nh + nc + f * nfb + jj] += s;
            } else {
                const int j = jj - nfb;
                for (int q = 0; q < nqf; q++) {
                    const double* st = A + q * SPf;
                    s += st[0] * st[2 + nr + i] * st[2 + j];
                }
                P[i * nh + j] -= s;
            }
        }
        for (int e = t; e < nfb * (nfb + nr); e += kTG) {
            const int aa = e / (nfb + nr), jj = e - aa * (nfb + nr);
            double s = 0;
            if (jj < nfb) {
                for (int q = 0; q < nqf; q++) {
                    const double* st = A + q * SPf;
                    s += st[1] * st[2 + 2 * nr + aa] * st[2 + 2 * nr + jj];
                }
                Mf[(f * nfb + aa) * nfb + jj] = s;
            } else {
                const int i = jj - nfb;
                for (int q = 0; q < nqf; q++) {
                    const double* st = A + q * SPf;
                    s += st[1] * st[2 + 2 * nr + aa] * st[2 + i];
                }
                TF[(f * nfb + aa) * nr + i] = s;
            }
        }
        __syncthreads();
    }
    for (int j = t; j < nh; j += kTG) P[j] = 0.0;  // degree-0 row
    // ---- reconstruction: rows 1.. solve S' P' = B'
    for (int e = t; e < n1 * n1; e += kTG) {
        const int i = e / n1, j = e - i * n1;
        T[e] = S[(i + 1) * nr + j + 1];
    }
    __syncthreads();
    if (n1 > 0) {
        gen_chol(T, n1, t, err);
        gen_chol_solve(T, n1, P + nh, nh, nh, t);
    }
    __syncthreads();
    // row 0 from the mean value
    for (int j = t; j < nh; j += kTG) {
        double v = j < nc ? Lm[j] : 0.0;
        for (int i = 1; i < nr; i++) v -= Lm[i] * P[i * nh + j];
        P[j] = v / Lm[0];
    }
    __syncthreads();
    // ---- W = S P (in R), A = P^T W
    double* W = R;
    for (int e = t; e < nr * nh; e += kTG) {
        const int i = e / nh, j = e - i * nh;
        double s = 0;
        for (int l = 0; l < nr; l++) s += S[i * nr + l] * P[l * nh + j];
        W[e] = s;
    }
    __syncthreads();
    for (int e = t; e < nh * nh; e += kTG) {
        const int r = e / nh, j = e - r * nh;
        double s = 0;
        for (int i = 0; i < nr; i++) s += P[i * nh + r] * W[i * nh + j];
        A[e] = s;
    }
    __syncthreads();
    // ---- Dt = M_T^-1 (N_T P) - [I 0]
    double* Dt = R;
    double* DF = R + nc * nh;
    double* MD = DF + nfb * nh;
    for (int e = t; e < nc * nc; e += kTG) {
        const int i = e / nc, j = e - i * nc;
        T[e] = MT[i * nr + j];
    }
    for (int e = t; e < nc * nh; e += kTG) {
        const int i = e / nh, j = e - i * nh;
        double s = 0;
        for (int l = 0; l < nr; l++) s += MT[i * nr + l] * P[l * nh + j];
        Dt[e] = s;
    }
    __syncthreads();
    gen_chol(T, nc, t, err);
    gen_chol_solve(T, nc, Dt, nh, nh, t);
    __syncthreads();
    for (int i = t; i < nc; i += kTG) Dt[i * nh + i] -= 1.0;
    __syncthreads();
    // ---- stabilisation, face by face
    for (int f = 0; f < m; f++) {
        double* Mff = Mf + f * nfb * nfb;
        double* PF = TF + f * nfb * nr;
        for (int e = t; e < nfb * nfb; e += kTG) T[e] = Mff[e];
        __syncthreads();
        gen_chol(T, nfb, t, err);
        gen_chol_solve(T, nfb, PF, nr, nr, t);
        __syncthreads();
        for (int e = t; e < nfb * nh; e += kTG) {
            const int aa = e / nh, j = e - aa * nh;
            double v = 0;
            for (int i = 0; i < nr; i++) v += PF[aa * nr + i] * P[i * nh + j];
            if (j == nc + f * nfb + aa) v -= 1.0;
            for (int i = 0; i < nc; i++) v -= PF[aa * nr + i] * Dt[i * nh + j];
            DF[e] = v;
        }
        __syncthreads();
        for (int e = t; e < nfb * nh; e += kTG) {
            const int aa = e / nh, j = e - aa * nh;
            double v = 0;
            for (int bb = 0; bb < nfb; bb++) v += Mff[aa * nfb + bb] * DF[bb * nh + j];
            MD[e] = v;
        }
        __syncthreads();
        const double cf_ = Cf[f];
        for (int e = t; e < nh * nh; e += kTG) {
            const int r = e / nh, j = e - r * nh;
            double v = 0;
            for (int aa = 0; aa < nfb; aa++) v += (cf_ * DF[aa * nh + r]) * MD[aa * nh + j];
            A[e] += v;
        }
        __syncthreads();
    }
    // ---- static condensation
    double* X = R;
    for (int e = t; e < nc * nc; e += kTG) {
        const int i = e / nc, j = e - i * nc;
        T[e] = A[i * nh + j];
    }
    for (int e = t; e < nc * (nF + 1); e += kTG) {
        const int i = e / (nF + 1), j = e - i * (nF + 1);
        X[e] = j < nF ? A[i * nh + nc + j] : BT[i];
    }
    __syncthreads();
    gen_chol(T, nc, t, err);
    gen_chol_solve(T, nc, X, nF + 1, nF + 1, t);
    __syncthreads();
    const int* frow = d.frow + (size_t)c * m;
    if (d.post) {
        double* wv = A;       // nh: [v_T; v_F] (A is free once X holds the condensation)
        double* pc = A + nh;  // nr
        for (int r = t; r < nF; r += kTG) {
            const int f = r / nfb, aa = r - f * nfb, row = frow[f];
            wv[nc + r] = row >= 0 ? d.xsol[(size_t)row * nfb + aa] : 0.0;
        }
        __syncthreads();
        for (int i = t; i < nc; i += kTG) {
            double s = X[i * (nF + 1) + nF];
            for (int j = 0; j < nF; j++) s -= X[i * (nF + 1) + j] * wv[nc + j];
            wv[i] = s;
        }
        __syncthreads();
        for (int i = t; i < nr; i += kTG) {
            double s = 0;
            for (int j = 0; j < nh; j++) s += P[i * nh + j] * wv[j];
            pc[i] = s;
        }
        __syncthreads();
        const int nl = d.nl1;
        const double* lx = d.l1;
        const double* lw = d.l1 + nl;
        const int nql = dim == 3 ? nl * nl * nl : nl * nl;
        double e2 = 0, u2 = 0;
        for (int q = t; q < nql; q += kTG) {
            const int i0 = q % nl, i1 = (q / nl) % nl, i2 = dim == 3 ? q / (nl * nl) : 0;
            const double xi[3] = {lx[i0], lx[i1], dim == 3 ? lx[i2] : 0.0};
            const double w = lw[i0] * lw[i1] * (dim == 3 ? lw[i2] : 1.0);
            double x[3];
            const double det = gen_cell_map(dim, CV, xi, x);
            double psi[35];
            gen_basis(d, ctr, hw, x, psi, nullptr);
            double ph = 0;
            for (int i = 0; i < nr; i++) ph += pc[i] * psi[i];
            const double u = exact_at(d, x[0], x[1], x[2]);
            e2 += w * det * (u - ph) * (u - ph);
            u2 += w * det * u * u;
        }
        // workgroup reduction in a fixed order (per wavefront, then wavefronts)
        for (int o = 32; o > 0; o >>= 1) {
            e2 += __shfl_xor(e2, o, 64);
            u2 += __shfl_xor(u2, o, 64);
        }
        if ((t & 63) == 0) {
            red[2 * (t >> 6)] = e2;
            red[2 * (t >> 6) + 1] = u2;
        }
        __syncthreads();
        if (t == 0) {
            double se = 0, su = 0;
            for (int wv2 = 0; wv2 < kTG / 64; wv2++) {
                se += red[2 * wv2];
                su += red[2 * wv2 + 1];
            }
            d.cell_err[2 * (size_t)c] = se;
            d.cell_err[2 * (size_t)c + 1] = su;
        }
        return;
    }
    // ---- condensed blocks and right-hand side, added into the global system
    const int* blk = d.blk + (size_t)c * m * m;
    const int bsz = nfb * nfb;
    for (int e = t; e < nF * nF; e += kTG) {
        const int r = e / nF, s2 = e - r * nF;
        const int f = r / nfb, aa = r - f * nfb, g = s2 / nfb, bb = s2 - g * nfb;
        const int p = blk[f * m + g];
        const int ep = d.elem_vals != nullptr ? elem_entry(nF, r, s2) : -1;
        if (p < 0 && ep < 0) continue;
        double v = A[(nc + r) * nh + nc + s2];
        for (int i = 0; i < nc; i++) v -= A[i * nh + nc + r] * X[i * (nF + 1) + s2];
        if (ep >= 0) d.elem_vals[(size_t)d.elem_pos[c] * d.elem_stride + ep] = v;
        if (p >= 0) atomicAdd(d.vals + (size_t)p * bsz + aa * nfb + bb, v);
    }
    for (int r = t; r < nF; r += kTG) {
        const int f = r / nfb, aa = r - f * nfb;
        const int row = frow[f];
        if (row < 0) continue;
        double v = 0;
        for (int i = 0; i < nc; i++) v -= A[i * nh + nc + r] * X[i * (nF + 1) + nF];
        atomicAdd(d.b + (size_t)row * nfb + aa, v);
    }
}

}  // namespace

size_t hho_local_lds_bytes(const HhoDev& d) {
    if (d.kind == 2) return (size_t)gen_layout(d).total * sizeof(double);
    const size_t m = d.nfc, nr = d.nr, nc = d.nc, nfb = d.nfb, nh = nc + m * nfb, nF = m * nfb, n1 = nr - 1;
    const size_t n = nr * nr + 2 * nr * nh + nh * nh + m * nfb * nr + m * nfb * nh + std::max(n1 * n1, nc * nc) +
                     nc * (nF + 1) + d.nq + m * 8 + nc;
    return n * sizeof(double);
}

static void launch_hho_general(Ctx& c, const HhoDev& d) {
    const GenLayout Lo = gen_layout(d);
    const int nh = d.nc + d.nfc * d.nfb;
    FH_CHECK(d.nr <= 35 && d.nfb <= 10 && d.nvc <= 8 && d.nfc <= 6 && d.ng1 <= 6 && d.nl1 <= 8, -3,
             "general cells: degree k <= 3 in 3-D");
    FH_CHECK(d.nvf <= 4 && d.nfc <= 6, -3, "general cells: at most 6 faces of at most 4 corners");
    FH_CHECK(!d.post || (d.xsol && d.cell_err), -2, "post-processing needs the face solution and the cell output");
    FH_CHECK(nh * nh >= nh + d.nr, -3, "post-processing scratch does not fit");
    const size_t lds = (size_t)Lo.total * sizeof(double);
    int dev = 0, maxlds = 0;
    FH_HIP(hipGetDevice(&dev));
    FH_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    FH_CHECK(lds <= (size_t)maxlds, -3,
             "cell matrices (" + std::to_string(lds) + " bytes) exceed the LDS of a workgroup (" + std::to_string(maxlds) + ")");
    if (lds > 65536)
        FH_HIP(hipFuncSetAttribute((const void*)k_hho_general, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    DevBuf err(sizeof(int));
    FH_HIP(hipMemsetAsync(err.p, 0, sizeof(int), c.stream));
    hipLaunchKernelGGL(k_hho_general, dim3(d.ncell), dim3(kTG), lds, c.stream, d, err.as<int>());
    FH_HIP(hipGetLastError());
    int h = 0;
    FH_HIP(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    FH_HIP(hipStreamSynchronize(c.stream));
    FH_CHECK(!(h & 2), -6, "a cell of the HHO assembly is inverted at a quadrature point");
    FH_CHECK(h == 0, -6, "a local matrix of the HHO assembly is not SPD");
}

void launch_hho_local(Ctx& c, const HhoDev& d) {
    if (d.kind == 2) {
        launch_hho_general(c, d);
        return;
    }
    FH_CHECK(d.nfc <= 8, -3, "at most 8 faces per cell");
    {
        const size_t nh = (size_t)d.nc + (size_t)d.nfc * d.nfb;
        FH_CHECK((size_t)d.nfc * d.nfb * nh >= nh + (size_t)d.nr, -3, "post-processing scratch does not fit");
        FH_CHECK(!d.post || (d.xsol && d.cell_err), -2, "post-processing needs the face solution and cell output");
    }
    const size_t lds = hho_local_lds_bytes(d);
    int dev = 0, maxlds = 0;
    FH_HIP(hipGetDevice(&dev));
    FH_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    FH_CHECK(lds + 64 <= (size_t)maxlds, -3,
             "cell matrices (" + std::to_string(lds) + " bytes) exceed the LDS of a workgroup (" +
                 std::to_string(maxlds) + "); use the host assembly");
    if (lds > 65536)
        FH_HIP(hipFuncSetAttribute((const void*)k_hho_local, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    DevBuf err(sizeof(int));
    FH_HIP(hipMemsetAsync(err.p, 0, sizeof(int), c.stream));
    hipLaunchKernelGGL(k_hho_local, dim3(d.ncell), dim3(kT), lds, c.stream, d, err.as<int>());
    FH_HIP(hipGetLastError());
    int h = 0;
    FH_HIP(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    FH_HIP(hipStreamSynchronize(c.stream));
    FH_CHECK(h == 0, -6, "a local matrix of the HHO assembly is not SPD");
}

}  // namespace fh
