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
        if (p < 0) continue;
        double v = A[(nc + r) * nh + nc + s];
        for (int i = 0; i < nc; i++) v -= A[i * nh + nc + r] * X[i * (nF + 1) + s];
        atomicAdd(d.vals + (size_t)p * bsz + a * nfb + bb, v);
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

}  // namespace

size_t hho_local_lds_bytes(const HhoDev& d) {
    const size_t m = d.nfc, nr = d.nr, nc = d.nc, nfb = d.nfb, nh = nc + m * nfb, nF = m * nfb, n1 = nr - 1;
    const size_t n = nr * nr + 2 * nr * nh + nh * nh + m * nfb * nr + m * nfb * nh + std::max(n1 * n1, nc * nc) +
                     nc * (nF + 1) + d.nq + m * 8 + nc;
    return n * sizeof(double);
}

void launch_hho_local(Ctx& c, const HhoDev& d) {
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
