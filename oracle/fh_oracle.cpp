// CPU oracle of the condensed-HHO solve path -- TEST INFRASTRUCTURE ONLY.
//
// A plain restatement of the reference algorithms, written for this
// repository from their description (no reference source is copied).  Each
// function names the reference file:line it follows.  Tests, the smoke check
// and the CPU-baseline leg of bench.py are its only users.
#include "fh_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

struct Csr {
    int n = 0, m = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
    const int* prp = nullptr;
    const int* pci = nullptr;
    const double* pv = nullptr;
    // block storage instead of the scalar arrays (bs > 0): n / bs block rows,
    // block row pointers brp, block columns bci, dense bs x bs row-major
    // blocks bv (64-bit value offsets: the finest level of the hexahedral
    // k = 3 system holds 2.9e9 values)
    int bs = 0;
    const int* brp = nullptr;
    const int* bci = nullptr;
    const double* bv = nullptr;
    void view(int n_, int m_, const int* r, const int* c, const double* val) {
        n = n_; m = m_; prp = r; pci = c; pv = val;
    }
    void own() { prp = rp.data(); pci = ci.data(); pv = v.data(); }
};

// y = A x over the block rows of a block-stored matrix.
void bspmv(const Csr& A, const double* x, double* y) {
    const int bs = A.bs, nb = A.n / bs;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < nb; I++) {
        double acc[64];
        for (int a = 0; a < bs; a++) acc[a] = 0;
        for (int p = A.brp[I]; p < A.brp[I + 1]; p++) {
            const double* blk = A.bv + (size_t)p * bs * bs;
            const double* xj = x + (size_t)A.bci[p] * bs;
            for (int a = 0; a < bs; a++) {
                double t = 0;
                for (int c = 0; c < bs; c++) t += blk[a * bs + c] * xj[c];
                acc[a] += t;
            }
        }
        for (int a = 0; a < bs; a++) y[(size_t)I * bs + a] = acc[a];
    }
}

double dot(int n, const double* a, const double* b) {
    double s = 0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}
double nrm2(int n, const double* a) { return std::sqrt(dot(n, a, a)); }

void spmv(const Csr& A, const double* x, double* y) {
    if (A.bs) {
        bspmv(A, x, y);
        return;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A.n; i++) {
        double s = 0;
        for (int p = A.prp[i]; p < A.prp[i + 1]; p++) s += A.pv[p] * x[A.pci[p]];
        y[i] = s;
    }
}

// Gauss-Jordan inverse with partial pivoting of an nf x nf row-major block.
bool invert_block(int nf, double* a, double* inv) {
    for (int i = 0; i < nf; i++)
        for (int j = 0; j < nf; j++) inv[i * nf + j] = i == j ? 1.0 : 0.0;
    for (int c = 0; c < nf; c++) {
        int p = c;
        for (int r = c + 1; r < nf; r++)
            if (std::fabs(a[r * nf + c]) > std::fabs(a[p * nf + c])) p = r;
        if (a[p * nf + c] == 0.0) return false;
        if (p != c)
            for (int j = 0; j < nf; j++) {
                std::swap(a[p * nf + j], a[c * nf + j]);
                std::swap(inv[p * nf + j], inv[c * nf + j]);
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
    return true;
}

int block_inverse(const Csr& A, int nf, std::vector<double>& invD) {
    int nb = A.n / nf;
    invD.assign((size_t)nb * nf * nf, 0.0);
    int bad = 0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nb; i++) {
        std::vector<double> D(nf * nf, 0.0);
        if (A.bs) {
            for (int p = A.brp[i]; p < A.brp[i + 1]; p++)
                if (A.bci[p] == i) std::copy(A.bv + (size_t)p * nf * nf, A.bv + (size_t)(p + 1) * nf * nf, D.begin());
        } else
        for (int k = 0; k < nf; k++) {
            int row = i * nf + k;
            for (int p = A.prp[row]; p < A.prp[row + 1]; p++) {
                int j = A.pci[p];
                if (j / nf == i) D[k * nf + (j - i * nf)] = A.pv[p];
            }
        }
        if (!invert_block(nf, D.data(), &invD[(size_t)i * nf * nf])) {
#pragma omp critical
            if (!bad) bad = -(1 + i);
        }
    }
    return bad;
}

// BlockJacobi::ProcessBlockRow (BlockJacobi.h:101-127)
void bj_sweep(const Csr& A, int nf, const double* invD, double omega, const double* b, double* x) {
    int nb = A.n / nf;
    std::vector<double> xold(x, x + A.n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nb; i++) {
        double tmp[64];
        for (int k = 0; k < nf; k++) {
            int row = i * nf + k;
            double t = omega * b[row];
            for (int p = A.prp[row]; p < A.prp[row + 1]; p++) {
                int j = A.pci[p];
                if (j / nf == i) t += (1 - omega) * A.pv[p] * xold[j];
                else t += -omega * A.pv[p] * xold[j];
            }
            tmp[k] = t;
        }
        const double* Di = invD + (size_t)i * nf * nf;
        for (int k = 0; k < nf; k++) {
            double s = 0;
            for (int l = 0; l < nf; l++) s += Di[k * nf + l] * tmp[l];
            x[i * nf + k] = s;
        }
    }
}

// BlockGaussSeidel::ProcessBlockRow (BlockGaussSeidel.h:193-223); with
// omega != 1 BlockSOR::ProcessBlockRow (BlockSOR.h:193-223):
// x_i <- D_i^-1 (omega b_i + (1 - omega) D_i x_i - omega sum_{j != i} A_ij x_j)
inline void sor_rhs(int nf, const int* rp, const int* ci, const double* v, int i, double omega, const double* b,
                    const double* x, double* tmp) {
    for (int k = 0; k < nf; k++) {
        int row = i * nf + k;
        if (omega == 1.0) {
            double t = b[row];
            for (int p = rp[row]; p < rp[row + 1]; p++)
                if (ci[p] / nf != i) t -= v[p] * x[ci[p]];
            tmp[k] = t;
        } else {
            double t = omega * b[row];
            for (int p = rp[row]; p < rp[row + 1]; p++) {
                if (ci[p] / nf == i) t += (1 - omega) * v[p] * x[ci[p]];
                else t += -omega * v[p] * x[ci[p]];
            }
            tmp[k] = t;
        }
    }
}

inline void bgs_row(const Csr& A, int nf, const double* invD, int i, const double* b, double* x,
                    double omega = 1.0) {
    double tmp[64];
    sor_rhs(nf, A.prp, A.pci, A.pv, i, omega, b, x, tmp);
    const double* Di = invD + (size_t)i * nf * nf;
    for (int k = 0; k < nf; k++) {
        double s = 0;
        for (int l = 0; l < nf; l++) s += Di[k * nf + l] * tmp[l];
        x[i * nf + k] = s;
    }
}

// BlockGaussSeidel::ExecuteOneIteration, non-hybrid (BlockGaussSeidel.h:84-123),
// BlockSOR::ExecuteOneIteration (BlockSOR.h:98-188); dir 0 forward,
// 1 backward, 2 symmetric
void bgs_sweep(const Csr& A, int nf, const double* invD, int dir, const double* b, double* x, double omega = 1.0) {
    int nb = A.n / nf;
    if (dir == 0 || dir == 2)
        for (int i = 0; i < nb; i++) bgs_row(A, nf, invD, i, b, x, omega);
    if (dir == 1 || dir == 2)
        for (int i = nb - 1; i >= 0; i--) bgs_row(A, nf, invD, i, b, x, omega);
}

// Dense Cholesky (coarse solver; the reference uses SimplicialLDLT,
// EigenSparseCholesky.h:19-41, which solves the same SPD system).
struct Dense {
    int n = 0;
    std::vector<double> L;
    bool factor(const Csr& A) {
        n = A.n;
        L.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; i++)
            for (int p = A.prp[i]; p < A.prp[i + 1]; p++) L[(size_t)i * n + A.pci[p]] = A.pv[p];
        for (int j = 0; j < n; j++) {
            double d = L[(size_t)j * n + j];
            for (int k = 0; k < j; k++) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
            if (d <= 0) return false;
            d = std::sqrt(d);
            L[(size_t)j * n + j] = d;
            for (int i = j + 1; i < n; i++) {
                double s = L[(size_t)i * n + j];
                for (int k = 0; k < j; k++) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
                L[(size_t)i * n + j] = s / d;
            }
        }
        return true;
    }
    void solve(const double* b, double* x) const {
        std::vector<double> y(n);
        for (int i = 0; i < n; i++) {
            double s = b[i];
            for (int k = 0; k < i; k++) s -= L[(size_t)i * n + k] * y[k];
            y[i] = s / L[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = y[i];
            for (int k = i + 1; k < n; k++) s -= L[(size_t)k * n + i] * x[k];
            x[i] = s / L[(size_t)i * n + i];
        }
    }
};

// ----------------------------------------------------- pairwise aggregation
struct AElem {
    std::vector<std::pair<int, double>> nbr;  // sorted by ascending coupling
    std::vector<int> strong;
    int count = 0;   // NElementsIAmStrongNeighbourOf
    int agg = -1;
};

// AlgebraicMesh::Build (AlgebraicMesh.h:54-113), CouplingValue (:133-148),
// IsStronglyCoupled (:150-153)
void build_algebraic_mesh(const Csr& A, int nf, double theta, std::vector<AElem>& E) {
    int nb = A.n / nf;
    E.assign(nb, AElem());
    for (int i = 0; i < nb; i++) {
        AElem& e = E[i];
        for (int k = 0; k < nf; k++) {
            int row = i * nf + k;
            for (int p = A.prp[row]; p < A.prp[row + 1]; p++) {
                int nbn = A.pci[p] / nf;
                if (nbn == i) continue;
                bool seen = false;
                for (auto& q : e.nbr) if (q.first == nbn) { seen = true; break; }
                if (seen) continue;
                // trace of block (i, nbn), summed over the diagonal in order
                double tr = 0;
                bool first = true;
                for (int a = 0; a < nf; a++) {
                    int r = i * nf + a, c = nbn * nf + a;
                    double val = 0;
                    for (int q = A.prp[r]; q < A.prp[r + 1]; q++) if (A.pci[q] == c) { val = A.pv[q]; break; }
                    if (first) { tr = val; first = false; } else tr += val;
                }
                e.nbr.push_back({nbn, tr});
            }
        }
        std::sort(e.nbr.begin(), e.nbr.end(),
                  [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.second < b.second; });
    }
    for (int i = 0; i < nb; i++) {
        AElem& e = E[i];
        for (auto& q : e.nbr) {
            if (q.second < 0 && q.second < theta * e.nbr[0].second) {
                e.strong.push_back(q.first);
                E[q.first].count++;
            } else
                break;
        }
    }
}

// PairwiseAggregation::Perform / OrderElementsByPriority / NumberNeighbours /
// StrongestAvailableNeighbour / AddToAggregate (PairwiseAggregation.h:9-205)
int pairwise_aggregate(std::vector<AElem>& E, bool& coarsest) {
    int n = (int)E.size();
    std::vector<int> order(n, -1);
    std::vector<char> numbered(n, 0);
    int on = 0;
    int start = 0, mn = E[0].count;
    for (int i = 0; i < n; i++) if (E[i].count < mn) { start = i; mn = E[i].count; }
    order[on++] = start;
    numbered[start] = 1;
    for (int i = 0; i < n; i++) {
        if (i >= on) {  // disconnected graph: restart from the next unnumbered element
            int s = 0;
            while (numbered[s]) s++;
            order[on++] = s;
            numbered[s] = 1;
        }
        const AElem& e = E[order[i]];
        std::vector<int> un;
        for (auto& q : e.nbr) if (!numbered[q.first]) un.push_back(q.first);
        if (un.empty()) continue;
        std::sort(un.begin(), un.end(), [&E](int a, int b) { return E[a].count < E[b].count; });
        for (int u : un) { order[on++] = u; numbered[u] = 1; }
    }
    int nagg = 0;
    int remaining = n;
    auto add = [&E](int el, int a) {
        E[el].agg = a;
        for (int s : E[el].strong) if (E[s].agg < 0) E[s].count--;
    };
    for (int idx = 0; idx < n; idx++) {
        int el = order[idx];
        if (E[el].agg >= 0) continue;
        int a = nagg++;
        add(el, a);
        remaining--;
        double mnc = 100000000;
        int chosen = -1;
        for (int s : E[el].strong)
            if (E[s].agg < 0 && E[s].count < mnc) { chosen = s; mnc = E[s].count; }
        if (chosen >= 0) { add(chosen, a); remaining--; }
        if (remaining == 0) break;
    }
    coarsest = (double)n / (double)nagg <= 1.01;
    return nagg;
}

// AggregLevel::GalerkinOperator (AggregAMG.h:156-180) extended to blocks:
// Ac(I,J) = sum_{k in I, l in J} A_kl, entries |a| <= 1e-15 dropped before
// summation (NonZeroCoefficients.h:36-40), summed in row-major order.
void galerkin(const Csr& A, int nf, const std::vector<int>& agg, int nagg, Csr& Ac) {
    int nb = A.n / nf;
    std::vector<std::vector<int>> members(nagg);
    for (int i = 0; i < nb; i++) members[agg[i]].push_back(i);
    Ac.n = Ac.m = nagg * nf;
    Ac.rp.assign(Ac.n + 1, 0);
    Ac.ci.clear();
    Ac.v.clear();
    std::vector<double> acc((size_t)Ac.n, 0.0);
    std::vector<char> used((size_t)Ac.n, 0);
    std::vector<int> cols;
    for (int I = 0; I < nagg; I++) {
        for (int a = 0; a < nf; a++) {
            cols.clear();
            for (int k : members[I]) {
                int row = k * nf + a;
                for (int p = A.prp[row]; p < A.prp[row + 1]; p++) {
                    double val = A.pv[p];
                    if (!(std::fabs(val) > 1e-15)) continue;
                    int l = A.pci[p];
                    int c = agg[l / nf] * nf + (l % nf);
                    if (!used[c]) { used[c] = 1; acc[c] = val; cols.push_back(c); }
                    else acc[c] += val;
                }
            }
            std::sort(cols.begin(), cols.end());
            for (int c : cols) { Ac.ci.push_back(c); Ac.v.push_back(acc[c]); used[c] = 0; }
            Ac.rp[I * nf + a + 1] = (int)Ac.ci.size();
        }
    }
    Ac.own();
}

// ----------------------------------------------------------------- multigrid
enum { SM_BJ = 1, SM_BGS_F = 2, SM_BGS_B = 3, SM_BGS_S = 4, SM_MC_F = 5, SM_MC_B = 6, SM_MC_S = 7,
       SM_CHEB2 = 8, SM_CHEB3 = 9, SM_CHEB4 = 10, SM_CHEB1 = 11,
       // alternating directions (forward first / backward first), lexicographic and multicolour
       SM_BGS_AF = 12, SM_BGS_AB = 13, SM_MC_AF = 14, SM_MC_AB = 15 };
inline bool sm_is_cheb(int k) { return k >= SM_CHEB2 && k <= SM_CHEB1; }

// out = D^-1 (b - A x), block by block
void dinv_resid(const Csr& A, int nf, const double* invD, const double* b, const double* x, double* out) {
    int nb = A.n / nf;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nb; i++) {
        double tmp[64];
        if (A.bs) {
            for (int k = 0; k < nf; k++) tmp[k] = b ? b[(size_t)i * nf + k] : 0.0;
            for (int p = A.brp[i]; p < A.brp[i + 1]; p++) {
                const double* blk = A.bv + (size_t)p * nf * nf;
                const double* xj = x + (size_t)A.bci[p] * nf;
                for (int k = 0; k < nf; k++) {
                    double t = 0;
                    for (int c = 0; c < nf; c++) t += blk[k * nf + c] * xj[c];
                    tmp[k] -= t;
                }
            }
        } else
        for (int k = 0; k < nf; k++) {
            int row = i * nf + k;
            double t = b ? b[row] : 0.0;
            for (int p = A.prp[row]; p < A.prp[row + 1]; p++) t -= A.pv[p] * x[A.pci[p]];
            tmp[k] = t;
        }
        const double* Di = invD + (size_t)i * nf * nf;
        for (int k = 0; k < nf; k++) {
            double s = 0;
            for (int l = 0; l < nf; l++) s += Di[k * nf + l] * tmp[l];
            out[i * nf + k] = s;
        }
    }
}

// Power iteration for lambda_max(D^-1 A): 20 steps from v_i = 1 + (7919 i mod 13) / 13.
double cheb_lmax(const Csr& A, int nf, const double* invD) {
    const int n = A.n;
    std::vector<double> v(n), w(n);
    for (int i = 0; i < n; i++) v[i] = 1.0 + (double)(((int64_t)i * 7919) % 13) / 13.0;
    double lam = 0;
    for (int it = 0; it < 20; it++) {
        dinv_resid(A, nf, invD, nullptr, v.data(), w.data());  // w = -D^-1 A v
        double nw = 0, nv = 0;
        for (int i = 0; i < n; i++) { w[i] = -w[i]; nw += w[i] * w[i]; nv += v[i] * v[i]; }
        nw = std::sqrt(nw); nv = std::sqrt(nv);
        lam = nw / nv;
        for (int i = 0; i < n; i++) v[i] = w[i] / nw;
    }
    return lam;
}

int greedy_colour_blocks(const Csr& A, int nf, std::vector<int>& cptr, std::vector<int>& rows);

struct Level {
    Csr A;
    int nf = 1;          // block size of this level
    Csr P, R;            // GMG transfer (level -> coarser)
    std::vector<int> agg; int nagg = 0;  // AMG transfer
    bool amg = false;
    bool ptrunc = false; // p-transfer: keep the first nfc modes of every block (hierarchical bases)
    int nfc = 1;
    std::vector<double> invD;
    std::vector<int> cptr, crows;   // multicolour ordering (greedy, block-row order)
    double lmax = 0;                 // Chebyshev: lambda_max(D^-1 A) estimate
};

struct MG {
    std::vector<Level> lv;
    Dense coarse;
    const or_solver_cfg* cfg = nullptr;

    void smooth(Level& L, int kind, double omega, int nu, const double* b, double* x) {
        const int nf = L.nf;
        for (int s = 0; s < nu; s++) {
            if (sm_is_cheb(kind)) {
                // Chebyshev polynomial of D^-1 A on [0.25, 1.1] x lambda_max estimate
                if (L.lmax == 0) L.lmax = cheb_lmax(L.A, nf, L.invD.data());
                const int deg = kind == SM_CHEB1 ? 1 : 2 + (kind - SM_CHEB2);
                const double hi = 1.1 * L.lmax, lo = 0.25 * L.lmax;
                const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
                double rho = 1.0 / sigma;
                const int n = L.A.n;
                std::vector<double> r(n), d(n, 0.0);
                for (int k = 0; k < deg; k++) {
                    dinv_resid(L.A, nf, L.invD.data(), b, x, r.data());
                    if (k == 0)
                        for (int i = 0; i < n; i++) { d[i] = r[i] / theta; x[i] += d[i]; }
                    else {
                        const double rn = 1.0 / (2.0 * sigma - rho);
                        for (int i = 0; i < n; i++) { d[i] = rn * rho * d[i] + 2.0 * rn / delta * r[i]; x[i] += d[i]; }
                        rho = rn;
                    }
                }
                continue;
            }
            if (L.A.bs) {  // block-stored levels: Chebyshev smoothing only
                std::fprintf(stderr, "oracle: block-stored levels are smoothed by Chebyshev only\n");
                std::abort();
            }
            // alternating: forward on even smoothing iterations when forward
            // comes first (IterationCount % 2, BlockGaussSeidel.h:105-120)
            int k2 = kind;
            if (kind == SM_BGS_AF || kind == SM_BGS_AB)
                k2 = s % 2 == (kind == SM_BGS_AF ? 0 : 1) ? SM_BGS_F : SM_BGS_B;
            if (kind == SM_MC_AF || kind == SM_MC_AB)
                k2 = s % 2 == (kind == SM_MC_AF ? 0 : 1) ? SM_MC_F : SM_MC_B;
            if (k2 == SM_BJ) bj_sweep(L.A, nf, L.invD.data(), omega, b, x);
            else if (k2 >= SM_MC_F && k2 <= SM_MC_S) {
                if (L.cptr.empty()) greedy_colour_blocks(L.A, nf, L.cptr, L.crows);
                int nc = (int)L.cptr.size() - 1;
                if (k2 != SM_MC_B)
                    or_mcsor_sweep(L.A.n, nf, L.A.prp, L.A.pci, L.A.pv, L.invD.data(), nc, L.cptr.data(), L.crows.data(),
                                   0, omega, b, x);
                if (k2 != SM_MC_F)
                    or_mcsor_sweep(L.A.n, nf, L.A.prp, L.A.pci, L.A.pv, L.invD.data(), nc, L.cptr.data(), L.crows.data(),
                                   1, omega, b, x);
            } else bgs_sweep(L.A, nf, L.invD.data(), k2 - SM_BGS_F, b, x, omega);
        }
    }
    void restrict_(int l, const std::vector<double>& r, std::vector<double>& rc) {
        Level& L = lv[l];
        const int nf = L.nf;
        if (L.ptrunc) {
            // p-restriction = truncation to the first nfc modes (LevelForHHO.h:586-620)
            int nb = L.A.n / nf;
            rc.assign((size_t)nb * L.nfc, 0.0);
            for (int i = 0; i < nb; i++)
                for (int a = 0; a < L.nfc; a++) rc[(size_t)i * L.nfc + a] = r[(size_t)i * nf + a];
            return;
        }
        if (!L.amg) {
            rc.assign(L.R.n, 0.0);
            spmv(L.R, r.data(), rc.data());
        } else {
            // AggregLevel::Restrict (AggregAMG.h:237-271), blockwise
            rc.assign((size_t)L.nagg * nf, 0.0);
            int nb = L.A.n / nf;
            std::vector<char> first((size_t)L.nagg, 1);
            for (int i = 0; i < nb; i++)
                for (int a = 0; a < nf; a++) rc[(size_t)L.agg[i] * nf + a] += r[(size_t)i * nf + a];
        }
    }
    void prolong_add(int l, const std::vector<double>& ec, double* x) {
        Level& L = lv[l];
        const int nf = L.nf;
        if (L.ptrunc) {
            // p-prolongation = injection of the low modes (LevelForHHO.h:586-620)
            int nb = L.A.n / nf;
            for (int i = 0; i < nb; i++)
                for (int a = 0; a < L.nfc; a++) x[(size_t)i * nf + a] += ec[(size_t)i * L.nfc + a];
            return;
        }
        if (!L.amg) {
            std::vector<double> t(L.P.n);
            spmv(L.P, ec.data(), t.data());
            for (int i = 0; i < L.P.n; i++) x[i] += t[i];
        } else {
            int nb = L.A.n / nf;
            for (int i = 0; i < nb; i++)
                for (int a = 0; a < nf; a++) x[(size_t)i * nf + a] += ec[(size_t)L.agg[i] * nf + a];
        }
    }
    // smoothing iterations of level l (Multigrid::Setup, Multigrid.h:224-242:
    // each coarse level's count is its finer level's changed by op, coeff)
    // (not in the reference: coarse_nu, the given count of the levels below
    // the finest, when coarse_nu_given)
    int level_nu(int nu, int l, int coarse_nu) const {
        if (l > 0 && cfg->coarse_nu_given) return coarse_nu;
        for (int k = 0; k < l; k++) {
            if (cfg->nu_change_op == '+') nu += cfg->nu_change_coeff;
            else if (cfg->nu_change_op == '-') nu -= cfg->nu_change_coeff;
            else if (cfg->nu_change_op == '*') nu *= cfg->nu_change_coeff;
        }
        return nu < 0 ? 0 : nu;
    }
    // Multigrid::MultigridCycle (Multigrid.h:438-608)
    void cycle(int l, const double* b, double* x, double* Ax) {
        Level& L = lv[l];
        int n = L.A.n;
        if (l == (int)lv.size() - 1) {
            coarse.solve(b, x);
            if (Ax) spmv(L.A, x, Ax);
            return;
        }
        const bool cs = l > 0 && cfg->coarse_kind > 0;
        smooth(L, cs ? cfg->coarse_kind : cfg->pre_kind, cs ? cfg->coarse_omega : cfg->pre_omega, level_nu(cfg->nu1, l, cfg->coarse_nu1),
               b, x);
        std::vector<double> r(n);
        spmv(L.A, x, r.data());
        for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
        std::vector<double> rc, ec;
        restrict_(l, r, rc);
        int nc = (int)rc.size();
        ec.assign(nc, 0.0);
        if (cfg->cycle == 'K' && l + 1 < (int)lv.size() - 1 && (cfg->kcycle_levels <= 0 || l + 1 <= cfg->kcycle_levels))
            fcg_kcycle(l + 1, rc, ec);
        else {
            int loops = cfg->cycle == 'W' ? std::max(2, cfg->wloops) : 1;
            for (int w = 0; w < loops; w++) {
                cycle(l + 1, rc.data(), ec.data(), nullptr);
                if (l + 1 == (int)lv.size() - 1) break;
            }
        }
        prolong_add(l, ec, x);
        smooth(L, cs ? cfg->coarse_kind : cfg->post_kind, cs ? cfg->coarse_omega : cfg->post_omega,
               level_nu(cfg->nu2, l, cfg->coarse_nu2), b, x);
        if (Ax) spmv(L.A, x, Ax);
    }
    // Multigrid::FCGForKCycle (Multigrid.h:611-656); r is overwritten
    void fcg_kcycle(int l, std::vector<double>& r, std::vector<double>& out) {
        const double t = 0.25;
        int n = (int)r.size();
        std::vector<double> c(n, 0.0), v(n);
        cycle(l, r.data(), c.data(), v.data());
        double rho1 = dot(n, c.data(), v.data());
        double alpha1 = dot(n, c.data(), r.data());
        double rold = nrm2(n, r.data());
        for (int i = 0; i < n; i++) r[i] -= (alpha1 / rho1) * v[i];
        double rnew = nrm2(n, r.data());
        if (rnew <= t * rold) {
            for (int i = 0; i < n; i++) out[i] = (alpha1 / rho1) * c[i];
            return;
        }
        std::vector<double> d(n, 0.0), w(n);
        cycle(l, r.data(), d.data(), w.data());
        double gamma = dot(n, d.data(), v.data());
        double beta = dot(n, d.data(), w.data());
        double alpha2 = dot(n, d.data(), r.data());
        double rho2 = beta - gamma * gamma / rho1;
        double ca = alpha1 / rho1 - gamma * alpha2 / (rho1 * rho2), cb = alpha2 / rho2;
        for (int i = 0; i < n; i++) out[i] = ca * c[i] + cb * d[i];
    }
};

double g_setup_s = 0, g_solve_s = 0;

// Greedy colouring of the block graph of A (block rows in increasing order
// take the smallest colour unused by lower-numbered neighbours); rows grouped
// by colour in increasing order.
int greedy_colour_blocks(const Csr& A, int nf, std::vector<int>& cptr, std::vector<int>& rows) {
    int nb = A.n / nf;
    std::vector<int> brp(nb + 1, 0), bci;
    std::vector<int> mark(nb, -1);
    for (int i = 0; i < nb; i++) {
        std::vector<int> cols;
        for (int k = 0; k < nf; k++)
            for (int p = A.prp[i * nf + k]; p < A.prp[i * nf + k + 1]; p++) {
                int j = A.pci[p] / nf;
                if (mark[j] != i) { mark[j] = i; cols.push_back(j); }
            }
        std::sort(cols.begin(), cols.end());
        bci.insert(bci.end(), cols.begin(), cols.end());
        brp[i + 1] = (int)bci.size();
    }
    std::vector<int> col(nb);
    int nc = or_greedy_colouring(nb, brp.data(), bci.data(), col.data());
    cptr.assign(nc + 1, 0);
    for (int c : col) cptr[c + 1]++;
    for (int c = 0; c < nc; c++) cptr[c + 1] += cptr[c];
    rows.assign(nb, 0);
    std::vector<int> pos(cptr.begin(), cptr.end() - 1);
    for (int i = 0; i < nb; i++) rows[pos[col[i]]++] = i;
    return nc;
}

// Aggregation hierarchy.  For blocks (nf > 1) with pfirst, the first
// coarsening is a p-coarsening to the lowest mode of every block
// (hierarchical face bases; HighOrderAggregAMG, HighOrderAggregAMG.h:28-45,
// with unsmoothed injection) and scalar pairwise aggregation follows.
int build_amg(MG& mg, const Csr& A0, int nf, double theta, int max_coarse, int max_levels, bool pfirst) {
    mg.lv.clear();
    mg.lv.emplace_back();
    mg.lv[0].A = A0;
    mg.lv[0].nf = nf;
    if (nf > 1 && pfirst && A0.n > max_coarse) {
        Level& L = mg.lv[0];
        L.ptrunc = true;
        L.nfc = 1;
        Level C;
        int nb = A0.n / nf;
        C.A.n = C.A.m = nb;
        C.A.rp.assign(nb + 1, 0);
        for (int i = 0; i < nb; i++) {
            if (A0.bs) {
                for (int p = A0.brp[i]; p < A0.brp[i + 1]; p++) {
                    C.A.ci.push_back(A0.bci[p]);
                    C.A.v.push_back(A0.bv[(size_t)p * nf * nf]);
                }
            } else {
                int row = i * nf;
                for (int p = A0.prp[row]; p < A0.prp[row + 1]; p++)
                    if (A0.pci[p] % nf == 0) { C.A.ci.push_back(A0.pci[p] / nf); C.A.v.push_back(A0.pv[p]); }
            }
            C.A.rp[i + 1] = (int)C.A.ci.size();
        }
        C.A.own();
        C.nf = 1;
        mg.lv.push_back(std::move(C));
        for (auto& q : mg.lv) if (!q.A.rp.empty()) q.A.own();
        nf = 1;
    }
    while ((int)mg.lv.size() < max_levels && mg.lv.back().A.n > max_coarse) {
        Level& L = mg.lv.back();
        // DoublePairwiseAggregation (AggregAMG.h:40-100)
        std::vector<AElem> E1;
        build_algebraic_mesh(L.A, nf, theta, E1);
        bool coarsest = false;
        int n1 = pairwise_aggregate(E1, coarsest);
        if (coarsest) break;
        std::vector<int> a1(E1.size());
        for (size_t i = 0; i < E1.size(); i++) a1[i] = E1[i].agg;
        Csr A1;
        galerkin(L.A, nf, a1, n1, A1);
        std::vector<AElem> E2;
        build_algebraic_mesh(A1, nf, theta, E2);
        int n2 = pairwise_aggregate(E2, coarsest);
        if (coarsest) break;
        std::vector<int> a2(E2.size());
        for (size_t i = 0; i < E2.size(); i++) a2[i] = E2[i].agg;
        L.amg = true;
        L.agg.resize(a1.size());
        for (size_t i = 0; i < a1.size(); i++) L.agg[i] = a2[a1[i]];
        L.nagg = n2;
        Level C;
        galerkin(A1, nf, a2, n2, C.A);
        C.nf = nf;
        mg.lv.push_back(std::move(C));
        // re-own pointers after the move of the vector storage
        for (auto& q : mg.lv) if (!q.A.rp.empty()) q.A.own();
    }
    return (int)mg.lv.size();
}

}  // namespace

extern "C" {

void or_spmv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    Csr A; A.view(n, n, rp, ci, v);
    spmv(A, x, y);
}

int or_block_inverse(int n, int nf, const int* rp, const int* ci, const double* v, double* invD) {
    Csr A; A.view(n, n, rp, ci, v);
    std::vector<double> inv;
    int rc = block_inverse(A, nf, inv);
    std::memcpy(invD, inv.data(), inv.size() * sizeof(double));
    return rc;
}

void or_bj_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                 double omega, const double* b, double* x) {
    Csr A; A.view(n, n, rp, ci, v);
    bj_sweep(A, nf, invD, omega, b, x);
}

void or_bgs_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                  int dir, const double* b, double* x) {
    Csr A; A.view(n, n, rp, ci, v);
    bgs_sweep(A, nf, invD, dir, b, x);
}

void or_bsor_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                   int dir, double omega, const double* b, double* x) {
    Csr A; A.view(n, n, rp, ci, v);
    bgs_sweep(A, nf, invD, dir, b, x, omega);
}

void or_mcbgs_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                    int ncol, const int* colour_ptr, const int* colour_rows, int dir,
                    const double* b, double* x) {
    or_mcsor_sweep(n, nf, rp, ci, v, invD, ncol, colour_ptr, colour_rows, dir, 1.0, b, x);
}

void or_mcsor_sweep(int n, int nf, const int* rp, const int* ci, const double* v, const double* invD,
                    int ncol, const int* colour_ptr, const int* colour_rows, int dir, double omega,
                    const double* b, double* x) {
    for (int cc = 0; cc < ncol; cc++) {
        int c = dir == 0 ? cc : ncol - 1 - cc;
        // rows of one colour are not coupled: their updates read only the
        // values of the other colours and may run in any order
#pragma omp parallel for schedule(static)
        for (int q = colour_ptr[c]; q < colour_ptr[c + 1]; q++) {
            int i = colour_rows[q];
            double tmp[64];
            sor_rhs(nf, rp, ci, v, i, omega, b, x, tmp);
            const double* Di = invD + (size_t)i * nf * nf;
            for (int k = 0; k < nf; k++) {
                double s = 0;
                for (int l = 0; l < nf; l++) s += Di[k * nf + l] * tmp[l];
                x[i * nf + k] = s;
            }
        }
    }
}

int or_greedy_colouring(int nb, const int* brp, const int* bci, int* colour) {
    int ncol = 0;
    std::vector<int> mark;
    for (int i = 0; i < nb; i++) {
        mark.assign(ncol + 1, 0);
        for (int p = brp[i]; p < brp[i + 1]; p++) {
            int j = bci[p];
            if (j < i) mark[colour[j]] = 1;
        }
        int c = 0;
        while (mark[c]) c++;
        colour[i] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    return ncol;
}

int or_pairwise_aggregate(int n, int nf, const int* rp, const int* ci, const double* v, double theta,
                          int* agg, int* coarsest) {
    Csr A; A.view(n, n, rp, ci, v);
    std::vector<AElem> E;
    build_algebraic_mesh(A, nf, theta, E);
    bool c = false;
    int na = pairwise_aggregate(E, c);
    for (size_t i = 0; i < E.size(); i++) agg[i] = E[i].agg;
    *coarsest = c;
    return na;
}

int or_amg_hierarchy(int n, int nf, const int* rp, const int* ci, const double* v, double theta,
                     int max_coarse, int* agg0, int* level_n, int max_levels) {
    Csr A; A.view(n, n, rp, ci, v);
    MG mg;
    int nl = build_amg(mg, A, nf, theta, max_coarse, max_levels, true);
    for (int l = 0; l < nl; l++) level_n[l] = mg.lv[l].A.n;
    if (nl > 1) std::memcpy(agg0, mg.lv[0].agg.data(), mg.lv[0].agg.size() * sizeof(int));
    return nl;
}

void or_last_times(double* setup_s, double* solve_s) {
    *setup_s = g_setup_s;
    *solve_s = g_solve_s;
}

static int solve_impl(const or_solver_cfg* cfg, const Csr& A, const or_level_desc* levels, const double* b, double* x,
                      double* hist, or_solver_out* out);

int or_solve(const or_solver_cfg* cfg, const or_level_desc* levels, const double* b, double* x,
             double* hist, or_solver_out* out) {
    Csr A; A.view(levels[0].n, levels[0].n, levels[0].rp, levels[0].ci, levels[0].v);
    return solve_impl(cfg, A, levels, b, x, hist, out);
}

int or_solve_bcsr(const or_solver_cfg* cfg, int nb, const int* brp, const int* bci, const double* bv, const double* b,
                  double* x, double* hist, or_solver_out* out) {
    if (cfg->prec != 0 && cfg->prec != 4) return -5;  // identity or the aggregation AMG (p-first)
    if (cfg->prec == 0 && cfg->krylov == 0) return -5;
    Csr A;
    A.n = A.m = nb * cfg->nf;
    A.bs = cfg->nf;
    A.brp = brp; A.bci = bci; A.bv = bv;
    return solve_impl(cfg, A, nullptr, b, x, hist, out);
}

static int solve_impl(const or_solver_cfg* cfg, const Csr& A, const or_level_desc* levels, const double* b, double* x,
                      double* hist, or_solver_out* out) {
    auto t0 = std::chrono::steady_clock::now();
    const int nf = cfg->nf;
    const int n = A.n;
    MG mg;
    mg.cfg = cfg;
    std::vector<double> invD;
    if (cfg->prec == 1 || cfg->prec == 2) {
        if (block_inverse(A, nf, invD)) return -2;
    } else if (cfg->prec == 3) {
        mg.lv.resize(cfg->nlevels);
        for (int l = 0; l < cfg->nlevels; l++) {
            const or_level_desc& d = levels[l];
            mg.lv[l].A.view(d.n, d.n, d.rp, d.ci, d.v);
            mg.lv[l].nf = d.nf > 0 ? d.nf : nf;
            if (l + 1 < cfg->nlevels && d.pmodes > 0) {
                // p-injection of hierarchical bases (P_MultigridForHHO.h:62-86)
                mg.lv[l].ptrunc = true;
                mg.lv[l].nfc = d.pmodes;
                continue;
            }
            if (l + 1 < cfg->nlevels) {
                Level& L = mg.lv[l];
                L.P.view(d.n, d.nc, d.prp, d.pci, d.pv);
                // R = P^T (LevelForHHO.h:535-543)
                L.R.n = d.nc; L.R.m = d.n;
                L.R.rp.assign(d.nc + 1, 0);
                for (int p = 0; p < d.prp[d.n]; p++) L.R.rp[d.pci[p] + 1]++;
                for (int i = 0; i < d.nc; i++) L.R.rp[i + 1] += L.R.rp[i];
                L.R.ci.resize(d.prp[d.n]); L.R.v.resize(d.prp[d.n]);
                std::vector<int> pos(L.R.rp.begin(), L.R.rp.end() - 1);
                for (int i = 0; i < d.n; i++)
                    for (int p = d.prp[i]; p < d.prp[i + 1]; p++) {
                        int q = pos[d.pci[p]]++;
                        L.R.ci[q] = i; L.R.v[q] = d.pv[p];
                    }
                L.R.own();
            }
        }
    } else if (cfg->prec == 4) {
        build_amg(mg, A, nf, cfg->theta, cfg->max_coarse, 32, true);
    }
    if (cfg->prec >= 3) {
        for (size_t l = 0; l + 1 < mg.lv.size(); l++)
            if (block_inverse(mg.lv[l].A, mg.lv[l].nf, mg.lv[l].invD)) return -2;
        if (!mg.coarse.factor(mg.lv.back().A)) return -3;
        out->nlevels = (int)mg.lv.size();
        for (size_t l = 0; l < mg.lv.size() && l < 32; l++) out->level_n[l] = mg.lv[l].A.n;
    }
    // MG::CanOptimizeResidualComputation (Multigrid.h:74-77): the point
    // Gauss-Seidel post-smoother returns A x (GaussSeidel.h:53-56).
    bool fused = cfg->prec >= 3 && mg.lv.size() > 1 && nf == 1 && cfg->post_kind >= SM_BGS_F && !sm_is_cheb(cfg->post_kind) &&
                 cfg->nu2 > 0;
    auto t1 = std::chrono::steady_clock::now();

    auto apply = [&](const double* r, double* z, double* Az) {
        std::fill(z, z + n, 0.0);
        if (cfg->prec == 0) { std::copy(r, r + n, z); return; }
        if (cfg->prec == 1) { bj_sweep(A, nf, invD.data(), cfg->omega, r, z); return; }
        if (cfg->prec == 2) { bgs_sweep(A, nf, invD.data(), 0, r, z); return; }
        mg.cycle(0, r, z, Az);
    };

    double bn = nrm2(n, b);
    auto nres = [&](const double* r) { double rn = nrm2(n, r); return bn > 0 ? rn / bn : rn; };
    int it = 0;
    int nh = 0;
    double res = 1.0;
    int status = 1;
    std::vector<double> r(b, b + n), z(n), d(n), Ad(n), Az(n);
    // first residual: b from x = 0, else b - A x (ConjugateGradient.h:36-45)
    if (cfg->x_given) {
        spmv(A, x, Ad.data());
        for (int i = 0; i < n; i++) r[i] = b[i] - Ad[i];
        res = nres(r.data());
    } else {
        std::fill(x, x + n, 0.0);
    }
    hist[nh++] = res;
    auto stop = [&]() {
        // IterativeSolver::StoppingCriteriaReached (IterativeSolver.h:224-247)
        if (cfg->max_iter == 0) return true;
        if (it == 0) return false;
        if (it >= cfg->max_iter) { status = res < cfg->tol ? 0 : 1; return true; }
        if (!std::isfinite(res)) { status = -4; return true; }
        if (res < cfg->tol) { status = 0; return true; }
        return false;
    };
    if (cfg->krylov == 1) {
        // ConjugateGradient::Solve (ConjugateGradient.h:27-104)
        apply(r.data(), z.data(), nullptr);
        d = z;
        double rz = dot(n, r.data(), z.data()), beta = 0;
        while (!stop()) {
            if (it > 0) for (int i = 0; i < n; i++) d[i] = z[i] + beta * d[i];
            spmv(A, d.data(), Ad.data());
            double alpha = rz / dot(n, d.data(), Ad.data());
            for (int i = 0; i < n; i++) x[i] += alpha * d[i];
            double old = rz;
            for (int i = 0; i < n; i++) r[i] -= alpha * Ad[i];
            apply(r.data(), z.data(), nullptr);
            rz = dot(n, r.data(), z.data());
            beta = rz / old;
            it++;
            res = nres(r.data());
            hist[nh++] = res;
        }
    } else if (cfg->krylov == 2) {
        // FlexibleConjugateGradient::Solve (FlexibleConjugateGradient.h:47-171):
        // A-orthogonalisation of z against the kept directions (oldest
        // first), the oldest dropped beyond `truncation` (all with restart)
        struct Dir { std::vector<double> d, Ad; double dAd; };
        std::vector<Dir> kept;
        const int m = cfg->fcg_truncation > 0 ? cfg->fcg_truncation : 1;
        while (!stop()) {
            apply(r.data(), z.data(), fused ? Az.data() : nullptr);
            d = z;
            if (fused) Ad = Az;
            for (const Dir& k : kept) {
                double zAdk = dot(n, z.data(), k.Ad.data());
                for (int i = 0; i < n; i++) d[i] -= (zAdk / k.dAd) * k.d[i];
                if (fused) for (int i = 0; i < n; i++) Ad[i] -= (zAdk / k.dAd) * k.Ad[i];
            }
            if (!fused) spmv(A, d.data(), Ad.data());
            double dAd = dot(n, d.data(), Ad.data());
            double alpha = dot(n, r.data(), d.data()) / dAd;
            for (int i = 0; i < n; i++) x[i] += alpha * d[i];
            for (int i = 0; i < n; i++) r[i] -= alpha * Ad[i];
            if ((int)kept.size() == m) {
                if (cfg->fcg_restart) kept.clear();
                else kept.erase(kept.begin());
            }
            kept.push_back(Dir{d, Ad, dAd});
            it++;
            res = nres(r.data());
            hist[nh++] = res;
        }
    } else {
        // Stand-alone iterations of the preconditioner (IterativeSolver.h:148-179)
        std::vector<double> xx(n);
        while (!stop()) {
            if (cfg->prec >= 3) mg.cycle(0, b, x, nullptr);
            else if (cfg->prec == 1) bj_sweep(A, nf, invD.data(), cfg->omega, b, x);
            else bgs_sweep(A, nf, invD.data(), 0, b, x);
            spmv(A, x, xx.data());
            for (int i = 0; i < n; i++) r[i] = b[i] - xx[i];
            it++;
            res = nres(r.data());
            hist[nh++] = res;
        }
    }
    auto t2 = std::chrono::steady_clock::now();
    g_setup_s = std::chrono::duration<double>(t1 - t0).count();
    g_solve_s = std::chrono::duration<double>(t2 - t1).count();
    out->iterations = it;
    out->final_res = res;
    return status;
}

}  // extern "C"

// ================================================================ learned post-smoother oracle
namespace {
const int NH = 32;
struct MpShape {
    int nf, L;
    int din() const { return nf + 2; }
    long We() const { return 0; }
    long be() const { return (long)NH * din(); }
    long S(int l) const { return be() + NH + (long)l * (2 * NH * NH + NH); }
    long N(int l) const { return S(l) + NH * NH; }
    long c(int l) const { return N(l) + NH * NH; }
    long Wo() const { return S(L); }
    long bo() const { return Wo() + (long)nf * NH; }
    long total() const { return bo() + nf; }
};

struct MpState {
    int nb, nf, L;
    std::vector<double> w, invD, v, u, x0, h, m;  // h: (L+1) x nb x NH, m: L x nb x NH
    double sigma;
};

void mp_forward(const MpShape& S, int nb, const int* brp, const int* bci, const double* bval, const double* feat,
                const double* P, const double* r1, double* delta, MpState& st) {
    const int nf = S.nf, L = S.L, din = S.din();
    const size_t bs = (size_t)nf * nf;
    st.nb = nb; st.nf = nf; st.L = L;
    // edge weights
    st.w.assign(brp[nb], 0.0);
    for (int i = 0; i < nb; i++) {
        double sum = 0;
        for (int p = brp[i]; p < brp[i + 1]; p++) {
            if (bci[p] == i) continue;
            double tr = 0;
            for (int a = 0; a < nf; a++) tr += bval[p * bs + a * nf + a];
            sum += std::fabs(tr);
        }
        for (int p = brp[i]; p < brp[i + 1]; p++) {
            double tr = 0;
            for (int a = 0; a < nf; a++) tr += bval[p * bs + a * nf + a];
            st.w[p] = (bci[p] == i || sum == 0) ? 0.0 : std::fabs(tr) / sum;
        }
    }
    // block inverses (Gauss-Jordan, partial pivoting)
    st.invD.assign((size_t)nb * bs, 0.0);
    for (int i = 0; i < nb; i++) {
        int pd = -1;
        for (int p = brp[i]; p < brp[i + 1]; p++) if (bci[p] == i) pd = p;
        std::vector<double> a(bval + pd * bs, bval + (pd + 1) * bs), inv(bs, 0.0);
        for (int r = 0; r < nf; r++) inv[r * nf + r] = 1;
        for (int c = 0; c < nf; c++) {
            int piv = c;
            for (int r = c + 1; r < nf; r++) if (std::fabs(a[r * nf + c]) > std::fabs(a[piv * nf + c])) piv = r;
            for (int j = 0; j < nf; j++) { std::swap(a[piv * nf + j], a[c * nf + j]); std::swap(inv[piv * nf + j], inv[c * nf + j]); }
            double d = 1.0 / a[c * nf + c];
            for (int j = 0; j < nf; j++) { a[c * nf + j] *= d; inv[c * nf + j] *= d; }
            for (int r = 0; r < nf; r++) {
                if (r == c) continue;
                double f = a[r * nf + c];
                for (int j = 0; j < nf; j++) { a[r * nf + j] -= f * a[c * nf + j]; inv[r * nf + j] -= f * inv[c * nf + j]; }
            }
        }
        std::copy(inv.begin(), inv.end(), st.invD.begin() + (size_t)i * bs);
    }
    const size_t N = (size_t)nb * nf;
    st.v.assign(N, 0.0);
    double vv = 0;
    for (int i = 0; i < nb; i++)
        for (int a = 0; a < nf; a++) {
            double s = 0;
            for (int b = 0; b < nf; b++) s += st.invD[i * bs + a * nf + b] * r1[i * nf + b];
            st.v[i * nf + a] = s;
            vv += s * s;
        }
    st.sigma = std::sqrt(vv / (double)N);
    const double inv = st.sigma > 0 ? 1.0 / st.sigma : 0.0;
    st.u.assign(N, 0.0);
    st.x0.assign((size_t)nb * din, 0.0);
    for (int i = 0; i < nb; i++) {
        for (int a = 0; a < nf; a++) st.u[i * nf + a] = st.x0[(size_t)i * din + a] = st.v[i * nf + a] * inv;
        st.x0[(size_t)i * din + nf] = feat[2 * i];
        st.x0[(size_t)i * din + nf + 1] = feat[2 * i + 1];
    }
    const size_t nH = (size_t)nb * NH;
    st.h.assign((L + 1) * nH, 0.0);
    st.m.assign(L * nH, 0.0);
    for (int i = 0; i < nb; i++)
        for (int c = 0; c < NH; c++) {
            double acc = P[S.be() + c];
            for (int k = 0; k < din; k++) acc += P[S.We() + c * din + k] * st.x0[(size_t)i * din + k];
            st.h[(size_t)i * NH + c] = acc > 0 ? acc : 0;
        }
    for (int l = 0; l < L; l++) {
        const double* hin = &st.h[l * nH];
        double* hout = &st.h[(l + 1) * nH];
        double* m = &st.m[l * nH];
        for (int i = 0; i < nb; i++) {
            for (int c = 0; c < NH; c++) {
                double s = 0;
                for (int p = brp[i]; p < brp[i + 1]; p++) s += st.w[p] * hin[(size_t)bci[p] * NH + c];
                m[(size_t)i * NH + c] = s;
            }
            for (int c = 0; c < NH; c++) {
                double acc = P[S.c(l) + c];
                for (int k = 0; k < NH; k++)
                    acc += P[S.S(l) + c * NH + k] * hin[(size_t)i * NH + k] + P[S.N(l) + c * NH + k] * m[(size_t)i * NH + k];
                hout[(size_t)i * NH + c] = acc > 0 ? acc : 0;
            }
        }
    }
    const double* hL = &st.h[L * nH];
    for (int i = 0; i < nb; i++)
        for (int a = 0; a < nf; a++) {
            double y = P[S.bo() + a];
            for (int k = 0; k < NH; k++) y += P[S.Wo() + a * NH + k] * hL[(size_t)i * NH + k];
            delta[i * nf + a] = st.sigma * (st.u[i * nf + a] + y);
        }
}

void bspmv(int nb, int nf, const int* brp, const int* bci, const double* bval, const double* x, double* y) {
    const size_t bs = (size_t)nf * nf;
    for (int i = 0; i < nb; i++)
        for (int a = 0; a < nf; a++) {
            double s = 0;
            for (int p = brp[i]; p < brp[i + 1]; p++)
                for (int b = 0; b < nf; b++) s += bval[p * bs + a * nf + b] * x[(size_t)bci[p] * nf + b];
            y[i * nf + a] = s;
        }
}
}  // namespace

extern "C" int or_mpnn_num_params(int nf, int layers) { return (int)MpShape{nf, layers}.total(); }

extern "C" void or_mpnn_forward(int nb, int nf, const int* brp, const int* bci, const double* bval, const double* feat,
                                int layers, const double* params, const double* r1, double* delta) {
    MpState st;
    mp_forward(MpShape{nf, layers}, nb, brp, bci, bval, feat, params, r1, delta, st);
}

extern "C" double or_mpnn_loss_grad(int nb, int nf, const int* brp, const int* bci, const double* bval,
                                    const double* feat, int layers, const double* params, const double* r1,
                                    const double* e1, double* grad) {
    const MpShape S{nf, layers};
    const int L = layers, din = S.din();
    const size_t N = (size_t)nb * nf, nH = (size_t)nb * NH;
    MpState st;
    std::vector<double> delta(N), diff(N), Ad(N);
    mp_forward(S, nb, brp, bci, bval, feat, params, r1, delta.data(), st);
    double E0 = 0, E = 0;
    for (size_t k = 0; k < N; k++) { E0 += e1[k] * r1[k]; diff[k] = e1[k] - delta[k]; }
    bspmv(nb, nf, brp, bci, bval, diff.data(), Ad.data());
    for (size_t k = 0; k < N; k++) E += diff[k] * Ad[k];
    if (!grad) return E / E0;
    std::fill(grad, grad + S.total(), 0.0);
    // d loss / d delta = -2 A diff / E0 ; dy = sigma * that
    std::vector<double> gy(N), dh(nH, 0.0), dprev(nH), t(nH);
    for (size_t k = 0; k < N; k++) gy[k] = st.sigma * (-2.0 * Ad[k] / E0);
    const double* hL = &st.h[L * nH];
    for (int i = 0; i < nb; i++)
        for (int a = 0; a < nf; a++) {
            const double g = gy[i * nf + a];
            grad[S.bo() + a] += g;
            for (int k = 0; k < NH; k++) {
                grad[S.Wo() + a * NH + k] += g * hL[(size_t)i * NH + k];
                dh[(size_t)i * NH + k] += params[S.Wo() + a * NH + k] * g;
            }
        }
    for (int l = L - 1; l >= 0; l--) {
        const double* hout = &st.h[(l + 1) * nH];
        const double* hin = &st.h[l * nH];
        const double* m = &st.m[l * nH];
        for (size_t k = 0; k < nH; k++) if (!(hout[k] > 0)) dh[k] = 0;  // dpre
        std::fill(dprev.begin(), dprev.end(), 0.0);
        for (int i = 0; i < nb; i++)
            for (int c = 0; c < NH; c++) {
                const double d = dh[(size_t)i * NH + c];
                grad[S.c(l) + c] += d;
                for (int k = 0; k < NH; k++) {
                    grad[S.S(l) + c * NH + k] += d * hin[(size_t)i * NH + k];
                    grad[S.N(l) + c * NH + k] += d * m[(size_t)i * NH + k];
                }
            }
        for (int i = 0; i < nb; i++)
            for (int k = 0; k < NH; k++) {
                double a = 0, b = 0;
                for (int c = 0; c < NH; c++) {
                    a += params[S.S(l) + c * NH + k] * dh[(size_t)i * NH + c];
                    b += params[S.N(l) + c * NH + k] * dh[(size_t)i * NH + c];
                }
                dprev[(size_t)i * NH + k] = a;
                t[(size_t)i * NH + k] = b;
            }
        // m_i = sum_p w_p h_{col p}  ->  dh_j += sum_{i, p=(i,j)} w_p t_i
        for (int i = 0; i < nb; i++)
            for (int p = brp[i]; p < brp[i + 1]; p++)
                for (int k = 0; k < NH; k++) dprev[(size_t)bci[p] * NH + k] += st.w[p] * t[(size_t)i * NH + k];
        dh = dprev;
    }
    const double* h0 = &st.h[0];
    for (size_t k = 0; k < nH; k++) if (!(h0[k] > 0)) dh[k] = 0;
    for (int i = 0; i < nb; i++)
        for (int c = 0; c < NH; c++) {
            const double d = dh[(size_t)i * NH + c];
            grad[S.be() + c] += d;
            for (int k = 0; k < din; k++) grad[S.We() + c * din + k] += d * st.x0[(size_t)i * din + k];
        }
    return E / E0;
}
