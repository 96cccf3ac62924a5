// Uncondensed aggregation AMG set-up (host).  Follows, with the reference's
// default options of -s uamg (main.cpp:1847-1868: chained coarsening
// prolongations, BoundaryAggregatesInteriorAverage face prolongation,
// ReconstructSmoothedTraceOrInject coarsening prolongation, multiple
// pairwise aggregation up to a coarsening factor of 3.8, Galerkin
// operators):
//   HybridAlgebraicMesh::Build / Coarsen / Theta   (HybridAlgebraicMesh.h:134-399)
//   UncondensedLevel::CoarsenMesh / Coarsen         (UncondensedLevel.h:74-185, 287-361)
//   UncondensedLevel::BuildProlongation (case 6)    (UncondensedLevel.h:424-462)
//   BuildQ_T / BuildQ_F / BuildCoarseTraceOnFineRemovedFaces (UncondensedLevel.h:679-823)
//   BlockJacobi::IterationMatrix                    (BlockJacobi.h:130-170)
// Coefficient lists are filled as NonZeroCoefficients fills them: entries
// with |v| <= 1e-15 are dropped (NonZeroCoefficients.h:36-40), duplicates
// summed in insertion order (setFromTriplets).  Sparse products keep every
// structural entry and accumulate in the order of Eigen's conservative
// product, so that the patterns the next pass reads (element faces, face
// elements) are the reference's.
#include "fh_uamg.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>

#include "fh_matrix.h"

namespace fh {

// ------------------------------------------------------------ sparse helpers
Csr csr_copy(int nr, int nc, const int* rp, const int* ci, const double* v) {
    Csr A;
    A.nr = nr; A.nc = nc;
    A.rp.assign(rp, rp + nr + 1);
    const int64_t nnz = rp[nr];
    A.ci.assign(ci, ci + nnz);
    A.v.assign(v, v + nnz);
    for (int i = 0; i < nr; i++)
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            FH_CHECK(A.ci[p] >= 0 && A.ci[p] < nc, -2, "column index out of range");
            FH_CHECK(p == A.rp[i] || A.ci[p] > A.ci[p - 1], -2, "columns must be strictly ascending in every row");
        }
    return A;
}

Csr csr_transpose(const Csr& A) {
    Csr T;
    T.nr = A.nc; T.nc = A.nr;
    T.rp.assign(A.nc + 1, 0);
    for (int c : A.ci) T.rp[c + 1]++;
    for (int i = 0; i < A.nc; i++) T.rp[i + 1] += T.rp[i];
    T.ci.resize(A.ci.size());
    T.v.resize(A.v.size());
    std::vector<int> fill(T.rp.begin(), T.rp.end() - 1);
    for (int i = 0; i < A.nr; i++)
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            const int q = fill[A.ci[p]]++;
            T.ci[q] = i;
            T.v[q] = A.v[p];
        }
    return T;
}

Csr csr_mul(const Csr& A, const Csr& B) {
    FH_CHECK(A.nc == B.nr, -2, "sparse product: inner dimensions differ");
    Csr C;
    C.nr = A.nr; C.nc = B.nc;
    std::vector<std::vector<int>> rc(A.nr);
    std::vector<std::vector<double>> rv(A.nr);
#pragma omp parallel
    {
        std::vector<int> mark(B.nc, -1);
        std::vector<double> acc(B.nc, 0.0);
        std::vector<int> idx;
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < A.nr; i++) {
            idx.clear();
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
                const int k = A.ci[p];
                const double a = A.v[p];
                for (int q = B.rp[k]; q < B.rp[k + 1]; q++) {
                    const int j = B.ci[q];
                    if (mark[j] != i) {
                        mark[j] = i;
                        acc[j] = a * B.v[q];
                        idx.push_back(j);
                    } else
                        acc[j] += a * B.v[q];
                }
            }
            std::sort(idx.begin(), idx.end());
            rc[i] = idx;
            rv[i].resize(idx.size());
            for (size_t t = 0; t < idx.size(); t++) rv[i][t] = acc[idx[t]];
        }
    }
    C.rp.assign(A.nr + 1, 0);
    for (int i = 0; i < A.nr; i++) C.rp[i + 1] = C.rp[i] + (int)rc[i].size();
    C.ci.resize(C.rp[A.nr]);
    C.v.resize(C.rp[A.nr]);
    for (int i = 0; i < A.nr; i++) {
        std::copy(rc[i].begin(), rc[i].end(), C.ci.begin() + C.rp[i]);
        std::copy(rv[i].begin(), rv[i].end(), C.v.begin() + C.rp[i]);
    }
    return C;
}

Csr csr_sym_lower(const Csr& A) {
    FH_CHECK(A.nr == A.nc, -2, "selfadjoint view of a non-square matrix");
    // row i: mirrored entries (j < i taken from row j's... no: from column i
    // of the lower part above the diagonal) then row i's own lower entries;
    // both in ascending column order.
    Csr L;  // strictly lower part transposed = strictly upper part of the full matrix
    std::vector<int> cnt(A.nr + 1, 0);
    for (int i = 0; i < A.nr; i++)
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++)
            if (A.ci[p] < i) cnt[A.ci[p] + 1]++;
    for (int i = 0; i < A.nr; i++) cnt[i + 1] += cnt[i];
    std::vector<int> uc(cnt[A.nr]);
    std::vector<double> uv(cnt[A.nr]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < A.nr; i++)
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++)
            if (A.ci[p] < i) {
                const int q = fill[A.ci[p]]++;
                uc[q] = i;
                uv[q] = A.v[p];
            }
    Csr F;
    F.nr = F.nc = A.nr;
    F.rp.assign(A.nr + 1, 0);
    for (int i = 0; i < A.nr; i++) {
        int own = 0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) own += A.ci[p] <= i;
        F.rp[i + 1] = F.rp[i] + own + (cnt[i + 1] - cnt[i]);
    }
    F.ci.resize(F.rp[A.nr]);
    F.v.resize(F.rp[A.nr]);
    for (int i = 0; i < A.nr; i++) {
        int q = F.rp[i];
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++)
            if (A.ci[p] <= i) { F.ci[q] = A.ci[p]; F.v[q] = A.v[p]; q++; }
        for (int p = cnt[i]; p < cnt[i + 1]; p++) { F.ci[q] = uc[p]; F.v[q] = uv[p]; q++; }
    }
    return F;
}

namespace {

// Coefficient list with NonZeroCoefficients semantics.
struct Coeffs {
    struct T {
        int i, j;
        double v;
    };
    std::vector<T> t;
    void add(int i, int j, double v) {
        if (std::fabs(v) > 1e-15) t.push_back({i, j, v});
    }
    // rows [r0, r0 + n) of M copied (NonZeroCoefficients::CopyRows)
    void copy_rows(int r0, int n, const Csr& M) {
        for (int r = r0; r < r0 + n; r++)
            for (int p = M.rp[r]; p < M.rp[r + 1]; p++) add(r, M.ci[p], M.v[p]);
    }
    // setFromTriplets: duplicates summed in insertion order
    Csr fill(int nr, int nc) const {
        Csr C;
        C.nr = nr; C.nc = nc;
        std::vector<std::vector<std::pair<int, double>>> rows(nr);
        for (const T& x : t) {
            FH_CHECK(x.i >= 0 && x.i < nr && x.j >= 0 && x.j < nc, -10, "coefficient outside the matrix");
            rows[x.i].push_back({x.j, x.v});
        }
        C.rp.assign(nr + 1, 0);
        for (int i = 0; i < nr; i++) {
            auto& r = rows[i];
            std::stable_sort(r.begin(), r.end(),
                             [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
            for (size_t k = 0; k < r.size(); k++) {
                if (k > 0 && r[k].first == C.ci.back()) {
                    C.v.back() += r[k].second;
                } else {
                    C.ci.push_back(r[k].first);
                    C.v.push_back(r[k].second);
                }
            }
            C.rp[i + 1] = (int)C.ci.size();
        }
        return C;
    }
};

// Entry (i, j) of A (0 when not stored).
double at(const Csr& A, int i, int j) {
    const int* b = A.ci.data() + A.rp[i];
    const int* e = A.ci.data() + A.rp[i + 1];
    const int* p = std::lower_bound(b, e, j);
    return (p != e && *p == j) ? A.v[p - A.ci.data()] : 0.0;
}

// Dense block (row-major) A(r0 : r0 + nr, c0 : c0 + nc).
std::vector<double> block(const Csr& A, int r0, int c0, int nr, int nc) {
    std::vector<double> B((size_t)nr * nc, 0.0);
    for (int i = 0; i < nr; i++) {
        const int r = r0 + i;
        const int* b = A.ci.data() + A.rp[r];
        const int* e = A.ci.data() + A.rp[r + 1];
        for (const int* p = std::lower_bound(b, e, c0); p != e && *p < c0 + nc; ++p)
            B[(size_t)i * nc + (*p - c0)] = A.v[p - A.ci.data()];
    }
    return B;
}

// Cholesky factor (lower, in place, row-major n x n) as Eigen's unblocked
// LLT: diagonal sqrt(a_kk - |L_k,0:k|^2), column below divided by it.
bool llt(std::vector<double>& M, int n) {
    for (int k = 0; k < n; k++) {
        double x = M[(size_t)k * n + k];
        if (k > 0) {
            double s = 0;
            for (int j = 0; j < k; j++) s += M[(size_t)k * n + j] * M[(size_t)k * n + j];
            x -= s;
        }
        if (!(x > 0)) return false;
        x = std::sqrt(x);
        M[(size_t)k * n + k] = x;
        for (int i = k + 1; i < n; i++) {
            double a = M[(size_t)i * n + k];
            for (int j = 0; j < k; j++) a -= M[(size_t)i * n + j] * M[(size_t)k * n + j];
            M[(size_t)i * n + k] = a / x;
        }
    }
    return true;
}

// X = (L L^T)^-1 B for B n x m (row-major), forward then backward
// substitution.
void llt_solve(const std::vector<double>& L, int n, std::vector<double>& B, int m) {
    for (int c = 0; c < m; c++) {
        for (int i = 0; i < n; i++) {
            double t = B[(size_t)i * m + c];
            for (int j = 0; j < i; j++) t -= L[(size_t)i * n + j] * B[(size_t)j * m + c];
            B[(size_t)i * m + c] = t / L[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double t = B[(size_t)i * m + c];
            for (int j = n - 1; j > i; j--) t -= L[(size_t)j * n + i] * B[(size_t)j * m + c];
            B[(size_t)i * m + c] = t / L[(size_t)i * n + i];
        }
    }
}

// Full-pivoting LU of a small n x n block (FullPivLU) and its solve.
struct FullLU {
    int n = 0;
    std::vector<double> lu;
    std::vector<int> rowp, colq;  // row / column permutations
    int rank = 0;
    void compute(const std::vector<double>& A, int n_) {
        n = n_;
        lu = A;
        rowp.resize(n);
        colq.resize(n);
        for (int i = 0; i < n; i++) rowp[i] = colq[i] = i;
        rank = n;
        for (int k = 0; k < n; k++) {
            // largest |a| of the trailing block (column-major scan, first hit)
            int pi = k, pj = k;
            double best = -1;
            for (int j = k; j < n; j++)
                for (int i = k; i < n; i++) {
                    const double a = std::fabs(lu[(size_t)i * n + j]);
                    if (a > best) { best = a; pi = i; pj = j; }
                }
            if (best == 0) { rank = k; break; }
            if (pi != k) { for (int j = 0; j < n; j++) std::swap(lu[(size_t)k * n + j], lu[(size_t)pi * n + j]); std::swap(rowp[k], rowp[pi]); }
            if (pj != k) { for (int i = 0; i < n; i++) std::swap(lu[(size_t)i * n + k], lu[(size_t)i * n + pj]); std::swap(colq[k], colq[pj]); }
            const double d = lu[(size_t)k * n + k];
            for (int i = k + 1; i < n; i++) lu[(size_t)i * n + k] /= d;
            for (int i = k + 1; i < n; i++)
                for (int j = k + 1; j < n; j++) lu[(size_t)i * n + j] -= lu[(size_t)i * n + k] * lu[(size_t)k * n + j];
        }
    }
    // X = A^-1 B, B n x m row-major
    std::vector<double> solve(const std::vector<double>& B, int m) const {
        FH_CHECK(rank == n, -6, "singular diagonal block in the block-Jacobi iteration matrix");
        std::vector<double> c((size_t)n * m), x((size_t)n * m);
        for (int i = 0; i < n; i++)
            for (int t = 0; t < m; t++) c[(size_t)i * m + t] = B[(size_t)rowp[i] * m + t];
        for (int t = 0; t < m; t++) {
            for (int i = 0; i < n; i++)
                for (int j = 0; j < i; j++) c[(size_t)i * m + t] -= lu[(size_t)i * n + j] * c[(size_t)j * m + t];
            for (int i = n - 1; i >= 0; i--) {
                double s = c[(size_t)i * m + t];
                for (int j = i + 1; j < n; j++) s -= lu[(size_t)i * n + j] * c[(size_t)j * m + t];
                c[(size_t)i * m + t] = s / lu[(size_t)i * n + i];
            }
        }
        for (int i = 0; i < n; i++)
            for (int t = 0; t < m; t++) x[(size_t)colq[i] * m + t] = c[(size_t)i * m + t];
        return x;
    }
};

// ------------------------------------------------------ hybrid algebraic mesh
struct HMesh {
    const Csr* ATT = nullptr;
    const Csr* ATF = nullptr;
    int cbs = 1, fbs = 1, ne = 0, nf = 0;
    std::vector<std::vector<int>> efaces;  // Elements[e].Faces
    std::vector<std::vector<int>> felems;  // Faces[f].Elements
    std::vector<AggElem> E;
    // coarsening
    std::vector<int> agg;
    std::vector<std::vector<int>> members;     // CoarseElements[c].FineElements
    std::vector<char> removed;                 // Faces[f].IsRemovedOnCoarseMesh
    std::vector<std::vector<int>> fce;         // Faces[f].CoarseElements
    std::vector<int> fcf;                      // Faces[f].CoarseFace (-1: none)
    std::vector<std::vector<int>> ce_fine;     // CoarseElements[c].FineFaces
    std::vector<std::vector<int>> ce_cfaces;   // CoarseElements[c].CoarseFaces
    int ncf = 0;

    // trace of the block A_TF(e, f) (DenseMatrix::trace of a cbs x fbs block)
    double trace_tf(int e, int f) const {
        double t = 0;
        const int m = std::min(cbs, fbs);
        for (int i = 0; i < m; i++) t += at(*ATF, e * cbs + i, f * fbs + i);
        return t;
    }
    // HybridAlgebraicMesh::CouplingValue(e1, e2, f) (HybridAlgebraicMesh.h:407-422)
    double coupling(int e1, int e2, int f) const {
        const double c = trace_tf(e1, f);
        const double k1 = std::fabs(c), k2 = std::fabs(trace_tf(e2, f));
        return c * (std::min(k1, k2) / std::max(k1, k2));
    }

    // HybridAlgebraicMesh::Build (HybridAlgebraicMesh.h:134-237)
    void build(double theta, bool with_neighbours) {
        ne = ATF->nr / cbs;
        nf = ATF->nc / fbs;
        efaces.assign(ne, {});
        felems.assign(nf, {});
#pragma omp parallel for schedule(dynamic, 256)
        for (int e = 0; e < ne; e++) {
            auto& fl = efaces[e];
            for (int k = 0; k < cbs; k++) {
                const int r = e * cbs + k;
                for (int p = ATF->rp[r]; p < ATF->rp[r + 1]; p++) {
                    const int f = ATF->ci[p] / fbs;
                    if (std::find(fl.begin(), fl.end(), f) == fl.end()) fl.push_back(f);
                }
            }
            FH_CHECK(!fl.empty(), -6, "element " + std::to_string(e) + " has no face in A_TF");
        }
        // faces' elements: column f*fbs of A_TF, rows ascending
        const Csr T = csr_transpose(*ATF);
#pragma omp parallel for schedule(dynamic, 256)
        for (int f = 0; f < nf; f++) {
            auto& el = felems[f];
            const int r = f * fbs;
            for (int p = T.rp[r]; p < T.rp[r + 1]; p++) {
                const int e = T.ci[p] / cbs;
                if (std::find(el.begin(), el.end(), e) == el.end()) el.push_back(e);
            }
        }
        if (!with_neighbours) return;
        E.assign(ne, AggElem{});
        std::vector<std::atomic<int>> counts(ne);
        for (auto& x : counts) x.store(0);
#pragma omp parallel for schedule(dynamic, 256)
        for (int e = 0; e < ne; e++) {
            AggElem& a = E[e];
            for (int f : efaces[e])
                for (int n : felems[f])
                    if (n != e) a.nbr.push_back({n, coupling(e, n, f)});
            std::sort(a.nbr.begin(), a.nbr.end(),
                      [](const std::pair<int, double>& x, const std::pair<int, double>& y) { return x.second < y.second; });
            for (auto& q : a.nbr) {
                if (q.second < 0 && q.second < theta * a.nbr[0].second) {
                    a.strong.push_back(q.first);
                    counts[q.first].fetch_add(1, std::memory_order_relaxed);
                } else
                    break;
            }
        }
        for (int e = 0; e < ne; e++) E[e].count = counts[e].load();
    }

    // HybridAlgebraicMesh::Coarsen with pairwise aggregation and interface
    // collapsing (HybridAlgebraicMesh.h:239-378).  Returns true when the
    // coarsest possible mesh is reached.
    bool coarsen() {
        bool coarsest = false;
        pairwise_pairing(E, agg, members, coarsest);
        if (coarsest) return true;
        const int nce = (int)members.size();
        removed.assign(nf, 0);
        fce.assign(nf, {});
        fcf.assign(nf, -1);
        ce_fine.assign(nce, {});
        ce_cfaces.assign(nce, {});
        std::vector<std::vector<int>> ce_removed(nce);
        for (int c = 0; c < nce; c++) {
            const auto& fe = members[c];
            for (size_t i = 0; i < fe.size(); i++) {
                for (int f : efaces[fe[i]]) {
                    for (size_t j = i + 1; j < fe.size(); j++) {
                        const auto& f2 = efaces[fe[j]];
                        if (std::find(f2.begin(), f2.end(), f) != f2.end()) {
                            removed[f] = 1;
                            ce_removed[c].push_back(f);
                            fce[f].push_back(c);
                            break;
                        }
                    }
                    if (!removed[f]) {
                        fce[f].push_back(c);
                        ce_fine[c].push_back(f);
                    }
                }
            }
        }
        // neighbours of the aggregates, by aggregate number, with the faces
        // they share in the order of the aggregate's faces
        std::vector<std::map<int, std::vector<int>>> nb(nce);
        for (int c = 0; c < nce; c++)
            for (int f : ce_fine[c])
                for (int n : fce[f])
                    if (n != c) nb[c][n].push_back(f);
        // interface collapsing
        ncf = 0;
        for (int c = 0; c < nce; c++)
            for (auto& kv : nb[c]) {
                const auto& ff = kv.second;
                if (fcf[ff.front()] >= 0) continue;
                const int id = ncf++;
                for (int f : ff) fcf[f] = id;
                ce_cfaces[c].push_back(id);
                ce_cfaces[kv.first].push_back(id);
            }
        return ncf == 0;
    }
};

// - A_TT^-1 A_TF block by block (HybridAlgebraicMesh::Theta, :381-399)
Csr theta_matrix(const HMesh& m) {
    const int cbs = m.cbs, fbs = m.fbs;
    std::vector<Coeffs> parts(m.ne);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < m.ne; i++) {
        std::vector<double> L = block(*m.ATT, i * cbs, i * cbs, cbs, cbs);
        FH_CHECK(llt(L, cbs), -6, "cell block of A_TT is not positive definite");
        for (int f : m.efaces[i]) {
            std::vector<double> B = block(*m.ATF, i * cbs, f * fbs, cbs, fbs);
            llt_solve(L, cbs, B, fbs);
            for (int a = 0; a < cbs; a++)
                for (int b = 0; b < fbs; b++) parts[i].add(i * cbs + a, f * fbs + b, -B[(size_t)a * fbs + b]);
        }
    }
    Coeffs all;
    for (auto& p : parts) all.t.insert(all.t.end(), p.t.begin(), p.t.end());
    return all.fill(m.ATF->nr, m.ATF->nc);
}

// J = I - omega D^-1 A with block diagonal D (BlockJacobi::IterationMatrix)
Csr jacobi_iteration_matrix(const Csr& A, int bs, double omega) {
    const int nb = A.nr / bs;
    std::vector<Coeffs> parts(nb);
#pragma omp parallel for schedule(dynamic, 256)
    for (int I = 0; I < nb; I++) {
        FullLU lu;
        lu.compute(block(A, I * bs, I * bs, bs, bs), bs);
        std::vector<int> done;
        for (int k = 0; k < bs; k++) {
            const int r = I * bs + k;
            for (int p = A.rp[r]; p < A.rp[r + 1]; p++) {
                const int J = A.ci[p] / bs;
                if (std::find(done.begin(), done.end(), J) != done.end()) continue;
                if (J == I) {
                    for (int a = 0; a < bs; a++)
                        for (int b = 0; b < bs; b++) parts[I].add(I * bs + a, I * bs + b, a == b ? 1.0 - omega : 0.0);
                } else {
                    const std::vector<double> X = lu.solve(block(A, I * bs, J * bs, bs, bs), bs);
                    for (int a = 0; a < bs; a++)
                        for (int b = 0; b < bs; b++) parts[I].add(I * bs + a, J * bs + b, -omega * X[(size_t)a * bs + b]);
                }
                done.push_back(J);
            }
        }
    }
    Coeffs all;
    for (auto& p : parts) all.t.insert(all.t.end(), p.t.begin(), p.t.end());
    return all.fill(A.nr, A.nc);
}

// One coarsening pass (UncondensedLevel::Coarsen, :287-361) with the
// default prolongations.
struct PassOut {
    Csr ATTc, ATFc, AFFc, Sc, P;
    bool coarsest = false;
};

PassOut coarsen_pass(HMesh& m, const Csr& AFF, const Csr& S, double theta, UamgPass& info) {
    PassOut o;
    const int cbs = m.cbs, fbs = m.fbs;
    m.build(theta, true);
    info.n_elems = m.ne;
    info.n_faces = m.nf;
    if (m.coarsen()) {
        o.coarsest = true;
        return o;
    }
    const int nce = (int)m.members.size();
    info.n_coarse_elems = nce;
    info.n_coarse_faces = m.ncf;
    info.elem_agg = m.agg;
    info.face_coarse = m.fcf;
    // Q_T: one identity block per element (BuildQ_T, :679-691)
    Coeffs qt;
    for (int e = 0; e < m.ne; e++)
        for (int a = 0; a < cbs; a++) qt.add(e * cbs + a, m.agg[e] * cbs + a, 1.0);
    const Csr QT = qt.fill(m.ne * cbs, nce * cbs);
    // Q_F: identity to the coarse face, or the average of the aggregate's
    // coarse faces on removed faces (BuildQ_F, :694-750)
    Coeffs qf;
    for (int f = 0; f < m.nf; f++) {
        if (m.removed[f] && m.fcf[f] < 0) {
            const int c = m.agg[m.felems[f][0]];
            const auto& cfs = m.ce_cfaces[c];
            for (int cf : cfs)
                for (int a = 0; a < fbs; a++)
                    for (int b = 0; b < fbs; b++)
                        qf.add(f * fbs + a, cf * fbs + b, a == b ? 1.0 / (double)cfs.size() : 0.0);
        } else {
            FH_CHECK(m.fcf[f] >= 0, -6, "face " + std::to_string(f) + " has no coarse face (boundary face of an aggregate)");
            for (int a = 0; a < fbs; a++) qf.add(f * fbs + a, m.fcf[f] * fbs + a, 1.0);
        }
    }
    const Csr QF = qf.fill(m.nf * fbs, m.ncf * fbs);
    const Csr QTt = csr_transpose(QT);
    // intermediate coarse operators
    o.ATTc = csr_mul(csr_mul(QTt, csr_sym_lower(*m.ATT)), QT);
    const Csr QTtATF = csr_mul(QTt, *m.ATF);
    const Csr ATFc_tmp = csr_mul(QTtATF, QF);
    HMesh aux;
    aux.ATT = &o.ATTc;
    aux.ATF = &ATFc_tmp;
    aux.cbs = cbs;
    aux.fbs = fbs;
    aux.build(theta, false);
    // ReconstructSmoothedTraceOrInject (:424-462)
    const Csr Theta = theta_matrix(aux);
    Coeffs pi;
    for (int f = 0; f < m.nf; f++)
        if (m.removed[f]) pi.add(f * fbs, m.fce[f][0] * cbs, 1.0);  // trace of the constant
    const Csr Pi = pi.fill(m.nf * fbs, nce * cbs);
    const Csr RT = csr_mul(Pi, Theta);
    Coeffs rti;
    for (int f = 0; f < m.nf; f++) rti.copy_rows(f * fbs, fbs, m.removed[f] ? RT : QF);
    const Csr RTI = rti.fill(QF.nr, QF.nc);
    const Csr J = jacobi_iteration_matrix(S, fbs, 2.0 / 3.0);
    const Csr RST = csr_mul(J, RTI);
    Coeffs pp;
    for (int f = 0; f < m.nf; f++) pp.copy_rows(f * fbs, fbs, m.removed[f] ? RST : QF);
    o.P = pp.fill(QF.nr, QF.nc);
    // coarse operators
    o.ATFc = csr_mul(QTtATF, o.P);
    const Csr Pt = csr_transpose(o.P);
    o.AFFc = csr_mul(csr_mul(Pt, csr_sym_lower(AFF)), o.P);
    o.Sc = csr_mul(csr_mul(Pt, csr_sym_lower(S)), o.P);
    return o;
}

}  // namespace

UamgHierarchy uamg_build(const Csr& A, const Csr& ATT, const Csr& ATF, const Csr& AFF, const UamgArgs& args) {
    const int cbs = args.cell_bs, fbs = args.face_bs;
    FH_CHECK(cbs > 0 && fbs > 0, -2, "block sizes must be positive");
    FH_CHECK(ATT.nr == ATT.nc && ATT.nr == ATF.nr && ATT.nr % cbs == 0, -2, "A_TT / A_TF row counts");
    FH_CHECK(AFF.nr == AFF.nc && AFF.nr == ATF.nc && ATF.nc % fbs == 0, -2, "A_FF / A_TF column counts");
    FH_CHECK(A.nr == A.nc && A.nr == AFF.nr, -2, "condensed operator size differs from A_FF");
    UamgHierarchy H;
    H.A.push_back(A);
    // current level's blocks (level 0: the inputs)
    Csr curTT = ATT, curTF = ATF, curFF = AFF;
    int level = 0;
    while (true) {
        const bool need = (args.max_levels <= 0 && curFF.nr > args.max_coarse) ||
                          (args.max_levels > 0 && level < args.max_levels - 1);
        if (!need) break;
        // UncondensedLevel::CoarsenMesh (:74-185), chained prolongations
        const Csr& S = H.A.back();
        Csr meshTT = curTT, meshTF = curTF, meshFF = curFF, meshS = S;
        Csr Pchain;
        double factor = 0;
        int pass = 0;
        bool coarsest = false;
        while (!(factor >= args.coarsening_factor)) {
            HMesh m;
            m.ATT = &meshTT;
            m.ATF = &meshTF;
            m.cbs = cbs;
            m.fbs = fbs;
            UamgPass info;
            info.level = level;
            info.pass = pass;
            PassOut o = coarsen_pass(m, meshFF, meshS, args.theta, info);
            H.passes.push_back(std::move(info));
            if (o.coarsest) {
                coarsest = true;
                break;
            }
            Pchain = pass == 0 ? o.P : csr_mul(Pchain, o.P);
            factor = (double)curTF.nc / (double)o.ATFc.nc;
            meshTT = std::move(o.ATTc);
            meshTF = std::move(o.ATFc);
            meshFF = std::move(o.AFFc);
            meshS = std::move(o.Sc);
            pass++;
        }
        if (coarsest) break;
        H.P.push_back(std::move(Pchain));
        H.A.push_back(std::move(meshS));
        curTT = std::move(meshTT);
        curTF = std::move(meshTF);
        curFF = std::move(meshFF);
        level++;
    }
    return H;
}

}  // namespace fh
