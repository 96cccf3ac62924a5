// Matrix construction and host-side set-up algorithms.
#include "fh_matrix.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>

#include "fh_kernels.h"

namespace fh {

static void upload(Ctx& c, DevBuf& d, const void* h, size_t bytes) {
    if (c.host_only) return;  // host set-up only: structure and values stay on the host
    d.alloc(bytes);
    if (bytes) FH_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, c.stream));
}

std::unique_ptr<Bcsr> bcsr_from_blocks(Ctx& c, int nbr, int nbc, int bs_r, int bs_c, const int* rowptr,
                                       const int* colidx, const double* vals, bool keep_host) {
    FH_CHECK(nbr >= 0 && nbc >= 0 && bs_r >= 1 && bs_c >= 1, -2, "invalid matrix dimensions");
    auto A = std::make_unique<Bcsr>();
    A->nbr = nbr; A->nbc = nbc; A->bs_r = bs_r; A->bs_c = bs_c;
    A->nnzb = rowptr[nbr];
    A->prec = Prec::F64;
    for (int i = 0; i < nbr; i++)
        FH_CHECK(rowptr[i + 1] >= rowptr[i], -2, "row pointer not monotone");
    upload(c, A->rowptr, rowptr, sizeof(int) * (nbr + 1));
    upload(c, A->colidx, colidx, sizeof(int) * A->nnzb);
    upload(c, A->vals, vals, sizeof(double) * (size_t)A->nnzb * bs_r * bs_c);
    A->h_rowptr.assign(rowptr, rowptr + nbr + 1);
    A->h_colidx.assign(colidx, colidx + A->nnzb);
    if (keep_host || c.host_only) A->h_vals.assign(vals, vals + (size_t)A->nnzb * bs_r * bs_c);
    c.sync();
    return A;
}

std::unique_ptr<Bcsr> bcsr_from_csr(Ctx& c, int n, int m, int bs, const int* rp, const int* ci, const double* v,
                                    bool keep_host) {
    FH_CHECK(bs >= 1 && n % bs == 0 && m % bs == 0, -2, "CSR dimensions are not multiples of the block size");
    int nbr = n / bs, nbc = m / bs;
    std::vector<int> brp(nbr + 1, 0), bci;
    std::vector<int> mark(nbc, -1);
    // block pattern, block columns sorted
    std::vector<int> cols;
    for (int I = 0; I < nbr; I++) {
        cols.clear();
        for (int a = 0; a < bs; a++) {
            int row = I * bs + a;
            for (int p = rp[row]; p < rp[row + 1]; p++) {
                int J = ci[p] / bs;
                FH_CHECK(ci[p] >= 0 && ci[p] < m, -2, "column index out of range");
                if (mark[J] != I) { mark[J] = I; cols.push_back(J); }
            }
        }
        std::sort(cols.begin(), cols.end());
        bci.insert(bci.end(), cols.begin(), cols.end());
        brp[I + 1] = (int)bci.size();
    }
    size_t bsz = (size_t)bs * bs;
    std::vector<double> bv(bci.size() * bsz, 0.0);
    std::vector<uint8_t> mask(keep_host || c.host_only ? bci.size() * bsz : 0, 0);
    std::vector<int> pos(nbc, -1);
    for (int I = 0; I < nbr; I++) {
        for (int q = brp[I]; q < brp[I + 1]; q++) pos[bci[q]] = q;
        for (int a = 0; a < bs; a++) {
            int row = I * bs + a;
            for (int p = rp[row]; p < rp[row + 1]; p++) {
                int J = ci[p] / bs, b = ci[p] % bs;
                size_t e = (size_t)pos[J] * bsz + a * bs + b;
                bv[e] = v[p];
                if (!mask.empty()) mask[e] = 1;
            }
        }
    }
    auto A = bcsr_from_blocks(c, nbr, nbc, bs, bs, brp.data(), bci.data(), bv.data(), keep_host);
    A->h_mask = std::move(mask);
    return A;
}

std::unique_ptr<Bcsr> bcsr_convert(Ctx& c, const Bcsr& A, Prec p) {
    FH_CHECK(!c.host_only, -9, "precision conversion needs a device");
    auto B = std::make_unique<Bcsr>();
    B->nbr = A.nbr; B->nbc = A.nbc; B->bs_r = A.bs_r; B->bs_c = A.bs_c; B->nnzb = A.nnzb; B->prec = p;
    B->rowptr.alloc(A.rowptr.bytes);
    B->colidx.alloc(A.colidx.bytes);
    if (A.rowptr.bytes) FH_HIP(hipMemcpyAsync(B->rowptr.p, A.rowptr.p, A.rowptr.bytes, hipMemcpyDeviceToDevice, c.stream));
    if (A.colidx.bytes) FH_HIP(hipMemcpyAsync(B->colidx.p, A.colidx.p, A.colidx.bytes, hipMemcpyDeviceToDevice, c.stream));
    int64_t nv = A.nnzb * A.bs_r * A.bs_c;
    B->vals.alloc((size_t)nv * prec_bytes(p));
    launch_convert(c, nv, A.vals.p, A.prec, B->vals.p, p);
    B->h_rowptr = A.h_rowptr;
    B->h_colidx = A.h_colidx;
    if (A.elem && A.elem->prec == Prec::F64 && p == Prec::F32) {
        // the element layout follows to FP32 (BF16 copies keep the block rows only)
        B->elem = elem_convert(c, *A.elem, p);
        set_elem(*B, A.elem_on);
    }
    c.sync();
    return B;
}

std::unique_ptr<Bcsr> bcsr_transpose(Ctx& c, const Bcsr& A) {
    FH_CHECK(!A.h_vals.empty(), -3, "transpose needs host values");
    int br = A.bs_r, bc = A.bs_c;
    std::vector<int> rp(A.nbc + 1, 0), ci(A.nnzb);
    for (int64_t p = 0; p < A.nnzb; p++) rp[A.h_colidx[p] + 1]++;
    for (int j = 0; j < A.nbc; j++) rp[j + 1] += rp[j];
    std::vector<int> pos(rp.begin(), rp.end() - 1);
    std::vector<double> v((size_t)A.nnzb * br * bc);
    for (int i = 0; i < A.nbr; i++)
        for (int p = A.h_rowptr[i]; p < A.h_rowptr[i + 1]; p++) {
            int q = pos[A.h_colidx[p]]++;
            ci[q] = i;
            const double* s = &A.h_vals[(size_t)p * br * bc];
            double* d = &v[(size_t)q * br * bc];
            for (int a = 0; a < br; a++)
                for (int b = 0; b < bc; b++) d[b * br + a] = s[a * bc + b];
        }
    return bcsr_from_blocks(c, A.nbc, A.nbr, bc, br, rp.data(), ci.data(), v.data(), true);
}

std::vector<int> diag_blocks(const Bcsr& A) {
    std::vector<int> d(A.nbr, -1);
    for (int i = 0; i < A.nbr; i++) {
        for (int p = A.h_rowptr[i]; p < A.h_rowptr[i + 1]; p++)
            if (A.h_colidx[p] == i) { d[i] = p; break; }
        FH_CHECK(d[i] >= 0, -4, "block row " + std::to_string(i) + " has no diagonal block");
    }
    return d;
}

int greedy_colouring(const Bcsr& A, std::vector<int>& colour) {
    colour.assign(A.nbr, 0);
    int ncol = 0;
    std::vector<int> mark(64, -1);
    for (int i = 0; i < A.nbr; i++) {
        for (int p = A.h_rowptr[i]; p < A.h_rowptr[i + 1]; p++) {
            int j = A.h_colidx[p];
            if (j < i) {
                int cj = colour[j];
                if (cj >= (int)mark.size()) mark.resize(cj + 1, -1);
                mark[cj] = i;
            }
        }
        int cc = 0;
        while (cc < (int)mark.size() && mark[cc] == i) cc++;
        colour[i] = cc;
        ncol = std::max(ncol, cc + 1);
    }
    return ncol;
}

// --------------------------------------------------------------- aggregation
namespace {
}  // namespace

int pairwise_aggregate(const Bcsr& A, double theta, std::vector<int>& agg, std::vector<std::vector<int>>& members,
                       bool& coarsest) {
    FH_CHECK(!A.h_vals.empty(), -3, "aggregation needs host values");
    FH_CHECK(A.bs_r == A.bs_c, -3, "aggregation needs square blocks");
    const int nb = A.nbr, bs = A.bs_r;
    const size_t bsz = (size_t)bs * bs;
    std::vector<AggElem> E(nb);
    std::vector<std::atomic<int>> counts(nb);
    for (auto& x : counts) x.store(0);
    // Neighbours in the order the scalar rows of the block row list them:
    // block j first appears in the first row k with a stored (non-zero)
    // entry in A_ij.  Coupling = trace(A_ij).
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < nb; i++) {
        AggElem& e = E[i];
        const int s = A.h_rowptr[i], t = A.h_rowptr[i + 1];
        std::vector<char> seen(t - s, 0);
        for (int k = 0; k < bs; k++) {
            for (int p = s; p < t; p++) {
                int j = A.h_colidx[p];
                if (j == i || seen[p - s]) continue;
                const double* blk = &A.h_vals[(size_t)p * bsz];
                bool nz = false;
                if (A.h_mask.empty()) {
                    for (int b = 0; b < bs; b++) if (blk[k * bs + b] != 0.0) { nz = true; break; }
                } else {
                    const uint8_t* m = &A.h_mask[(size_t)p * bsz];
                    for (int b = 0; b < bs; b++) if (m[k * bs + b]) { nz = true; break; }
                }
                if (!nz) continue;
                seen[p - s] = 1;
                double tr = blk[0];
                for (int a = 1; a < bs; a++) tr += blk[a * bs + a];
                e.nbr.push_back({j, tr});
            }
        }
        std::sort(e.nbr.begin(), e.nbr.end(),
                  [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.second < b.second; });
        for (auto& q : e.nbr) {
            if (q.second < 0 && q.second < theta * e.nbr[0].second) {
                e.strong.push_back(q.first);
                counts[q.first].fetch_add(1, std::memory_order_relaxed);
            } else
                break;
        }
    }
    for (int i = 0; i < nb; i++) E[i].count = counts[i].load();
    return pairwise_pairing(E, agg, members, coarsest);
}

int pairwise_pairing(std::vector<AggElem>& E, std::vector<int>& agg, std::vector<std::vector<int>>& members,
                     bool& coarsest) {
    const int nb = (int)E.size();
    // Priority ordering: breadth-first from the element with the fewest
    // "I am strong neighbour of" counts, neighbours numbered by ascending count.
    std::vector<int> order(nb, -1);
    std::vector<char> numbered(nb, 0);
    int on = 0;
    if (nb > 0) {
        int start = 0, mn = E[0].count;
        for (int i = 0; i < nb; i++) if (E[i].count < mn) { start = i; mn = E[i].count; }
        order[on++] = start;
        numbered[start] = 1;
    }
    std::vector<int> un;
    int scan = 0;
    for (int i = 0; i < nb; i++) {
        if (i >= on) {  // disconnected graph: restart at the next unnumbered element
            while (numbered[scan]) scan++;
            order[on++] = scan;
            numbered[scan] = 1;
        }
        const AggElem& e = E[order[i]];
        un.clear();
        for (auto& q : e.nbr) if (!numbered[q.first]) un.push_back(q.first);
        if (un.empty()) continue;
        std::sort(un.begin(), un.end(), [&E](int a, int b) { return E[a].count < E[b].count; });
        // an element listed twice (two shared faces) is numbered once
        for (int u : un) if (!numbered[u]) { order[on++] = u; numbered[u] = 1; }
    }
    // Pairing
    int nagg = 0, remaining = nb;
    members.clear();
    members.reserve(nb / 2 + 1);
    auto add = [&E](int el, int a) {
        E[el].agg = a;
        for (int s : E[el].strong) if (E[s].agg < 0) E[s].count--;
    };
    for (int idx = 0; idx < nb; idx++) {
        int el = order[idx];
        if (E[el].agg >= 0) continue;
        int a = nagg++;
        members.emplace_back();
        add(el, a);
        members.back().push_back(el);
        remaining--;
        double mnc = 100000000;
        int chosen = -1;
        for (int s : E[el].strong)
            if (E[s].agg < 0 && E[s].count < mnc) { chosen = s; mnc = E[s].count; }
        if (chosen >= 0) { add(chosen, a); members.back().push_back(chosen); remaining--; }
        if (remaining == 0) break;
    }
    agg.resize(nb);
    for (int i = 0; i < nb; i++) agg[i] = E[i].agg;
    coarsest = nb == 0 || (double)nb / (double)nagg <= 1.01;
    return nagg;
}

std::unique_ptr<Bcsr> galerkin_aggregate(Ctx& c, const Bcsr& A, const std::vector<int>& agg, int nagg) {
    FH_CHECK(!A.h_vals.empty(), -3, "Galerkin operator needs host values");
    const int bs = A.bs_r;
    const size_t bsz = (size_t)bs * bs;
    std::vector<std::vector<int>> mem(nagg);
    for (int i = 0; i < A.nbr; i++) mem[agg[i]].push_back(i);
    std::vector<std::vector<int>> rcols(nagg);
    std::vector<std::vector<double>> rvals(nagg);
    std::vector<std::vector<uint8_t>> rmask(nagg);
#pragma omp parallel
    {
        std::vector<int> slot(nagg, -1);
        std::vector<int> cols;
        std::vector<double> acc;
        std::vector<char> used;
#pragma omp for schedule(dynamic, 256)
        for (int I = 0; I < nagg; I++) {
            cols.clear();
            acc.clear();
            used.clear();
            for (int k : mem[I]) {
                for (int p = A.h_rowptr[k]; p < A.h_rowptr[k + 1]; p++) {
                    int J = agg[A.h_colidx[p]];
                    const double* blk = &A.h_vals[(size_t)p * bsz];
                    int q = slot[J];
                    for (size_t e = 0; e < bsz; e++) {
                        double val = blk[e];
                        if (!(std::fabs(val) > 1e-15)) continue;  // NonZeroCoefficients.h:36-40
                        if (q < 0) {  // a coarse block exists once a contribution survives the drop
                            q = slot[J] = (int)cols.size();
                            cols.push_back(J);
                            acc.resize(acc.size() + bsz, 0.0);
                            used.resize(used.size() + bsz, 0);
                        }
                        double* dst = &acc[(size_t)q * bsz];
                        char* u = &used[(size_t)q * bsz];
                        if (!u[e]) { dst[e] = val; u[e] = 1; }
                        else dst[e] += val;
                    }
                }
            }
            std::vector<int> perm(cols.size());
            std::iota(perm.begin(), perm.end(), 0);
            std::sort(perm.begin(), perm.end(), [&cols](int a, int b) { return cols[a] < cols[b]; });
            rcols[I].resize(cols.size());
            rvals[I].resize(cols.size() * bsz);
            rmask[I].resize(cols.size() * bsz);
            for (size_t t = 0; t < perm.size(); t++) {
                rcols[I][t] = cols[perm[t]];
                std::memcpy(&rvals[I][t * bsz], &acc[(size_t)perm[t] * bsz], bsz * sizeof(double));
                for (size_t e = 0; e < bsz; e++) rmask[I][t * bsz + e] = used[(size_t)perm[t] * bsz + e];
            }
            for (int J : cols) slot[J] = -1;
        }
    }
    std::vector<int> rp(nagg + 1, 0);
    for (int I = 0; I < nagg; I++) rp[I + 1] = rp[I] + (int)rcols[I].size();
    std::vector<int> ci(rp[nagg]);
    std::vector<double> v((size_t)rp[nagg] * bsz);
    std::vector<uint8_t> mask((size_t)rp[nagg] * bsz);
    for (int I = 0; I < nagg; I++) {
        std::copy(rcols[I].begin(), rcols[I].end(), ci.begin() + rp[I]);
        std::copy(rvals[I].begin(), rvals[I].end(), v.begin() + (size_t)rp[I] * bsz);
        std::copy(rmask[I].begin(), rmask[I].end(), mask.begin() + (size_t)rp[I] * bsz);
    }
    auto Ac = bcsr_from_blocks(c, nagg, nagg, bs, bs, rp.data(), ci.data(), v.data(), true);
    Ac->h_mask = std::move(mask);
    return Ac;
}

std::unique_ptr<Bcsr> galerkin_rap(Ctx& c, const Bcsr& A, const Bcsr& P) {
    FH_CHECK(!A.h_vals.empty() && !P.h_vals.empty(), -3, "Galerkin product needs host values");
    FH_CHECK(A.bs_r == A.bs_c && P.bs_r == A.bs_r && P.bs_c == A.bs_r && P.nbr == A.nbr, -2,
             "Galerkin product: incompatible block shapes");
    const int bs = A.bs_r, nc = P.nbc;
    const size_t bsz = (size_t)bs * bs;
    // R = P^T by block rows of the coarse space
    std::vector<int> rrp(nc + 1, 0), rci(P.nnzb), rpos(P.nnzb);
    for (int64_t p = 0; p < P.nnzb; p++) rrp[P.h_colidx[p] + 1]++;
    for (int J = 0; J < nc; J++) rrp[J + 1] += rrp[J];
    {
        std::vector<int> fill(rrp.begin(), rrp.end() - 1);
        for (int i = 0; i < P.nbr; i++)
            for (int p = P.h_rowptr[i]; p < P.h_rowptr[i + 1]; p++) {
                int q = fill[P.h_colidx[p]]++;
                rci[q] = i;
                rpos[q] = p;  // block P_iJ; R_Ji = P_iJ^T
            }
    }
    std::vector<std::vector<int>> ccols(nc);
    std::vector<std::vector<double>> cvals(nc);
#pragma omp parallel
    {
        std::vector<int> slot(nc, -1);
        std::vector<double> T(bsz), U(bsz);
#pragma omp for schedule(dynamic, 64)
        for (int I = 0; I < nc; I++) {
            std::vector<int>& cols = ccols[I];
            std::vector<double>& acc = cvals[I];
            for (int q = rrp[I]; q < rrp[I + 1]; q++) {
                const int i = rci[q];
                const double* Pi = &P.h_vals[(size_t)rpos[q] * bsz];  // P_iI (bs x bs); R_Ii = Pi^T
                for (int pa = A.h_rowptr[i]; pa < A.h_rowptr[i + 1]; pa++) {
                    const int j = A.h_colidx[pa];
                    const double* Aij = &A.h_vals[(size_t)pa * bsz];
                    // T = Pi^T Aij
                    for (int r = 0; r < bs; r++)
                        for (int s = 0; s < bs; s++) {
                            double v = 0;
                            for (int t = 0; t < bs; t++) v += Pi[t * bs + r] * Aij[t * bs + s];
                            T[r * bs + s] = v;
                        }
                    for (int pp = P.h_rowptr[j]; pp < P.h_rowptr[j + 1]; pp++) {
                        const int J = P.h_colidx[pp];
                        const double* Pj = &P.h_vals[(size_t)pp * bsz];
                        int sl = slot[J];
                        if (sl < 0) {
                            sl = slot[J] = (int)cols.size();
                            cols.push_back(J);
                            acc.resize(acc.size() + bsz, 0.0);
                        }
                        double* dst = &acc[(size_t)sl * bsz];
                        for (int r = 0; r < bs; r++)
                            for (int s = 0; s < bs; s++) {
                                double v = 0;
                                for (int t = 0; t < bs; t++) v += T[r * bs + t] * Pj[t * bs + s];
                                dst[r * bs + s] += v;
                            }
                    }
                }
            }
            // sort the row by column
            std::vector<int> perm(cols.size());
            std::iota(perm.begin(), perm.end(), 0);
            std::sort(perm.begin(), perm.end(), [&cols](int a, int b) { return cols[a] < cols[b]; });
            std::vector<int> sc(cols.size());
            std::vector<double> sv(acc.size());
            for (size_t t = 0; t < perm.size(); t++) {
                sc[t] = cols[perm[t]];
                std::memcpy(&sv[t * bsz], &acc[(size_t)perm[t] * bsz], bsz * sizeof(double));
            }
            for (int J : cols) slot[J] = -1;
            cols.swap(sc);
            acc.swap(sv);
        }
    }
    std::vector<int> rp(nc + 1, 0);
    for (int I = 0; I < nc; I++) rp[I + 1] = rp[I] + (int)ccols[I].size();
    std::vector<int> ci(rp[nc]);
    std::vector<double> v((size_t)rp[nc] * bsz);
    for (int I = 0; I < nc; I++) {
        std::copy(ccols[I].begin(), ccols[I].end(), ci.begin() + rp[I]);
        std::copy(cvals[I].begin(), cvals[I].end(), v.begin() + (size_t)rp[I] * bsz);
    }
    // exact symmetry (the product is symmetric up to rounding)
    for (int I = 0; I < nc; I++)
        for (int p = rp[I]; p < rp[I + 1]; p++) {
            int J = ci[p];
            if (J <= I) continue;
            int q = (int)(std::lower_bound(ci.begin() + rp[J], ci.begin() + rp[J + 1], I) - ci.begin());
            if (q >= rp[J + 1] || ci[q] != I) continue;
            for (int r = 0; r < bs; r++)
                for (int s = 0; s < bs; s++) {
                    double m = 0.5 * (v[(size_t)p * bsz + r * bs + s] + v[(size_t)q * bsz + s * bs + r]);
                    v[(size_t)p * bsz + r * bs + s] = m;
                    v[(size_t)q * bsz + s * bs + r] = m;
                }
        }
    return bcsr_from_blocks(c, nc, nc, bs, bs, rp.data(), ci.data(), v.data(), true);
}

std::unique_ptr<Bcsr> bcsr_mode_submatrix(Ctx& c, const Bcsr& A, int nfc) {
    FH_CHECK(nfc >= 1 && nfc <= A.bs_r && A.bs_r == A.bs_c, -2, "invalid number of kept modes");
    const int bs = A.bs_r;
    if (A.h_vals.empty()) {
        // operator on the device only (the GPU assembly): gather on the
        // device, keep a host copy of the (small) coarse values for the
        // set-up steps below it (Galerkin products, aggregation)
        FH_CHECK(!c.host_only && A.prec == Prec::F64 && A.vals.p, -3, "p-coarsening needs FP64 values");
        auto S = std::make_unique<Bcsr>();
        S->nbr = A.nbr; S->nbc = A.nbc; S->bs_r = S->bs_c = nfc; S->nnzb = A.nnzb; S->prec = Prec::F64;
        S->h_rowptr = A.h_rowptr;
        S->h_colidx = A.h_colidx;
        S->rowptr.alloc(sizeof(int) * (A.nbr + 1));
        S->colidx.alloc(sizeof(int) * std::max<int64_t>(1, A.nnzb));
        FH_HIP(hipMemcpyAsync(S->rowptr.p, A.rowptr.p, sizeof(int) * (A.nbr + 1), hipMemcpyDeviceToDevice, c.stream));
        FH_HIP(hipMemcpyAsync(S->colidx.p, A.colidx.p, sizeof(int) * A.nnzb, hipMemcpyDeviceToDevice, c.stream));
        const size_t nv = (size_t)A.nnzb * nfc * nfc;
        S->vals.alloc(sizeof(double) * std::max<size_t>(1, nv));
        launch_mode_submatrix(c, A.nnzb, bs, nfc, A.vals.as<double>(), S->vals.as<double>());
        S->h_vals.resize(nv);
        FH_HIP(hipMemcpyAsync(S->h_vals.data(), S->vals.p, sizeof(double) * nv, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        return S;
    }
    std::vector<double> v((size_t)A.nnzb * nfc * nfc);
    for (int64_t p = 0; p < A.nnzb; p++)
        for (int a = 0; a < nfc; a++)
            for (int b = 0; b < nfc; b++)
                v[(size_t)p * nfc * nfc + a * nfc + b] = A.h_vals[(size_t)p * bs * bs + a * bs + b];
    auto S = bcsr_from_blocks(c, A.nbr, A.nbc, nfc, nfc, A.h_rowptr.data(), A.h_colidx.data(), v.data(), true);
    if (!A.h_mask.empty()) {
        S->h_mask.resize(v.size());
        for (int64_t p = 0; p < A.nnzb; p++)
            for (int a = 0; a < nfc; a++)
                for (int b = 0; b < nfc; b++)
                    S->h_mask[(size_t)p * nfc * nfc + a * nfc + b] = A.h_mask[(size_t)p * bs * bs + a * bs + b];
    }
    return S;
}

std::vector<double> dense_spd_inverse(const Bcsr& A) {
    FH_CHECK(!A.h_vals.empty(), -3, "coarse solver needs host values");
    const int bs = A.bs_r;
    const int n = A.nbr * bs;
    FH_CHECK(n <= 6000, -5, "coarsest level too large for the dense coarse solver (" + std::to_string(n) + ")");
    std::vector<double> L((size_t)n * n, 0.0);
    for (int I = 0; I < A.nbr; I++)
        for (int p = A.h_rowptr[I]; p < A.h_rowptr[I + 1]; p++) {
            int J = A.h_colidx[p];
            for (int a = 0; a < bs; a++)
                for (int b = 0; b < bs; b++)
                    L[(size_t)(I * bs + a) * n + J * bs + b] = A.h_vals[(size_t)p * bs * bs + a * bs + b];
        }
    // Cholesky, lower triangle in place
    for (int j = 0; j < n; j++) {
        double d = L[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        FH_CHECK(d > 0, -6, "coarse matrix is not positive definite");
        d = std::sqrt(d);
        L[(size_t)j * n + j] = d;
#pragma omp parallel for schedule(static)
        for (int i = j + 1; i < n; i++) {
            double s = L[(size_t)i * n + j];
            for (int k = 0; k < j; k++) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = s / d;
        }
    }
    // inverse = L^-T L^-1 : solve for each unit vector (columns in parallel)
    std::vector<double> inv((size_t)n * n, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
    for (int col = 0; col < n; col++) {
        std::vector<double> y(n, 0.0), x(n, 0.0);
        for (int i = col; i < n; i++) {
            double s = (i == col) ? 1.0 : 0.0;
            for (int k = col; k < i; k++) s -= L[(size_t)i * n + k] * y[k];
            y[i] = s / L[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = y[i];
            for (int k = i + 1; k < n; k++) s -= L[(size_t)k * n + i] * x[k];
            x[i] = s / L[(size_t)i * n + i];
        }
        for (int i = 0; i < n; i++) inv[(size_t)i * n + col] = x[i];
    }
    // symmetrise (rounding)
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            double m = 0.5 * (inv[(size_t)i * n + j] + inv[(size_t)j * n + i]);
            inv[(size_t)i * n + j] = inv[(size_t)j * n + i] = m;
        }
    return inv;
}

}  // namespace fh
