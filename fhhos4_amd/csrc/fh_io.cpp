// Matrix Market reader / writer of the exported linear system.
#include "fh_io.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

namespace fh {

namespace {

struct File {
    FILE* f = nullptr;
    File(const std::string& path, const char* mode) {
        f = std::fopen(path.c_str(), mode);
        FH_CHECK(f != nullptr, -11, "cannot open " + path);
    }
    ~File() {
        if (f) std::fclose(f);
    }
};

// Values as std::ostream prints them under std::scientific, precision 64.
inline int put_value(char* buf, size_t cap, double v) { return std::snprintf(buf, cap, "%.64e", v); }

// One text line, without the end-of-line; false at end of file.
bool get_line(FILE* f, std::string& line) {
    line.clear();
    char buf[4096];
    while (std::fgets(buf, sizeof buf, f)) {
        line += buf;
        if (!line.empty() && line.back() == '\n') {
            line.pop_back();
            if (!line.empty() && line.back() == '\r') line.pop_back();
            return true;
        }
    }
    return !line.empty();
}

}  // namespace

void market_write_matrix(Ctx& c, const Bcsr& A, const std::string& path) {
    FH_CHECK(A.bs_r == A.bs_c, -2, "Matrix Market export needs square blocks");
    const int bs = A.bs_r;
    const size_t bsz = (size_t)bs * bs;
    std::vector<double> dl;
    const double* V = A.h_vals.empty() ? nullptr : A.h_vals.data();
    if (!V) {
        FH_CHECK(!c.host_only && A.prec == Prec::F64 && A.vals.p, -2,
                 "Matrix Market export needs FP64 values (host copy or FP64 device matrix)");
        dl.resize((size_t)A.nnzb * bsz);
        FH_HIP(hipMemcpyAsync(dl.data(), A.vals.p, dl.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        c.sync();
        V = dl.data();
    }
    const uint8_t* M = A.h_mask.empty() ? nullptr : A.h_mask.data();
    auto stored = [&](size_t e) { return M ? M[e] != 0 : V[e] != 0.0; };
    // blocks of every row in ascending column order
    std::vector<int> order(A.nnzb);
    std::iota(order.begin(), order.end(), 0);
    for (int I = 0; I < A.nbr; I++) {
        const int s = A.h_rowptr[I], e = A.h_rowptr[I + 1];
        std::sort(order.begin() + s, order.begin() + e,
                  [&](int p, int q) { return A.h_colidx[p] < A.h_colidx[q]; });
    }
    int64_t nnz = 0;
    for (size_t e = 0; e < (size_t)A.nnzb * bsz; e++) nnz += stored(e);

    File out(path, "wb");
    std::fprintf(out.f, "%%%%MatrixMarket matrix coordinate  real general\n%lld %lld %lld\n", (long long)A.rows(),
                 (long long)A.cols(), (long long)nnz);
    // rows are formatted in parallel chunks and written in order
    const int chunk = 256;
    const int nchunks = (A.nbr + chunk - 1) / chunk;
    const int batch = 64;
    std::vector<std::string> text(batch);
    for (int b0 = 0; b0 < nchunks; b0 += batch) {
        const int nb = std::min(batch, nchunks - b0);
#pragma omp parallel for schedule(dynamic)
        for (int t = 0; t < nb; t++) {
            std::string& s = text[t];
            s.clear();
            char buf[160];
            const int I0 = (b0 + t) * chunk, I1 = std::min(A.nbr, I0 + chunk);
            for (int I = I0; I < I1; I++)
                for (int i = 0; i < bs; i++)
                    for (int q = A.h_rowptr[I]; q < A.h_rowptr[I + 1]; q++) {
                        const int p = order[q];
                        for (int j = 0; j < bs; j++) {
                            const size_t e = (size_t)p * bsz + (size_t)i * bs + j;
                            if (!stored(e)) continue;
                            int n = std::snprintf(buf, sizeof buf, "%lld %lld ", (long long)I * bs + i + 1,
                                                  (long long)A.h_colidx[p] * bs + j + 1);
                            n += put_value(buf + n, sizeof buf - n, V[e]);
                            buf[n++] = '\n';
                            s.append(buf, n);
                        }
                    }
        }
        for (int t = 0; t < nb; t++)
            FH_CHECK(std::fwrite(text[t].data(), 1, text[t].size(), out.f) == text[t].size(), -11,
                     "write error on " + path);
    }
}

void market_read_csr(const std::string& path, int& nrows, int& ncols, std::vector<int>& rowptr,
                     std::vector<int>& colidx, std::vector<double>& vals) {
    File in(path, "rb");
    std::string line;
    FH_CHECK(get_line(in.f, line), -11, path + ": empty file");
    FH_CHECK(line.rfind("%%MatrixMarket matrix coordinate", 0) == 0 && line.find("real") != std::string::npos &&
                 line.find("general") != std::string::npos,
             -11, path + ": not a 'coordinate real general' Matrix Market file");
    long long m = -1, n = -1, nnz = -1;
    while (get_line(in.f, line)) {
        if (line.empty() || line[0] == '%') continue;
        FH_CHECK(std::sscanf(line.c_str(), "%lld %lld %lld", &m, &n, &nnz) == 3 && m > 0 && n > 0 && nnz >= 0, -11,
                 path + ": bad size line");
        break;
    }
    FH_CHECK(m > 0 && m < (1ll << 31) && n < (1ll << 31), -11, path + ": missing or too large size line");
    std::vector<int> ri, ci;
    std::vector<double> v;
    ri.reserve(nnz); ci.reserve(nnz); v.reserve(nnz);
    while (get_line(in.f, line)) {
        if (line.empty() || line[0] == '%') continue;
        char* p = &line[0];
        char* q;
        long long i = std::strtoll(p, &q, 10);
        FH_CHECK(q != p, -11, path + ": bad entry line");
        p = q;
        long long j = std::strtoll(p, &q, 10);
        FH_CHECK(q != p, -11, path + ": bad entry line");
        p = q;
        double x = std::strtod(p, &q);
        FH_CHECK(q != p, -11, path + ": bad entry line");
        FH_CHECK(i >= 1 && i <= m && j >= 1 && j <= n, -11, path + ": entry index out of range");
        ri.push_back((int)(i - 1)); ci.push_back((int)(j - 1)); v.push_back(x);
    }
    FH_CHECK((long long)v.size() == nnz, -11, path + ": entry count differs from the size line");
    nrows = (int)m; ncols = (int)n;
    // counting sort by row, stable sort by column, duplicates summed
    std::vector<int64_t> cnt(nrows + 1, 0);
    for (int r : ri) cnt[r + 1]++;
    for (int r = 0; r < nrows; r++) cnt[r + 1] += cnt[r];
    std::vector<int64_t> perm(v.size()), fill(cnt.begin(), cnt.end() - 1);
    for (size_t e = 0; e < v.size(); e++) perm[fill[ri[e]]++] = (int64_t)e;
    rowptr.assign(nrows + 1, 0);
    colidx.clear(); vals.clear();
    colidx.reserve(v.size()); vals.reserve(v.size());
    for (int r = 0; r < nrows; r++) {
        std::stable_sort(perm.begin() + cnt[r], perm.begin() + cnt[r + 1],
                         [&](int64_t a, int64_t b) { return ci[a] < ci[b]; });
        for (int64_t k = cnt[r]; k < cnt[r + 1]; k++) {
            const int64_t e = perm[k];
            if ((int64_t)colidx.size() > rowptr[r] && colidx.back() == ci[e])
                vals.back() += v[e];
            else {
                colidx.push_back(ci[e]);
                vals.push_back(v[e]);
            }
        }
        rowptr[r + 1] = (int)colidx.size();
    }
}

void market_write_csr(int nrows, int ncols, const int* rp, const int* ci, const double* v, const std::string& path) {
    File out(path, "wb");
    std::fprintf(out.f, "%%%%MatrixMarket matrix coordinate  real general\n%d %d %d\n", nrows, ncols, rp[nrows]);
    char buf[160];
    for (int i = 0; i < nrows; i++)
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            int k = std::snprintf(buf, sizeof buf, "%d %d ", i + 1, ci[p] + 1);
            k += put_value(buf + k, sizeof buf - k, v[p]);
            buf[k++] = '\n';
            FH_CHECK(std::fwrite(buf, 1, k, out.f) == (size_t)k, -11, "write error on " + path);
        }
}

void market_write_vector(int64_t n, const double* v, const std::string& path) {
    File out(path, "wb");
    std::fprintf(out.f, "%%%%MatrixMarket matrix array real general\n%lld 1\n", (long long)n);
    char buf[128];
    for (int64_t i = 0; i < n; i++) {
        int k = put_value(buf, sizeof buf, v[i]);
        buf[k++] = '\n';
        FH_CHECK(std::fwrite(buf, 1, k, out.f) == (size_t)k, -11, "write error on " + path);
    }
}

std::vector<double> market_read_vector(const std::string& path) {
    File in(path, "rb");
    std::string line;
    FH_CHECK(get_line(in.f, line), -11, path + ": empty file");
    FH_CHECK(line.rfind("%%MatrixMarket matrix array real", 0) == 0, -11,
             path + ": not an 'array real' Matrix Market file");
    long long n = -1, cols = -1;
    while (get_line(in.f, line)) {
        if (line.empty() || line[0] == '%') continue;
        FH_CHECK(std::sscanf(line.c_str(), "%lld %lld", &n, &cols) == 2 && n > 0 && cols == 1, -11,
                 path + ": bad size line (expected 'n 1')");
        break;
    }
    FH_CHECK(n > 0, -11, path + ": missing size line");
    std::vector<double> v;
    v.reserve(n);
    while ((long long)v.size() < n && get_line(in.f, line)) {
        char* q;
        double x = std::strtod(line.c_str(), &q);
        FH_CHECK(q != line.c_str(), -11, path + ": bad value line");
        v.push_back(x);
    }
    FH_CHECK((long long)v.size() == n, -11, path + ": fewer values than the size line states");
    return v;
}

}  // namespace fh
