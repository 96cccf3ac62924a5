// Shared definitions of the MI355X solve-path library (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace fh {

// Storage precision of matrix values and vectors.
enum class Prec : int { F64 = 0, F32 = 1, BF16 = 2 };

inline size_t prec_bytes(Prec p) { return p == Prec::F64 ? 8 : (p == Prec::F32 ? 4 : 2); }

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FH_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::fh::Error(-10, std::string("HIP error ") + hipGetErrorString(e_) +  \
                                       " at " __FILE__ ":" + std::to_string(__LINE__)); \
    } while (0)

#define FH_CHECK(cond, code, msg) \
    do {                          \
        if (!(cond)) throw ::fh::Error(code, msg); \
    } while (0)

// Execution context: one device, one stream (default priority), scratch for
// reductions and a pinned mailbox for scalars read back by the host.
struct Ctx {
    int device = 0;
    bool host_only = false;         // device < 0: host set-up algorithms only, no HIP calls
    hipStream_t stream = nullptr;
    double* d_partials = nullptr;   // reduction partials (kMaxPartials per slot)
    double* d_scalars = nullptr;    // device scalars (kScalars)
    double* h_scalars = nullptr;    // pinned host mirror
    static constexpr int kMaxPartials = 2048;
    static constexpr int kPartialSlots = 4;
    static constexpr int kScalars = 512;  // K-cycle uses 32 + 8 * level + 7 (levels < 40)
    // In-situ timing of the block-row kernel on one matrix (HIP events on
    // this stream around every launch of that operation).
    // set while an allocation-free entry point runs: no execution-layout
    // decision (which allocates) and no event-timed launches
    bool no_alloc = false;
    const void* prof_matrix = nullptr;
    int prof_mode = -1;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_events;
    explicit Ctx(int dev);
    ~Ctx();
    void sync() {
        if (!host_only) FH_HIP(hipStreamSynchronize(stream));
    }
};

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) { alloc(b); }
    void alloc(size_t b) {
        release();
        bytes = b;
        if (b) FH_HIP(hipMalloc(&p, b));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Device vector of a given precision.
struct DVec {
    int64_t n = 0;
    Prec prec = Prec::F64;
    DevBuf buf;
    DVec() = default;
    DVec(int64_t n_, Prec p) { resize(n_, p); }
    void resize(int64_t n_, Prec p) {
        if (n_ == n && p == prec && buf.p) return;
        n = n_; prec = p;
        buf.alloc((size_t)n * prec_bytes(p));
    }
    void* data() const { return buf.p; }
};

// Sliced execution layout of a square-block matrix whose rows fit one
// wavefront.  A wavefront owns RW consecutive block rows; lane
// (r, b, jp) = r * W * JP + b * JP + jp holds block slot b of row r and its
// column group jp (cw = 2 columns for even block sizes, else 1; JP = nf/cw).
// Values of slice s are stored [s][i][lane < ACT][cw] (ACT = RW * W * JP), so
// load i of a wavefront reads ACT * cw contiguous values; cols holds the
// block column of every (slice, row, slot), -1 for padding.
// Execution copies of a block-CSR operator's values for full-matrix
// products.  kind 0 (rows that fit a wavefront): slices of RW rows with W
// block slots each, over the nrows rows of this part; rowmap (empty: the
// identity) gives the operator row of part row r, and `next` holds a
// further part (rows split by length into classes of their own width).
// kind 1 (wide rows): every row split into chunks of W block slots
// (RW = 1), chunk_ptr[row] .. chunk_ptr[row + 1] being its chunks, one
// wavefront per row.  kind 2 (scalar square operators): SELL-64, slice s
// of 64 rows with chunk_ptr[s + 1] - chunk_ptr[s] slots, entries and
// columns slot-major from 64 chunk_ptr[s] (padding: value 0, column = row).
struct Sliced {
    int kind = 0;
    int W = 0, RW = 0, cw = 1, nslices = 0;  // kind 1: nslices = number of chunks
    int nrows = 0;
    DevBuf cols, vals, chunk_ptr, rowmap;
    std::unique_ptr<Sliced> next;
    int act(int nf) const { return RW * W * (nf / cw); }
};

// Element execution layout of a statically condensed operator
// A = sum_c P_c^T S_c P_c (S_c: the condensed matrix of cell c over its
// nfc faces, m = nfc * nf unknowns, m <= 64 and even), kept from the
// assembly.  Each S_c is symmetric and stored by cyclic diagonals: diagonal
// t = 0 .. m/2 holds S[l][(l + t) mod m] (l = 0 .. m-1), so every unordered
// pair of unknowns is stored once (twice on diagonal m/2).  Diagonal 0 sits
// at l, diagonals 2p + 1 and 2p + 2 interleaved at m + 2 m p + 2 l + {0, 1}
// (one two-value load per lane and pair; an unused slot when m/2 is odd);
// elem_stride(m) values per cell.  Cells are in two colours
// such that every unknown face belongs to exactly one cell of each colour (a
// bipartite cell graph, such as a hexahedral grid's); colour 0 first
// (ncol[0] cells), then colour 1.  rows holds the block row of every
// (cell, face), -1 for an eliminated face.
inline int elem_stride(int m) { return m * (1 + 2 * ((m / 2 + 1) / 2)); }
inline int elem_pos(int m, int t, int l) { return t == 0 ? l : m + 2 * m * ((t - 1) / 2) + 2 * l + (t - 1) % 2; }
struct ElemLayout {
    int nfc = 0, nf = 0, m = 0, stride = 0;
    int ncol[2] = {0, 0};
    Prec prec = Prec::F64;
    std::shared_ptr<DevBuf> rows;
    DevBuf vals;
    int64_t ncell() const { return (int64_t)ncol[0] + ncol[1]; }
    size_t value_bytes() const { return (size_t)ncell() * stride * prec_bytes(prec); }
};

// Block-CSR matrix with dense nf_r x nf_c blocks stored row-major, block
// rows contiguous.  Value offsets are 64-bit (block index * nf_r * nf_c).
struct Bcsr {
    int nbr = 0, nbc = 0, bs_r = 1, bs_c = 1;
    int64_t nnzb = 0;
    Prec prec = Prec::F64;
    DevBuf rowptr, colidx, vals;
    // host mirror of the structure (kept for set-up: aggregation, colouring)
    std::vector<int> h_rowptr, h_colidx;
    // host copy of FP64 values (kept only when a host set-up step needs it)
    std::vector<double> h_vals;
    // stored-entry mask of the scalar pattern inside the blocks (1 = the
    // entry exists in the scalar sparse matrix, even if its value is 0);
    // empty: an entry is stored iff its value is non-zero
    std::vector<uint8_t> h_mask;
    // Sliced copy of the values used by full-matrix products (built from the
    // block-CSR values; the values must not change afterwards).  State:
    // 0 undecided (decided at the first full-matrix product), 1 built,
    // -1 not used.
    mutable std::unique_ptr<Sliced> sliced;
    mutable int sliced_state = 0;
    // element layout (from the assembly; null if none) and whether full
    // products, residuals and Chebyshev steps use it
    std::shared_ptr<ElemLayout> elem;
    mutable bool elem_on = false;
    int64_t rows() const { return (int64_t)nbr * bs_r; }
    int64_t cols() const { return (int64_t)nbc * bs_c; }
    size_t value_bytes() const { return (size_t)nnzb * bs_r * bs_c * prec_bytes(prec); }
};

}  // namespace fh
