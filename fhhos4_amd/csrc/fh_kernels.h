// Host-side launchers of the gfx950 kernels (fh_kernels.hip).
#pragma once
#include "fh_common.h"

namespace fh {

// Operation of the block-row kernel.
enum BcsrMode : int {
    OP_SPMV = 0,    // y = alpha * A x + beta * y
    OP_RESID = 1,   // y = b - A x
    OP_SMOOTH = 2,  // y = x + omega * D^-1 (b - A x)   (block Jacobi / GS colour)
    OP_CHEB = 3,    // d = alpha * d + beta * D^-1 (b - A x), y = x + d   (Chebyshev step)
    OP_RESID_HEAD = 4,  // rc[i] = (b - A x)[i nf]: launch_resid_head (in-situ timing code only)
};

// Block-row operation over all block rows, or over `rows[0..nrows)` when
// rows != nullptr (a colour of the multicolour Gauss-Seidel; then y may
// alias x).  Vectors are in `vprec` (F64 or F32); matrix values in A.prec;
// invD (nbr blocks nf x nf) in dprec.
void launch_bcsr(Ctx& c, const Bcsr& A, BcsrMode mode, Prec vprec, const void* x, const void* b, void* y,
                 double alpha, double beta, double omega, const void* invD, Prec dprec,
                 const int* rows, int nrows);

// One Chebyshev step over all block rows, fused: d = a d + c D^-1 (b - A x)
// (d not read when a = 0) and xnew = x + d.  xnew must not alias x.
void launch_bcsr_cheb(Ctx& c, const Bcsr& A, Prec vprec, const void* x, const void* b, void* xnew, void* d,
                      double a, double cc, const void* invD, Prec dprec);

// rc[i] = (b - A x)[i nf] for every block row i: the residual truncated to
// the first mode of each block (the p-restriction to one mode per face),
// reading entry 0 of every block only.  Returns false, having launched
// nothing, unless A has a one-part sliced layout (and x is 2-value aligned
// for even nf); the caller then forms the residual and restricts it.
bool launch_resid_head(Ctx& c, const Bcsr& A, Prec vprec, const void* x, const void* b, void* rc);

// y = A x for an FP32-stored A with FP64 vectors and FP64 accumulation, on
// the one-part sliced copy (n_f >= 2) or the SELL-64 copy (n_f = 1).
// Returns false, having launched nothing, when A has neither.
bool launch_spmv_f32vals(Ctx& c, const Bcsr& A, const void* x, void* y);

// Builds (on = true) or drops the sliced execution layout of A used by
// full-matrix products; returns whether A now has one (false when a row of
// A does not fit one wavefront).  Without a call the layout is chosen at the
// first full-matrix product (rows fit, >= 4096 block rows, <= 25% padding).
bool set_sliced(Ctx& c, const Bcsr& A, bool on);
// Counter of execution-layout changes (built, dropped or decided) of any
// matrix; a graph captured before a change must not be replayed after it.
uint64_t layout_epoch();
void layout_epoch_bump();

// Element layout (ElemLayout): y = alpha A x + beta y (OP_SPMV), y = b - A x
// (OP_RESID) or the fused Chebyshev step (OP_CHEB, d in aux), as two
// launches, one per cell colour.  Returns false, having launched nothing,
// when A has no active element layout or the mode and precisions are not
// covered (FP64 values and vectors; FP32 values and vectors; FP32 values
// with FP64 vectors for OP_SPMV and OP_RESID).  x must not alias y.
bool launch_elem(Ctx& c, const Bcsr& A, BcsrMode mode, Prec vprec, const void* x, const void* b, void* y,
                 double alpha, double beta, const void* invD, Prec dprec, void* aux);
// Makes the element layout of A active (on, when A has one) or not; returns
// whether it is active.  The layout's data stay with A either way.
bool set_elem(const Bcsr& A, bool on);
// false when the environment sets FHGPU_ELEM=0 (element layouts unused)
bool elem_enabled();
// FP32 copy of an FP64 element layout (cell rows shared).
std::shared_ptr<ElemLayout> elem_convert(Ctx& c, const ElemLayout& E, Prec p);

// Scratch of the allocation-free entry points: 2^21 doubles (16 MiB) per
// device, a file-scope device array.  noalloc_scratch_init(device) looks up
// the device's copy (call it with that device current, once per device);
// noalloc_scratch(device) returns that address.
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 21;
constexpr int kMaxDevices = 64;
void noalloc_scratch_init(int device);
void* noalloc_scratch(int device);

// y = omega * D^-1 b over all block rows (smoothing from a zero guess); the
// same values also into y2 when it is not null (first Chebyshev step).
void launch_bjac_zero(Ctx& c, int nbr, int nf, Prec vprec, const void* b, void* y, double omega,
                      const void* invD, Prec dprec, const int* rows, int nrows, void* y2 = nullptr);

// Batched inverse of the diagonal blocks of A (FP64 values, Gauss-Jordan
// with partial pivoting per block); diag[i] = block index of A_ii.  Writes
// nbr * nf * nf entries in dprec.  *d_flag (device int) is set to i+1 when
// block i is singular.
void launch_block_inverse(Ctx& c, const Bcsr& A, const int* diag, void* invD, Prec dprec, int* d_flag);

// Conversions / BLAS-1 over n entries.
void launch_convert(Ctx& c, int64_t n, const void* src, Prec sp, void* dst, Prec dp);
void launch_fill(Ctx& c, int64_t n, void* x, Prec p, double v);
// y = a*x + b*y, with a = ha * (ia >= 0 ? s[ia] : 1) * (ib >= 0 ? 1/s[ib] : 1)
// (device scalars s; ia, ib < 0 use the host value only); same for b with
// (hb, ic, id).
void launch_axpby(Ctx& c, int64_t n, Prec p, const void* x, void* y, double ha, int ia, int ib, double hb,
                  int ic, int id);
// FCG orthogonalisation (FP64): out = z - sum_k (s[num[k]] / s[den[k]]) vs[k],
// the nv terms added in order, one pass per 4 of them (out may alias z; a
// copy when nv = 0).
void launch_fcg_orth(Ctx& c, int64_t n, const double* z, int nv, const double* const* vs, const int* num,
                     const int* den, double* out);
// The FP64 update fused with s[slot_rr] = <r, r> of the updated residual
// (the value a separate launch_dots gives).
void launch_cg_update_rr(Ctx& c, int64_t n, double* x, const double* d, double* r, const double* q, int inum,
                         int iden, double sign, int slot_rr);
// x += a d ; r -= a q with a = s[inum] / s[iden]  (CG / FCG update).
void launch_cg_update(Ctx& c, int64_t n, Prec p, void* x, const void* d, void* r, const void* q, int inum, int iden,
                      double sign);
// Deterministic dot products: s[slot_k] = <x_k, y_k>, k < nd (nd <= 3),
// accumulated in FP64 with a fixed reduction tree.
void launch_dots(Ctx& c, int64_t n, Prec p, int nd, const void* const* xs, const void* const* ys, const int* slots);
// Device scalar arithmetic: s[dst] = op(s[a], s[b]); op 0: a/b, 1: sqrt(a),
// 2: a*b, 3: a-b, 4: copy a.
void launch_scalar_op(Ctx& c, int op, int dst, int a, int b);

// K-cycle coefficients on the device: slots first .. first+6 hold rho1,
// alpha1, |r|^2, |r_new|^2, gamma, beta, alpha2; ca and cb (cb = 0 when
// |r_new| <= t |r|: the first correction alone) replace first+2, first+3.
void launch_kcycle_coeffs(Ctx& c, int first, double t);

// Aggregation transfers (piecewise-constant prolongation, block identity).
// restrict: rc[I*nf+a] = sum_{k in members(I)} r[k*nf+a]
void launch_agg_restrict(Ctx& c, int nagg, int nf, Prec p, const int* mptr, const int* members, const void* r,
                         void* rc);
// prolong-add: x[k*nf+a] += ec[agg[k]*nf+a]
void launch_agg_prolong_add(Ctx& c, int nb, int nf, Prec p, const int* agg, const void* ec, void* x);

// p-transfers between block sizes nf (fine) and nfc <= nf (coarse) for
// hierarchical bases: restriction keeps the first nfc modes of every block,
// prolongation injects them (adds into x).
void launch_p_restrict(Ctx& c, int nb, int nf, int nfc, Prec p, const void* r, void* rc);
void launch_p_prolong_add(Ctx& c, int nb, int nf, int nfc, Prec p, const void* ec, void* x);

// S[p] = leading nfc x nfc sub-block of the nf x nf block A[p], p < nnzb (FP64).
void launch_mode_submatrix(Ctx& c, int64_t nnzb, int nf, int nfc, const double* A, double* S);

// y = M x with M dense symmetric n x n (precision mp), vectors in vp.
void launch_dense_symv(Ctx& c, int n, const void* M, Prec mp, Prec vp, const void* x, void* y);

}  // namespace fh
