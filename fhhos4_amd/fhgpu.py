"""ctypes binding of libfhgpu.so (the C ABI declared in include/fhgpu.h).

Device vectors are passed as raw HBM pointers; this module accepts torch
tensors on a ROCm device (``tensor.data_ptr()``) or integers.  The library
must have been built (``fhhos4_amd.build.build()``): there is no fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfhgpu.so")

F64, F32, BF16 = 0, 1, 2
_PREC = {"f64": F64, "fp64": F64, F64: F64, "f32": F32, "fp32": F32, F32: F32, "bf16": BF16, BF16: BF16}


class FhError(RuntimeError):
    pass


class SolverArgs(C.Structure):
    _fields_ = [("solver", C.c_char_p), ("prec", C.c_char_p), ("cycle", C.c_char), ("wloops", C.c_int),
                ("nu1", C.c_int), ("nu2", C.c_int), ("pre_smoother", C.c_char_p), ("post_smoother", C.c_char_p),
                ("relax", C.c_double), ("tol", C.c_double), ("max_iter", C.c_int), ("max_coarse", C.c_int),
                ("theta", C.c_double), ("mg_prec", C.c_int), ("mg_store", C.c_int), ("galerkin", C.c_int),
                ("coarse_smoother", C.c_char_p), ("coarse_relax", C.c_double), ("fcg_truncation", C.c_int),
                ("fcg_restart", C.c_int), ("coarse_store", C.c_int), ("kcycle_levels", C.c_int),
                ("nu_change_op", C.c_char), ("nu_change_coeff", C.c_int), ("outer_prec", C.c_int),
                ("replace_every", C.c_int), ("coarse_nu_given", C.c_int), ("coarse_nu1", C.c_int),
                ("coarse_nu2", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("levels", C.c_int), ("graph_mode", C.c_int),
                ("final_res", C.c_double), ("setup_s", C.c_double), ("solve_s", C.c_double),
                ("op_complexity", C.c_double), ("nn_used", C.c_int), ("nn_fallback", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes)
_SIGS = {
    "fh_ctx_create": [C.c_int, _pp],
    "fh_ctx_destroy": [_vp],
    "fh_ctx_sync": [_vp],
    "fh_mat_from_csr": [_vp, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_int, _pp],
    "fh_mat_from_blocks": [_vp, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_int, _pp],
    "fh_mat_convert": [_vp, _vp, C.c_int, _pp],
    "fh_mat_destroy": [_vp],
    "fh_mat_info": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                    C.POINTER(C.c_int)],
    "fh_mat_device_arrays": [_vp, _pp, _pp, _pp],
    "fh_mat_blocks": [_vp, _i32p, _i32p, _f64p],
    "fh_mat_set_sliced": [_vp, _vp, C.c_int, C.POINTER(C.c_int)],
    "fh_mat_set_elem": [_vp, _vp, C.c_int, C.POINTER(C.c_int)],
    "fh_mat_layout": [_vp, C.POINTER(C.c_int)],
    "fh_nn_create": [_vp, C.c_int, C.c_int, C.c_uint64, _pp],
    "fh_nn_destroy": [_vp],
    "fh_nn_num_params": [_vp, C.POINTER(C.c_int64)],
    "fh_nn_set_params": [_vp, _f32p],
    "fh_nn_get_params": [_vp, _f32p],
    "fh_nn_graph_create": [_vp, _vp, _vp, _pp],
    "fh_nn_graph_destroy": [_vp],
    "fh_nn_forward": [_vp, _vp, _vp, _vp, _vp, C.c_int],
    "fh_nn_loss_grad": [_vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_double), _vp],
    "fh_nn_train_step": [_vp, _vp, _vp, _vp, _vp, C.c_float, C.POINTER(C.c_double)],
    "fh_solver_set_nn": [_vp, _vp, _vp, C.c_double, C.c_int],
    "fh_solver_set_stop": [_vp, C.c_double, C.c_int],
    "fh_solver_set_level_p": [_vp, C.c_int, C.c_int],
    "fh_solver_export_levels": [_vp, C.c_char_p],
    "fh_hho_assemble_gpu": [_vp, _vp, C.c_int, _pp, _vp],
    "fh_solver_apply_prec": [_vp, _vp, _vp],
    "fh_mat_export_market": [_vp, _vp, C.c_char_p],
    "fh_mat_import_market": [_vp, C.c_char_p, C.c_int, C.c_int, _pp],
    "fh_vec_export_market": [C.c_int64, _f64p, C.c_char_p],
    "fh_vec_import_market": [C.c_char_p, C.c_int64, _vp, C.POINTER(C.c_int64)],
    "fh_spmv": [_vp, _vp, C.c_double, _vp, C.c_double, _vp, C.c_int],
    "fh_residual": [_vp, _vp, _vp, _vp, _vp, C.c_int],
    "fh_dot": [_vp, C.c_int64, _vp, _vp, C.c_int, C.POINTER(C.c_double)],
    "fh_axpby": [_vp, C.c_int64, C.c_double, _vp, C.c_double, _vp, C.c_int],
    "fh_convert": [_vp, C.c_int64, _vp, C.c_int, _vp, C.c_int],
    "fh_bjac_create": [_vp, _vp, C.c_double, C.c_int, _pp],
    "fh_bjac_destroy": [_vp],
    "fh_bjac_inverses": [_vp, _vp, _f64p],
    "fh_bjac_sweep": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int],
    "fh_bjac_sweep_noalloc": [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int],
    "fh_coloring_create": [_vp, _vp, _pp],
    "fh_coloring_destroy": [_vp],
    "fh_coloring_get": [_vp, C.POINTER(C.c_int), _i32p],
    "fh_mcgs_sweep": [_vp, _vp, _vp, _vp, C.c_int, _vp, _vp, C.c_int],
    "fh_aggregate_pairwise": [_vp, _vp, C.c_double, _i32p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "fh_galerkin_aggregate": [_vp, _vp, _i32p, C.c_int, _pp],
    "fh_galerkin_rap": [_vp, _vp, _vp, _pp],
    "fh_solver_create": [_vp, C.POINTER(SolverArgs), _pp],
    "fh_solver_destroy": [_vp],
    "fh_solver_set_level": [_vp, C.c_int, _vp, _vp],
    "fh_solver_setup": [_vp, _vp],
    "fh_solver_solve": [_vp, _vp, _vp, C.POINTER(SolveStats)],
    "fh_solver_solve_x0": [_vp, _vp, _vp, C.c_int, C.POINTER(SolveStats)],
    "fh_solver_history": [_vp, _f64p, C.c_int],
    "fh_solver_level_size": [_vp, C.c_int],
    "fh_profile_begin": [_vp, _vp, C.c_int],
    "fh_profile_end": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_double)],
    "fh_hho_create": [_vp, _pp],
    "fh_hho_destroy": [_vp],
    "fh_hho_get_info": [_vp, _vp],
    "fh_hho_matrix": [_vp, _vp, C.c_int, _pp],
    "fh_hho_rhs": [_vp, _f64p],
    "fh_hho_blocks": [_vp, _i32p, _i32p, _f64p],
    "fh_hho_l2_error": [_vp, _f64p, C.POINTER(C.c_double)],
    "fh_hho_prolongation": [_vp, _vp, _vp, C.c_int, _pp],
    "fh_hho_face_cells": [_vp, _i32p],
    "fh_hho_cell_kappa": [_vp, _f64p],
    "fh_condense_batched": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _vp, _vp, _vp, _vp, _pp],
    "fh_cell_recover_batched": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _vp, _vp, _vp, _vp],
    "fh_hho_uncondensed": [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64), _vp, _vp, _vp],
    "fh_hho_cell_unknowns": [_vp, C.POINTER(C.c_int)],
    "fh_hho_l2_error_gpu": [_vp, _vp, _vp, C.POINTER(C.c_double)],
    "fh_solver_setup_uncondensed": [_vp, _vp, _vp, _vp, _vp, C.c_int],
    "fh_uamg_build": [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_double, C.c_int, _pp],
    "fh_uamg_destroy": [_vp],
    "fh_uamg_info": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "fh_uamg_level": [_vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64), _vp, _vp,
                      _vp],
    "fh_uamg_pass": [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _vp, _vp, _vp],
}


class HhoArgs(C.Structure):
    _fields_ = [("dim", C.c_int), ("mesh", C.c_char_p), ("n", C.c_int), ("k", C.c_int), ("perturb", C.c_double),
                ("seed", C.c_uint64), ("source", C.c_char_p), ("kappa", C.c_char_p), ("sigma", C.c_double),
                ("subdomain", C.c_int), ("aniso", C.c_double), ("ref_n", C.c_int), ("source_seed", C.c_uint64),
                ("source_modes", C.c_int), ("defer_assembly", C.c_int)]


class HhoInfo(C.Structure):
    _fields_ = [("nb", C.c_int), ("nf", C.c_int), ("ncells", C.c_int), ("nfaces", C.c_int), ("nnzb", C.c_int64),
                ("mesh_s", C.c_double), ("local_s", C.c_double), ("assemble_s", C.c_double),
                ("redraws", C.c_int64), ("kept_vertices", C.c_int64), ("min_quality", C.c_double)]


def lib():
    """Loads libfhgpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FhError("libfhgpu.so is missing: build it with fhhos4_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.fh_last_error.restype = C.c_char_p
        L.fh_last_error.argtypes = []
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS) + ["fh_last_error"]


def _check(rc):
    if rc < 0:
        raise FhError("fhgpu error %d: %s" % (rc, lib().fh_last_error().decode()))
    return rc


def _device_sync():
    """Device-wide synchronisation.  The library runs on its own non-blocking
    stream: inputs written by torch must be complete before its kernels read
    them, and its outputs complete before torch (or the host) reads them or
    frees a temporary input."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.synchronize()


def _ordered(fn):
    """Wraps a call that passes device tensors to the library with device
    synchronisations before and after it."""
    import functools

    @functools.wraps(fn)
    def wrapper(*a, **k):
        _device_sync()
        try:
            return fn(*a, **k)
        finally:
            _device_sync()
    return wrapper


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _dev(t, n=None, like=None, dtype=None, name="vector"):
    """Device pointer of tensor t after checking what the C entry points
    assume but cannot see (they take no lengths): a contiguous tensor on the
    ROCm device, of dtype `dtype` (or that of `like`), with exactly n
    entries.  Raw integer pointers are passed through unchecked."""
    if isinstance(t, int):
        return t
    import torch
    if not isinstance(t, torch.Tensor):
        raise FhError("%s must be a torch tensor on the device (or a raw pointer)" % name)
    if not t.is_cuda:
        raise FhError("%s must be in device memory" % name)
    if not t.is_contiguous():
        raise FhError("%s must be contiguous" % name)
    want = dtype if dtype is not None else (like.dtype if like is not None and not isinstance(like, int) else None)
    if want is not None and t.dtype != want:
        raise FhError("%s has dtype %s, expected %s" % (name, t.dtype, want))
    if t.dtype not in (torch.float64, torch.float32):
        raise FhError("%s must be float64 or float32" % name)
    if n is not None and t.numel() != n:
        raise FhError("%s has %d entries, expected %d" % (name, t.numel(), n))
    return t.data_ptr()


def _f64():
    import torch
    return torch.float64


def _vprec(t):
    try:
        import torch
        if t.dtype == torch.float64:
            return F64
        if t.dtype == torch.float32:
            return F32
    except Exception:
        pass
    raise FhError("device vectors must be float64 or float32 tensors")


class Context:
    """Execution context on one device (device < 0: host set-up only)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().fh_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def sync(self):
        _check(lib().fh_ctx_sync(self.h))

    def close(self):
        if self.h:
            lib().fh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Matrix:
    """Block-CSR operator (dense bs x bs blocks) on the device."""

    def __init__(self, ctx, handle, keep=None):
        self.ctx, self.h, self._keep = ctx, handle, keep

    @classmethod
    def from_csr(cls, ctx, A, bs, keep_host=True):
        A = A.tocsr()
        A.sort_indices()
        rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(A.indices, dtype=np.int32)
        v = np.ascontiguousarray(A.data, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().fh_mat_from_csr(ctx.h, A.shape[0], A.shape[1], bs, rp, ci, v, int(keep_host), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_blocks(cls, ctx, nbc, bs, rowptr, colidx, vals, keep_host=True):
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        ci = np.ascontiguousarray(colidx, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1)
        h = C.c_void_p()
        _check(lib().fh_mat_from_blocks(ctx.h, len(rp) - 1, nbc, bs, rp, ci, v, int(keep_host), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def read_market(cls, ctx, path, bs, keep_host=True):
        """Matrix Market file (the reference's A.dat format) -> block matrix."""
        h = C.c_void_p()
        _check(lib().fh_mat_import_market(ctx.h, os.fsencode(path), bs, int(keep_host), C.byref(h)))
        return cls(ctx, h)

    def write_market(self, path):
        """Writes the scalar operator in the reference's A.dat format."""
        _check(lib().fh_mat_export_market(self.ctx.h, self.h, os.fsencode(path)))

    def blocks(self):
        """Host copies of the block arrays (row pointers, block columns,
        nnzb x bs x bs FP64 values; device-only matrices are read from HBM)."""
        i = self.info()
        rp = np.zeros(i["nbr"] + 1, dtype=np.int32)
        ci = np.zeros(i["nnzb"], dtype=np.int32)
        v = np.empty(i["nnzb"] * i["bs"] * i["bs"])
        _check(lib().fh_mat_blocks(self.h, rp, ci, v))
        return rp, ci, v.reshape(-1, i["bs"], i["bs"])

    def scipy(self):
        """Host copy as a scipy CSR matrix (matrix created with keep_host)."""
        import scipy.sparse as sp
        i = self.info()
        rp = np.zeros(i["nbr"] + 1, dtype=np.int32)
        ci = np.zeros(i["nnzb"], dtype=np.int32)
        v = np.zeros(i["nnzb"] * i["bs"] * i["bs"])
        _check(lib().fh_mat_blocks(self.h, rp, ci, v))
        bs = i["bs"]
        return sp.bsr_matrix((v.reshape(-1, bs, bs), ci, rp), shape=(i["nbr"] * bs, i["nbc"] * bs)).tocsr()

    @_ordered
    def set_sliced(self, on):
        """Builds (on=True) or drops the sliced execution layout of full-matrix
        products; returns whether the matrix now has one."""
        a = C.c_int()
        _check(lib().fh_mat_set_sliced(self.ctx.h, self.h, int(bool(on)), C.byref(a)))
        return bool(a.value)

    def set_elem(self, on):
        """Uses (on=True) or stops using the element layout (condensed cell
        matrices kept by the GPU assembly of hexahedral meshes) for full
        products, residuals and Chebyshev steps; returns whether it is used."""
        a = C.c_int()
        _check(lib().fh_mat_set_elem(self.ctx.h, self.h, int(bool(on)), C.byref(a)))
        return bool(a.value)

    def layout(self):
        """0 block CSR (or undecided), 1 sliced, 2 sliced in two parts, 3 chunked,
        4 SELL-64, 5 element layout."""
        a = C.c_int()
        _check(lib().fh_mat_layout(self.h, C.byref(a)))
        return a.value

    def convert(self, prec):
        h = C.c_void_p()
        _check(lib().fh_mat_convert(self.ctx.h, self.h, _PREC[prec], C.byref(h)))
        return Matrix(self.ctx, h)

    @property
    def rows(self):
        i = self.info()
        return i["nbr"] * i["bs"]

    @property
    def cols(self):
        i = self.info()
        return i["nbc"] * i["bs"]

    def info(self):
        nbr, nbc, bs, pr = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        nnzb = C.c_int64()
        _check(lib().fh_mat_info(self.h, C.byref(nbr), C.byref(nbc), C.byref(bs), C.byref(nnzb), C.byref(pr)))
        return {"nbr": nbr.value, "nbc": nbc.value, "bs": bs.value, "nnzb": nnzb.value, "prec": pr.value}

    def close(self):
        if self.h:
            lib().fh_mat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_market_vector(v, path):
    """Host vector -> the reference's b.dat / solutionFaces.dat format."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    _check(lib().fh_vec_export_market(v.size, v, os.fsencode(path)))


def read_market_vector(path):
    n = C.c_int64()
    _check(lib().fh_vec_import_market(os.fsencode(path), 0, None, C.byref(n)))
    v = np.zeros(n.value)
    _check(lib().fh_vec_import_market(os.fsencode(path), v.size, v.ctypes.data, C.byref(n)))
    return v


@_ordered
def spmv(ctx, A, x, y, alpha=1.0, beta=0.0):
    px = _dev(x, A.cols, name="x")
    py = _dev(y, A.rows, like=x, name="y")
    _check(lib().fh_spmv(ctx.h, A.h, alpha, px, beta, py, _vprec(x)))


@_ordered
def residual(ctx, A, b, x, r):
    n = A.rows
    px = _dev(x, n, name="x")
    pb = _dev(b, n, like=x, name="b")
    pr = _dev(r, n, like=x, name="r")
    _check(lib().fh_residual(ctx.h, A.h, pb, px, pr, _vprec(x)))


@_ordered
def dot(ctx, x, y):
    out = C.c_double()
    px = _dev(x, name="x")
    py = _dev(y, x.numel(), like=x, name="y")
    _check(lib().fh_dot(ctx.h, x.numel(), px, py, _vprec(x), C.byref(out)))
    return out.value


@_ordered
def axpby(ctx, a, x, b, y):
    px = _dev(x, name="x")
    py = _dev(y, x.numel(), like=x, name="y")
    _check(lib().fh_axpby(ctx.h, x.numel(), a, px, b, py, _vprec(x)))


class BlockJacobi:
    def __init__(self, ctx, A, omega=1.0, dprec=F64):
        h = C.c_void_p()
        _check(lib().fh_bjac_create(ctx.h, A.h, omega, _PREC[dprec], C.byref(h)))
        self.ctx, self.h, self.A = ctx, h, A

    def inverses(self, nbr, bs):
        out = np.zeros(nbr * bs * bs)
        _check(lib().fh_bjac_inverses(self.ctx.h, self.h, out))
        return out

    @_ordered
    def sweep(self, b, x, work, nsweeps=1, x_is_zero=False, A=None):
        M = A or self.A
        n = M.rows
        px = _dev(x, n, name="x")
        pb = _dev(b, n, like=x, name="b")
        pw = _dev(work, n, like=x, name="work")
        _check(lib().fh_bjac_sweep(self.ctx.h, self.h, M.h, pb, px, pw, nsweeps, int(x_is_zero), _vprec(x)))

    @_ordered
    def sweep_noalloc(self, b, x, nsweeps=1, x_is_zero=False, A=None):
        """The same sweeps without a work vector (fh_bjac_sweep_noalloc)."""
        M = A or self.A
        n = M.rows
        px = _dev(x, n, name="x")
        pb = _dev(b, n, like=x, name="b")
        _check(lib().fh_bjac_sweep_noalloc(self.ctx.h, self.h, M.h, pb, px, nsweeps, int(x_is_zero), _vprec(x)))

    def __del__(self):
        try:
            lib().fh_bjac_destroy(self.h)
        except Exception:
            pass


class Colouring:
    def __init__(self, ctx, A):
        h = C.c_void_p()
        _check(lib().fh_coloring_create(ctx.h, A.h, C.byref(h)))
        self.ctx, self.h, self.nbr = ctx, h, A.info()["nbr"]

    def get(self):
        nc = C.c_int()
        col = np.zeros(self.nbr, dtype=np.int32)
        _check(lib().fh_coloring_get(self.h, C.byref(nc), col))
        return nc.value, col

    def __del__(self):
        try:
            lib().fh_coloring_destroy(self.h)
        except Exception:
            pass


@_ordered
def mcgs_sweep(ctx, A, bj, col, direction, b, x):
    n = A.rows
    px = _dev(x, n, name="x")
    pb = _dev(b, n, like=x, name="b")
    _check(lib().fh_mcgs_sweep(ctx.h, A.h, bj.h, col.h, direction, pb, px, _vprec(x)))


def _cell_args(cell_faces, local, nc, nf):
    import torch
    faces = np.ascontiguousarray(cell_faces, dtype=np.int32)
    if faces.ndim != 2:
        raise FhError("cell_faces must be a (cells, faces per cell) array")
    ncells, nfc = faces.shape
    nl = nc + nfc * nf
    _dev(local, ncells * nl * nl, dtype=torch.float64, name="local")
    return faces, ncells, nfc


@_ordered
def condense_batched(ctx, cell_faces, local, nc, nf, nfaces, B_T=None, B_F=None, rhs=True):
    """Static condensation of per-cell local matrices on the GPU
    (fh_condense_batched): cell_faces (cells x faces per cell, host, -1 for
    an eliminated face), local (device FP64, cells x nl x nl with the cell
    unknowns first), optional device B_T (cells x nc) and B_F (nfaces x nf).
    Returns (Matrix, b) with b a device FP64 tensor, or None without rhs."""
    import torch
    faces, ncells, nfc = _cell_args(cell_faces, local, nc, nf)
    bt = None if B_T is None else _dev(B_T, ncells * nc, dtype=torch.float64, name="B_T")
    bf = None if B_F is None else _dev(B_F, nfaces * nf, dtype=torch.float64, name="B_F")
    b = torch.empty(nfaces * nf, dtype=torch.float64, device="cuda") if rhs else None
    h = C.c_void_p()
    _check(lib().fh_condense_batched(ctx.h, ncells, nfc, nc, nf, nfaces, faces.reshape(-1), _ptr(local), bt, bf,
                                     None if b is None else b.data_ptr(), C.byref(h)))
    return Matrix(ctx, h), b


@_ordered
def cell_recover_batched(ctx, cell_faces, local, nc, nf, nfaces, x_F, B_T=None):
    """Cell unknowns x_T = A_TT^-1 (B_T - A_TF x_F) of every cell
    (fh_cell_recover_batched); returns a device FP64 tensor (cells x nc)."""
    import torch
    faces, ncells, nfc = _cell_args(cell_faces, local, nc, nf)
    bt = None if B_T is None else _dev(B_T, ncells * nc, dtype=torch.float64, name="B_T")
    xf = _dev(x_F, nfaces * nf, dtype=torch.float64, name="x_F")
    xt = torch.empty(ncells * nc, dtype=torch.float64, device="cuda")
    _check(lib().fh_cell_recover_batched(ctx.h, ncells, nfc, nc, nf, nfaces, faces.reshape(-1), _ptr(local), bt, xf,
                                         xt.data_ptr()))
    return xt.reshape(ncells, nc)


def aggregate_pairwise(ctx, A, theta=0.25):
    nbr = A.info()["nbr"]
    agg = np.zeros(nbr, dtype=np.int32)
    na, co = C.c_int(), C.c_int()
    _check(lib().fh_aggregate_pairwise(ctx.h, A.h, theta, agg, C.byref(na), C.byref(co)))
    return na.value, agg, bool(co.value)


def galerkin_aggregate(ctx, A, agg, nagg):
    h = C.c_void_p()
    _check(lib().fh_galerkin_aggregate(ctx.h, A.h, np.ascontiguousarray(agg, dtype=np.int32), nagg, C.byref(h)))
    return Matrix(ctx, h)


def galerkin_rap(ctx, A, P):
    h = C.c_void_p()
    _check(lib().fh_galerkin_rap(ctx.h, A.h, P.h, C.byref(h)))
    return Matrix(ctx, h)


class Solver:
    """Solver created from the reference's string codes (SolverFactory)."""

    def __init__(self, ctx, solver="cg", prec="mg", cycle="V", nu=(1, 1), smoothers=("bj", "bj"), relax=1.0,
                 tol=1e-8, max_iter=200, max_coarse=1000, theta=0.25, mg_prec="f64", mg_store=None, wloops=0,
                 galerkin=False, coarse_smoother=None, coarse_relax=0.0, fcg_truncation=1, fcg_restart=False,
                 coarse_store=None, kcycle_levels=0, nu_change=None, outer_prec="f64", replace_every=10,
                 coarse_nu=None):
        a = SolverArgs()
        self._strs = [s.encode() for s in (solver, prec or "", smoothers[0], smoothers[1], coarse_smoother or "")]
        a.solver, a.prec, a.pre_smoother, a.post_smoother, a.coarse_smoother = self._strs
        a.coarse_relax = coarse_relax
        a.fcg_truncation, a.fcg_restart = fcg_truncation, int(fcg_restart)
        a.cycle = cycle.encode()
        a.wloops = wloops
        a.nu1, a.nu2 = nu
        a.relax, a.tol, a.max_iter = relax, tol, max_iter
        a.max_coarse, a.theta = max_coarse, theta
        a.mg_prec = _PREC[mg_prec]
        a.mg_store = _PREC[mg_store if mg_store is not None else mg_prec]
        a.galerkin = int(galerkin)
        a.coarse_store = 0 if coarse_store is None else _PREC[coarse_store]
        a.kcycle_levels = kcycle_levels
        a.outer_prec, a.replace_every = _PREC[outer_prec], replace_every
        if nu_change:  # e.g. "+1": Multigrid::CoarseLevelChangeSmoothingOperator / Coeff
            a.nu_change_op, a.nu_change_coeff = nu_change[0].encode(), int(nu_change[1:])
        if coarse_nu is not None:  # (pre, post) smoothing iterations below the finest level
            a.coarse_nu_given, (a.coarse_nu1, a.coarse_nu2) = 1, coarse_nu
        h = C.c_void_p()
        _check(lib().fh_solver_create(ctx.h, C.byref(a), C.byref(h)))
        self.ctx, self.h, self.max_iter = ctx, h, max_iter
        self._levels = []

    def export_levels(self, directory):
        """Writes level<l>_A/P/R.dat as the reference's -export mg does."""
        _check(lib().fh_solver_export_levels(self.h, os.fsencode(directory)))

    def set_stop(self, tol, max_iter):
        self.max_iter = max_iter
        _check(lib().fh_solver_set_stop(self.h, tol, max_iter))

    def _n(self):
        if getattr(self, "A", None) is None:
            raise FhError("setup() must be called first")
        return self.A.rows

    @_ordered
    def apply_prec(self, r, z):
        """z = M r with the preconditioner alone (device FP64 vectors)."""
        n = self._n()
        _check(lib().fh_solver_apply_prec(self.h, _dev(r, n, dtype=_f64(), name="r"),
                                          _dev(z, n, dtype=_f64(), name="z")))

    def set_learned_smoother(self, nn, graph, omega=1.0, unit_step=False):
        """Learned post-smoother inside FCG (before setup)."""
        self._nn = (nn, graph)
        _check(lib().fh_solver_set_nn(self.h, nn.h, graph.h, omega, int(unit_step)))

    def set_level_p(self, level, nfc):
        """p-injection from level+1 to level keeping nfc modes per face
        (Galerkin operator below unless set_level gives one)."""
        _check(lib().fh_solver_set_level_p(self.h, level, nfc))

    def set_level(self, level, A=None, P=None):
        self._levels.append((A, P))
        _check(lib().fh_solver_set_level(self.h, level, A.h if A is not None else None, P.h if P is not None else None))

    def setup(self, A):
        self.A = A
        _check(lib().fh_solver_setup(self.h, A.h))

    def setup_uncondensed(self, A, A_TT, A_TF, A_FF, cell_bs):
        """Set-up of the "uamg" multigrid from the uncondensed blocks (scipy
        matrices; A created with keep_host)."""
        self.A = A
        keep = [_CsrArg(M) for M in (A_TT, A_TF, A_FF)]
        _check(lib().fh_solver_setup_uncondensed(self.h, A.h, *[C.byref(k.s) for k in keep], int(cell_bs)))

    @_ordered
    def solve(self, b, x, guess=False):
        """Solves A x = b from x = 0, or from the initial guess held in x
        when guess is true (fh_solver_solve_x0)."""
        n = self._n()
        pb = _dev(b, n, dtype=_f64(), name="b")
        px = _dev(x, n, dtype=_f64(), name="x")
        st = SolveStats()
        if guess:
            _check(lib().fh_solver_solve_x0(self.h, pb, px, 0, C.byref(st)))
        else:
            _check(lib().fh_solver_solve(self.h, pb, px, C.byref(st)))
        return st.as_dict()

    def history(self):
        out = np.zeros(self.max_iter + 2)
        n = _check(lib().fh_solver_history(self.h, out, len(out)))
        return out[:n]

    def level_size(self, l):
        return _check(lib().fh_solver_level_size(self.h, l))

    def __del__(self):
        try:
            lib().fh_solver_destroy(self.h)
        except Exception:
            pass


class _Csr(C.Structure):
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("rowptr", C.POINTER(C.c_int)),
                ("colidx", C.POINTER(C.c_int)), ("vals", C.POINTER(C.c_double))]


class _CsrArg:
    """fh_csr view of a scipy matrix (sorted CSR; arrays kept alive)."""

    def __init__(self, M):
        M = M.tocsr()
        M.sort_indices()
        self.rp = np.ascontiguousarray(M.indptr, dtype=np.int32)
        self.ci = np.ascontiguousarray(M.indices, dtype=np.int32)
        self.v = np.ascontiguousarray(M.data, dtype=np.float64)
        self.s = _Csr(M.shape[0], M.shape[1], self.rp.ctypes.data_as(C.POINTER(C.c_int)),
                      self.ci.ctypes.data_as(C.POINTER(C.c_int)), self.v.ctypes.data_as(C.POINTER(C.c_double)))


def _csr_get(fn, *args):
    """Two-call CSR read-out: sizes, then arrays -> scipy CSR."""
    import scipy.sparse as sp
    nr, nc, nnz = C.c_int(), C.c_int(), C.c_int64()
    _check(fn(*args, C.byref(nr), C.byref(nc), C.byref(nnz), None, None, None))
    rp = np.zeros(nr.value + 1, dtype=np.int32)
    ci = np.zeros(nnz.value, dtype=np.int32)
    v = np.zeros(nnz.value)
    _check(fn(*args, None, None, None, rp.ctypes.data, ci.ctypes.data, v.ctypes.data))
    return sp.csr_matrix((v, ci, rp), shape=(nr.value, nc.value))


class UncondensedAMG:
    """U-AMG hierarchy built on the host (UncondensedAMG set-up): level
    operators A[l], prolongations P[l] (level l+1 -> l) and the aggregation of
    every coarsening pass."""

    def __init__(self, A, A_TT, A_TF, A_FF, cell_bs, face_bs, theta=0.25, max_coarse=1000):
        args = [_CsrArg(M) for M in (A, A_TT, A_TF, A_FF)]
        h = C.c_void_p()
        _check(lib().fh_uamg_build(*[C.byref(a.s) for a in args], int(cell_bs), int(face_bs), float(theta),
                                   int(max_coarse), C.byref(h)))
        self.h = h
        nl, npass = C.c_int(), C.c_int()
        _check(lib().fh_uamg_info(self.h, C.byref(nl), C.byref(npass)))
        self.A = [_csr_get(lib().fh_uamg_level, self.h, l, 0) for l in range(nl.value)]
        self.P = [_csr_get(lib().fh_uamg_level, self.h, l, 1) for l in range(nl.value - 1)]
        self.passes = []
        for k in range(npass.value):
            lv, ps = C.c_int(), C.c_int()
            cnt = np.zeros(4, dtype=np.int32)
            _check(lib().fh_uamg_pass(self.h, k, C.byref(lv), C.byref(ps), cnt.ctypes.data, None, None))
            agg = np.full(max(cnt[0], 1), -1, dtype=np.int32)
            fc = np.full(max(cnt[1], 1), -1, dtype=np.int32)
            _check(lib().fh_uamg_pass(self.h, k, None, None, None, agg.ctypes.data, fc.ctypes.data))
            self.passes.append({"level": lv.value, "pass": ps.value, "elems": int(cnt[0]), "faces": int(cnt[1]),
                                "coarse_elems": int(cnt[2]), "coarse_faces": int(cnt[3]),
                                "elem_agg": agg[:cnt[0]], "face_coarse": fc[:cnt[1]]})

    def __del__(self):
        try:
            lib().fh_uamg_destroy(self.h)
        except Exception:
            pass


class HhoProblem:
    """Condensed HHO face system of -div(K grad u) = f on a structured mesh."""

    def __init__(self, dim=3, mesh="stetra", n=8, k=2, perturb=0.0, seed=1, source="pisine", kappa="one",
                 sigma=0.0, subdomain=8, aniso=1e-2, ref_n=0, source_seed=1, source_modes=8, defer_assembly=False):
        a = HhoArgs()
        self._strs = [mesh.encode(), source.encode(), kappa.encode()]
        a.dim, a.mesh, a.n, a.k, a.perturb, a.seed = dim, self._strs[0], n, k, perturb, seed
        a.source, a.kappa, a.sigma, a.subdomain, a.aniso = self._strs[1], self._strs[2], sigma, subdomain, aniso
        a.ref_n, a.source_seed, a.source_modes = ref_n, source_seed, source_modes
        a.defer_assembly = int(defer_assembly)
        self.args = dict(dim=dim, mesh=mesh, n=n, k=k, perturb=perturb, seed=seed, source=source, kappa=kappa,
                         sigma=sigma, subdomain=subdomain, aniso=aniso, ref_n=ref_n, source_seed=source_seed,
                         source_modes=source_modes, defer_assembly=defer_assembly)
        h = C.c_void_p()
        _check(lib().fh_hho_create(C.byref(a), C.byref(h)))
        self.h = h
        info = HhoInfo()
        _check(lib().fh_hho_get_info(self.h, C.byref(info)))
        self.info = {k: getattr(info, k) for k, _ in info._fields_}
        self.nb, self.nf, self.nnzb = info.nb, info.nf, info.nnzb
        self.n_unknowns = info.nb * info.nf

    @_ordered
    def assemble_gpu(self, ctx, keep_host=False):
        """Local operators + static condensation + assembly on the GPU:
        returns (device Matrix, device right-hand side tensor)."""
        import torch
        b = torch.zeros(self.n_unknowns, dtype=torch.float64, device="cuda")
        h = C.c_void_p()
        _check(lib().fh_hho_assemble_gpu(ctx.h, self.h, int(keep_host), C.byref(h), _ptr(b)))
        return Matrix(ctx, h), b

    def matrix(self, ctx, keep_host=True):
        h = C.c_void_p()
        _check(lib().fh_hho_matrix(ctx.h, self.h, int(keep_host), C.byref(h)))
        return Matrix(ctx, h)

    def rhs(self):
        b = np.zeros(self.n_unknowns)
        _check(lib().fh_hho_rhs(self.h, b))
        return b

    def blocks(self):
        rp = np.zeros(self.nb + 1, dtype=np.int32)
        ci = np.zeros(self.nnzb, dtype=np.int32)
        v = np.zeros(self.nnzb * self.nf * self.nf)
        _check(lib().fh_hho_blocks(self.h, rp, ci, v))
        return rp, ci, v.reshape(-1, self.nf, self.nf)

    def scipy(self):
        import scipy.sparse as sp
        rp, ci, v = self.blocks()
        return sp.bsr_matrix((v, ci, rp), shape=(self.n_unknowns, self.n_unknowns)).tocsr()

    def uncondensed(self):
        """(A_TT, A_TF, A_FF) as scipy CSR, and the cell unknowns per cell."""
        out = tuple(_csr_get(lib().fh_hho_uncondensed, self.h, w) for w in range(3))
        nc = C.c_int()
        _check(lib().fh_hho_cell_unknowns(self.h, C.byref(nc)))
        return out + (nc.value,)

    @_ordered
    def l2_error_gpu(self, ctx, x):
        """Relative L2 error of the reconstruction on the GPU (x: device
        FP64 tensor of the face unknowns)."""
        e = C.c_double()
        _check(lib().fh_hho_l2_error_gpu(ctx.h, self.h, _dev(x, self.n_unknowns, dtype=_f64(), name="x"),
                                         C.byref(e)))
        return e.value

    def l2_error(self, x):
        e = C.c_double()
        _check(lib().fh_hho_l2_error(self.h, np.ascontiguousarray(x, dtype=np.float64), C.byref(e)))
        return e.value

    def prolongation(self, ctx, coarse, keep_host=True):
        """HHO prolongation from `coarse` (mesh n/2) to this problem."""
        h = C.c_void_p()
        _check(lib().fh_hho_prolongation(ctx.h, self.h, coarse.h, int(keep_host), C.byref(h)))
        return Matrix(ctx, h)

    def coarsened(self):
        """The same problem on the once-coarsened mesh (same finest grid)."""
        kw = dict(self.args)
        kw["ref_n"] = kw["ref_n"] or kw["n"]
        kw["n"] = kw["n"] // 2
        return HhoProblem(**kw)

    def face_cells(self):
        out = np.zeros(2 * self.nb, dtype=np.int32)
        _check(lib().fh_hho_face_cells(self.h, out))
        return out.reshape(-1, 2)

    def cell_kappa(self):
        out = np.zeros(self.info["ncells"])
        _check(lib().fh_hho_cell_kappa(self.h, out))
        return out

    def __del__(self):
        try:
            lib().fh_hho_destroy(self.h)
        except Exception:
            pass


class Mpnn:
    """Learned post-smoother network (message passing over the face graph)."""

    def __init__(self, ctx, nf, layers=3, seed=1):
        h = C.c_void_p()
        _check(lib().fh_nn_create(ctx.h, nf, layers, seed, C.byref(h)))
        self.ctx, self.h, self.nf, self.layers = ctx, h, nf, layers

    def num_params(self):
        n = C.c_int64()
        _check(lib().fh_nn_num_params(self.h, C.byref(n)))
        return n.value

    def get_params(self):
        p = np.zeros(self.num_params(), dtype=np.float32)
        _check(lib().fh_nn_get_params(self.h, p))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params()
        _check(lib().fh_nn_set_params(self.h, p))

    @_ordered
    def forward(self, graph, r1, delta, bf16=True):
        n = graph.A.rows
        _check(lib().fh_nn_forward(self.ctx.h, self.h, graph.h, _dev(r1, n, dtype=_f64(), name="r1"),
                                   _dev(delta, n, dtype=_f64(), name="delta"), int(bf16)))

    @_ordered
    def loss_grad(self, graph, r1, e1, want_grad=True):
        loss = C.c_double()
        g = np.zeros(self.num_params(), dtype=np.float32) if want_grad else None
        n = graph.A.rows
        _check(lib().fh_nn_loss_grad(self.ctx.h, self.h, graph.h, _dev(r1, n, dtype=_f64(), name="r1"),
                                      _dev(e1, n, dtype=_f64(), name="e1"), C.byref(loss),
                                      g.ctypes.data if g is not None else None))
        return loss.value, g

    @_ordered
    def train_step(self, graph, r1, e1, lr):
        loss = C.c_double()
        n = graph.A.rows
        _check(lib().fh_nn_train_step(self.ctx.h, self.h, graph.h, _dev(r1, n, dtype=_f64(), name="r1"),
                                      _dev(e1, n, dtype=_f64(), name="e1"), lr, C.byref(loss)))
        return loss.value

    def __del__(self):
        try:
            if self.h:
                lib().fh_nn_destroy(self.h)
                self.h = None
        except Exception:
            pass


class NnGraph:
    """Face graph of an FP64 operator with 2 features per block row."""

    def __init__(self, ctx, A, features):
        f = np.ascontiguousarray(features, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().fh_nn_graph_create(ctx.h, A.h, f.ctypes.data, C.byref(h)))
        self.ctx, self.h, self.A, self._f = ctx, h, A, f

    def __del__(self):
        try:
            if self.h:
                lib().fh_nn_graph_destroy(self.h)
                self.h = None
        except Exception:
            pass


def face_features(problem):
    """Mean and half-difference of log kappa of the two cells of every
    block row (face unknown) of an HHO problem."""
    fc = problem.face_cells()
    lk = np.log(problem.cell_kappa())
    c0, c1 = fc[:, 0], np.where(fc[:, 1] >= 0, fc[:, 1], fc[:, 0])
    return np.stack([(lk[c0] + lk[c1]) / 2, np.abs(lk[c0] - lk[c1]) / 2], axis=1)


def profile_begin(ctx, A, mode=0):
    """Time every launch of `mode` (0 spmv, 1 residual, 2 smoothing, 3 fused
    Chebyshev step) on A."""
    _check(lib().fh_profile_begin(ctx.h, A.h, mode))


def profile_end(ctx):
    n, ms = C.c_int(), C.c_double()
    _check(lib().fh_profile_end(ctx.h, C.byref(n), C.byref(ms)))
    return n.value, ms.value
