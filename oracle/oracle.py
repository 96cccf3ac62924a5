"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
bench.py as the checker; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class LevelDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("rp", C.c_void_p), ("ci", C.c_void_p), ("v", C.c_void_p),
                ("nc", C.c_int), ("prp", C.c_void_p), ("pci", C.c_void_p), ("pv", C.c_void_p),
                ("nf", C.c_int), ("pmodes", C.c_int)]


class SolverCfg(C.Structure):
    _fields_ = [("krylov", C.c_int), ("prec", C.c_int), ("nf", C.c_int), ("omega", C.c_double),
                ("pre_kind", C.c_int), ("post_kind", C.c_int), ("pre_omega", C.c_double),
                ("post_omega", C.c_double), ("nu1", C.c_int), ("nu2", C.c_int), ("cycle", C.c_char),
                ("wloops", C.c_int), ("max_coarse", C.c_int), ("theta", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_int), ("nlevels", C.c_int), ("coarse_kind", C.c_int), ("coarse_omega", C.c_double),
                ("fcg_truncation", C.c_int), ("fcg_restart", C.c_int), ("kcycle_levels", C.c_int),
                ("x_given", C.c_int), ("nu_change_op", C.c_char), ("nu_change_coeff", C.c_int),
                ("coarse_nu_given", C.c_int), ("coarse_nu1", C.c_int), ("coarse_nu2", C.c_int)]


class SolverOut(C.Structure):
    _fields_ = [("iterations", C.c_int), ("nlevels", C.c_int), ("final_res", C.c_double),
                ("level_n", C.c_int * 32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_spmv.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.or_block_inverse.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p]
        L.or_bj_sweep.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, C.c_double, _f64p, _f64p]
        L.or_bgs_sweep.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, C.c_int, _f64p, _f64p]
        L.or_mcbgs_sweep.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, C.c_int, _i32p, _i32p,
                                     C.c_int, _f64p, _f64p]
        L.or_bsor_sweep.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, C.c_int, C.c_double, _f64p, _f64p]
        L.or_mcsor_sweep.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, C.c_int, _i32p, _i32p,
                                     C.c_int, C.c_double, _f64p, _f64p]
        L.or_greedy_colouring.argtypes = [C.c_int, _i32p, _i32p, _i32p]
        L.or_pairwise_aggregate.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_double, _i32p,
                                            C.POINTER(C.c_int)]
        L.or_amg_hierarchy.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_double, C.c_int, _i32p,
                                       _i32p, C.c_int]
        L.or_solve.argtypes = [C.POINTER(SolverCfg), C.POINTER(LevelDesc), _f64p, _f64p, _f64p,
                               C.POINTER(SolverOut)]
        L.or_last_times.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_mpnn_num_params.argtypes = [C.c_int, C.c_int]
        L.or_mpnn_forward.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, C.c_int, _f64p, _f64p, _f64p]
        L.or_mpnn_loss_grad.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, C.c_int, _f64p, _f64p, _f64p,
                                        C.c_void_p]
        L.or_mpnn_loss_grad.restype = C.c_double
        _lib = L
    return _lib


def _csr(A):
    A = A.tocsr()
    A.sort_indices()
    return (np.ascontiguousarray(A.indptr, dtype=np.int32), np.ascontiguousarray(A.indices, dtype=np.int32),
            np.ascontiguousarray(A.data, dtype=np.float64))


def spmv(A, x):
    rp, ci, v = _csr(A)
    y = np.zeros(A.shape[0])
    lib().or_spmv(A.shape[0], rp, ci, v, np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def block_inverse(A, nf):
    rp, ci, v = _csr(A)
    inv = np.zeros((A.shape[0] // nf) * nf * nf)
    rc = lib().or_block_inverse(A.shape[0], nf, rp, ci, v, inv)
    if rc:
        raise ValueError("singular diagonal block %d" % (-rc - 1))
    return inv


def bj_sweep(A, nf, invD, omega, b, x):
    rp, ci, v = _csr(A)
    x = np.array(x, dtype=np.float64)
    lib().or_bj_sweep(A.shape[0], nf, rp, ci, v, invD, omega, np.ascontiguousarray(b, dtype=np.float64), x)
    return x


def bgs_sweep(A, nf, invD, direction, b, x):
    rp, ci, v = _csr(A)
    x = np.array(x, dtype=np.float64)
    lib().or_bgs_sweep(A.shape[0], nf, rp, ci, v, invD, direction, np.ascontiguousarray(b, dtype=np.float64), x)
    return x


def bsor_sweep(A, nf, invD, direction, omega, b, x):
    rp, ci, v = _csr(A)
    x = np.array(x, dtype=np.float64)
    lib().or_bsor_sweep(A.shape[0], nf, rp, ci, v, invD, direction, omega, np.ascontiguousarray(b, dtype=np.float64),
                        x)
    return x


def mcsor_sweep(A, nf, invD, colour_ptr, colour_rows, direction, omega, b, x):
    rp, ci, v = _csr(A)
    x = np.array(x, dtype=np.float64)
    lib().or_mcsor_sweep(A.shape[0], nf, rp, ci, v, invD, len(colour_ptr) - 1,
                         np.ascontiguousarray(colour_ptr, dtype=np.int32),
                         np.ascontiguousarray(colour_rows, dtype=np.int32), direction, omega,
                         np.ascontiguousarray(b, dtype=np.float64), x)
    return x


def mcbgs_sweep(A, nf, invD, colour_ptr, colour_rows, direction, b, x):
    rp, ci, v = _csr(A)
    x = np.array(x, dtype=np.float64)
    lib().or_mcbgs_sweep(A.shape[0], nf, rp, ci, v, invD, len(colour_ptr) - 1,
                         np.ascontiguousarray(colour_ptr, dtype=np.int32),
                         np.ascontiguousarray(colour_rows, dtype=np.int32), direction,
                         np.ascontiguousarray(b, dtype=np.float64), x)
    return x


def greedy_colouring(brp, bci):
    nb = len(brp) - 1
    col = np.zeros(nb, dtype=np.int32)
    nc = lib().or_greedy_colouring(nb, np.ascontiguousarray(brp, dtype=np.int32),
                                   np.ascontiguousarray(bci, dtype=np.int32), col)
    return nc, col


def pairwise_aggregate(A, nf, theta=0.25):
    rp, ci, v = _csr(A)
    agg = np.zeros(A.shape[0] // nf, dtype=np.int32)
    coarsest = C.c_int(0)
    na = lib().or_pairwise_aggregate(A.shape[0], nf, rp, ci, v, theta, agg, C.byref(coarsest))
    return na, agg, bool(coarsest.value)


def amg_hierarchy(A, nf, theta=0.25, max_coarse=1000):
    rp, ci, v = _csr(A)
    agg0 = np.zeros(A.shape[0] // nf, dtype=np.int32)
    ln = np.zeros(32, dtype=np.int32)
    nl = lib().or_amg_hierarchy(A.shape[0], nf, rp, ci, v, theta, max_coarse, agg0, ln, 32)
    return [int(t) for t in ln[:nl]], agg0


SMOOTHER_KIND = {"bj": 1, "bj23": 1, "bgs": 2, "rbgs": 3, "sbgs": 4, "gs": 2, "rgs": 3, "sgs": 4,
                 "abgs": 12, "rabgs": 13, "ags": 12, "rags": 13,
                 # block SOR: the Gauss-Seidel kinds with the relaxation `omega`
                 "bsor": 2, "rbsor": 3, "sbsor": 4, "absor": 12, "rabsor": 13,
                 # multicolour block Gauss-Seidel / SOR (greedy colouring), the ordering of the HIP path
                 "mcbgs": 5, "mcrbgs": 6, "mcsbgs": 7, "mcabgs": 14, "mcrabgs": 15,
                 "cheb2": 8, "cheb3": 9, "cheb4": 10, "cheb1": 11}
PREC_KIND = {"": 0, "none": 0, "bj": 1, "bgs": 2, "mg": 3, "aggregamg": 4}


def solve(A, b, solver="cg", prec="mg", nf=1, cycle="V", nu=(1, 1), smoothers=("bj", "bj"), omega=1.0,
          levels=None, tol=1e-8, max_iter=200, max_coarse=1000, theta=0.25, coarse_smoother=None, coarse_omega=0.0,
          fcg_truncation=1, fcg_restart=False, kcycle_levels=0, x0=None, nu_change=None, coarse_nu=None):
    """Runs the oracle solver; `levels` = [(A_l, P_l), ...] for GMG (P of the
    coarsest level is None).  A level may be (A_l, ("p", nfc), nf_l): a
    p-transfer keeping nfc modes of every nf_l-block; (A_l, P_l, nf_l) gives
    the block size of a level whose size differs from nf.  x0: initial guess
    (default zero).  Returns (x, hist, out)."""
    keep = []
    if levels is None:
        levels = [(A, None)]
    descs = (LevelDesc * len(levels))()
    for l, lev in enumerate(levels):
        Al, Pl = lev[0], lev[1]
        descs[l].nf = lev[2] if len(lev) > 2 else 0
        rp, ci, v = _csr(Al)
        keep += [rp, ci, v]
        descs[l].n = Al.shape[0]
        descs[l].rp, descs[l].ci, descs[l].v = rp.ctypes.data, ci.ctypes.data, v.ctypes.data
        if isinstance(Pl, tuple) and Pl[0] == "p":
            descs[l].pmodes = Pl[1]
        elif Pl is not None:
            prp, pci, pv = _csr(Pl)
            keep += [prp, pci, pv]
            descs[l].nc = Pl.shape[1]
            descs[l].prp, descs[l].pci, descs[l].pv = prp.ctypes.data, pci.ctypes.data, pv.ctypes.data
    cfg = SolverCfg()
    cfg.krylov = {"none": 0, "cg": 1, "fcg": 2}[solver]
    cfg.prec = PREC_KIND[prec]
    cfg.nf = nf
    cfg.omega = 2.0 / 3.0 if prec == "bj23" else omega
    cfg.pre_kind = SMOOTHER_KIND[smoothers[0]]
    cfg.post_kind = SMOOTHER_KIND[smoothers[1]]
    cfg.pre_omega = 2.0 / 3.0 if smoothers[0] == "bj23" else omega
    cfg.post_omega = 2.0 / 3.0 if smoothers[1] == "bj23" else omega
    cfg.nu1, cfg.nu2 = nu
    cfg.fcg_truncation, cfg.fcg_restart = fcg_truncation, int(fcg_restart)
    cfg.kcycle_levels = kcycle_levels
    if nu_change:  # e.g. "+1" (Multigrid.h:224-242)
        cfg.nu_change_op, cfg.nu_change_coeff = nu_change[0].encode(), int(nu_change[1:])
    if coarse_nu is not None:  # (pre, post) smoothing iterations below the finest level
        cfg.coarse_nu_given, (cfg.coarse_nu1, cfg.coarse_nu2) = 1, coarse_nu
    if coarse_smoother:
        cfg.coarse_kind = SMOOTHER_KIND[coarse_smoother]
        cfg.coarse_omega = 2.0 / 3.0 if coarse_smoother == "bj23" else (coarse_omega if coarse_omega > 0 else omega)
    cfg.cycle = cycle.encode()
    cfg.wloops = 2 if cycle == "W" else 1
    cfg.max_coarse = max_coarse
    cfg.theta = theta
    cfg.tol = tol
    cfg.max_iter = max_iter
    cfg.nlevels = len(levels)
    x = np.zeros(A.shape[0])
    if x0 is not None:
        x[:] = x0
        cfg.x_given = 1
    hist = np.zeros(max_iter + 2)
    out = SolverOut()
    st = lib().or_solve(C.byref(cfg), descs, np.ascontiguousarray(b, dtype=np.float64), x, hist, C.byref(out))
    if st < 0:
        raise RuntimeError("oracle solve failed with status %d" % st)
    return x, hist[: out.iterations + 1], out


def solve_bcsr(rowptr, colidx, vals, nf, b, solver="fcg", prec="aggregamg", cycle="K", nu=(1, 1),
               smoothers=("cheb2", "cheb2"), tol=1e-8, max_iter=200, max_coarse=1000, theta=0.25,
               coarse_smoother=None, kcycle_levels=0, fcg_truncation=1, coarse_nu=None):
    """Oracle solve with the matrix in block storage (block row pointers,
    block columns, nnzb x nf x nf values): no preconditioner or the p-first
    aggregation AMG with Chebyshev smoothing.  Returns (x, hist, out)."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(colidx, dtype=np.int32)
    v = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1)
    nb = len(rp) - 1
    cfg = SolverCfg()
    cfg.krylov = {"none": 0, "cg": 1, "fcg": 2}[solver]
    cfg.prec = PREC_KIND[prec]
    cfg.nf = nf
    cfg.omega = 1.0
    cfg.pre_kind = SMOOTHER_KIND[smoothers[0]]
    cfg.post_kind = SMOOTHER_KIND[smoothers[1]]
    cfg.pre_omega = cfg.post_omega = 1.0
    cfg.nu1, cfg.nu2 = nu
    cfg.fcg_truncation, cfg.fcg_restart = fcg_truncation, 0
    if coarse_nu is not None:
        cfg.coarse_nu_given, (cfg.coarse_nu1, cfg.coarse_nu2) = 1, coarse_nu
    if coarse_smoother:
        cfg.coarse_kind = SMOOTHER_KIND[coarse_smoother]
        cfg.coarse_omega = 1.0
    cfg.cycle = cycle.encode()
    cfg.wloops = 2 if cycle == "W" else 1
    cfg.max_coarse = max_coarse
    cfg.theta = theta
    cfg.tol = tol
    cfg.max_iter = max_iter
    cfg.nlevels = 0
    cfg.kcycle_levels = kcycle_levels
    x = np.zeros(nb * nf)
    hist = np.zeros(max_iter + 2)
    out = SolverOut()
    L = lib()
    L.or_solve_bcsr.argtypes = [C.POINTER(SolverCfg), C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                C.POINTER(SolverOut)]
    st = L.or_solve_bcsr(C.byref(cfg), nb, rp, ci, v, np.ascontiguousarray(b, dtype=np.float64), x, hist,
                         C.byref(out))
    if st < 0:
        raise RuntimeError("oracle solve failed with status %d" % st)
    return x, hist[: out.iterations + 1], out


def last_times():
    s, t = C.c_double(), C.c_double()
    lib().or_last_times(C.byref(s), C.byref(t))
    return s.value, t.value


def _bsr(A, nf):
    import scipy.sparse as sp
    B = sp.bsr_matrix(A, blocksize=(nf, nf))
    B.sort_indices()
    return (np.ascontiguousarray(B.indptr, dtype=np.int32), np.ascontiguousarray(B.indices, dtype=np.int32),
            np.ascontiguousarray(B.data, dtype=np.float64).reshape(-1))


def mpnn_num_params(nf, layers=3):
    return lib().or_mpnn_num_params(nf, layers)


def mpnn_forward(A, nf, feat, params, r1, layers=3):
    """Learned post-smoother correction delta of r1 (FP64)."""
    rp, ci, v = _bsr(A, nf)
    delta = np.zeros(A.shape[0])
    lib().or_mpnn_forward(len(rp) - 1, nf, rp, ci, v, np.ascontiguousarray(feat, dtype=np.float64).reshape(-1),
                          layers, np.ascontiguousarray(params, dtype=np.float64), np.ascontiguousarray(r1, dtype=np.float64),
                          delta)
    return delta


def mpnn_loss_grad(A, nf, feat, params, r1, e1, layers=3, want_grad=True):
    """Relative energy error of the corrected vector and its gradient."""
    rp, ci, v = _bsr(A, nf)
    g = np.zeros(len(params)) if want_grad else None
    loss = lib().or_mpnn_loss_grad(len(rp) - 1, nf, rp, ci, v,
                                   np.ascontiguousarray(feat, dtype=np.float64).reshape(-1), layers,
                                   np.ascontiguousarray(params, dtype=np.float64),
                                   np.ascontiguousarray(r1, dtype=np.float64), np.ascontiguousarray(e1, dtype=np.float64),
                                   g.ctypes.data if g is not None else None)
    return loss, g
