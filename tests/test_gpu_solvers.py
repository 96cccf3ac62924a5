"""End-to-end parity of the GPU solvers with the reference's own solver on
the same exported systems (golden fixtures) and with the CPU oracle.

Block-Jacobi-smoothed cycles must reproduce the reference's iteration count
exactly and its FP64 solution to 1e-10 relative.  Gauss-Seidel runs as a
multicolour sweep on the GPU, which changes the sweep order: those cases are
checked against the oracle run with the same multicolour ordering
(iterations equal, solution to 1e-10), and the reference's lexicographic
count is stated beside it.
"""
import numpy as np
import pytest

from conftest import csr, gmg_levels, load
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import fhhos4_amd as F  # noqa: E402

G = F.fhgpu


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "no ROCm device visible"
    return G.Context(0)


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    return t


def run_gpu(ctx, d, nf, levels=None, **kw):
    A = csr(d, "A")
    M = G.Matrix.from_csr(ctx, A, nf, keep_host=True)
    s = G.Solver(ctx, **kw)
    keep = [M]
    if levels:
        for l, (Al, Pl) in enumerate(levels):
            Ml = M if l == 0 else G.Matrix.from_csr(ctx, Al, nf)
            Pm = G.Matrix.from_csr(ctx, Pl, nf) if Pl is not None else None
            keep += [Ml, Pm]
            s.set_level(l, Ml if l > 0 else None, Pm)
    s.setup(M)
    b = dev(d["b"])
    x = torch.zeros_like(b)
    st = s.solve(b, x)
    ctx.sync()
    return x.cpu().numpy(), s.history(), st


BJ_CASES = [
    # fixture (reference run), system fixture, kwargs, nf
    ("c1_n64_cg_mg_bj", "c1_n64_cg_mg_bj", dict(solver="cg", prec="mg", smoothers=("bj", "bj")), 2),
    ("c1_n64_cg_mg_bj23", "c1_n64_cg_mg_bj", dict(solver="cg", prec="mg", smoothers=("bj23", "bj23")), 2),
    ("c1_n32_cg_mg_bj23", "c1_n32_cg_mg_bj23", dict(solver="cg", prec="mg", smoothers=("bj23", "bj23")), 2),
    ("c1_n32_cg_mg_bj", "c1_n32_cg_mg_bj23", dict(solver="cg", prec="mg", smoothers=("bj", "bj"), max_iter=15), 2),
    ("c1_n32_cg_bj", "c1_n32_cg_mg_bj23", dict(solver="cg", prec="bj"), 2),
    ("c1_n32_cg_none", "c1_n32_cg_mg_bj23", dict(solver="cg", prec=""), 2),
    ("c1_n32_fcg_mg_bj23", "c1_n32_cg_mg_bj23", dict(solver="fcg", prec="mg", smoothers=("bj23", "bj23")), 2),
    ("c1_n32_cg_mg_w_bj23", "c1_n32_cg_mg_bj23", dict(solver="cg", prec="mg", cycle="W", smoothers=("bj23", "bj23")), 2),
    ("cart8_k2_cg_mg_bj23", "cart8_k2_cg_mg_bj23", dict(solver="cg", prec="mg", smoothers=("bj23", "bj23")), 6),
    ("tet4_k2", "tet4_k2", dict(solver="cg", prec="bj"), 6),
]


@pytest.mark.parametrize("case", BJ_CASES, ids=[c[0] for c in BJ_CASES])
def test_solver_matches_reference(ctx, case):
    name, sysname, kw, nf = case
    print("case", name, flush=True)
    f, d = load(name), load(sysname)
    levels = gmg_levels(d) if kw.get("prec") == "mg" else None
    x, hist, st = run_gpu(ctx, d, nf, levels=levels, **kw)
    assert st["iterations"] == int(f["info_iterations"])
    assert np.linalg.norm(x - f["x"]) <= 1e-10 * np.linalg.norm(f["x"])
    k = min(6, len(hist))
    np.testing.assert_allclose(hist[:k], f["hist"][:k], rtol=1e-8)


GS_CASES = [
    ("c1_n32_cg_mg_bgs", "c1_n32_cg_mg_bj23", "cg", "mg", "V", 2),
    ("c1_n32_fcg_mg_k_bgs", "c1_n32_cg_mg_bj23", "fcg", "mg", "K", 2),
    ("tet8_k0_fcg_amg", "tet8_k0_fcg_amg", "fcg", "aggregamg", "K", 1),
    ("tet8_k0_cg_amg_v", "tet8_k0_fcg_amg", "cg", "aggregamg", "V", 1),
]


@pytest.mark.parametrize("case", GS_CASES, ids=[c[0] for c in GS_CASES])
def test_gauss_seidel_solver_matches_oracle_multicolour(ctx, case):
    name, sysname, solver, prec, cycle, nf = case
    print("case", name, flush=True)
    f, d = load(name), load(sysname)
    A = csr(d, "A")
    levels = gmg_levels(d) if prec == "mg" else None
    maxit = 25 if name == "tet8_k0_cg_amg_v" else 200
    x, hist, st = run_gpu(ctx, d, nf, levels=levels, solver=solver, prec=prec, cycle=cycle,
                          smoothers=("bgs", "rbgs"), max_iter=maxit)
    xo, ho, oo = O.solve(A, d["b"], solver, prec, nf=nf, cycle=cycle, smoothers=("mcbgs", "mcrbgs"),
                         levels=levels, max_iter=maxit)
    assert st["iterations"] == oo.iterations
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    # stated: reference (lexicographic) count and distance to its solution
    print("reference lexicographic iterations", int(f["info_iterations"]), "multicolour", st["iterations"],
          "rel. distance to reference solution", np.linalg.norm(x - f["x"]) / np.linalg.norm(f["x"]))
    assert abs(st["iterations"] - int(f["info_iterations"])) <= 3


MC_CASES = [
    # fixture (reference run), system, solver, prec, cycle, smoothers, oracle's multicolour kinds, nu, relax,
    # per-level smoothing change
    ("c1_n32_cg_mg_bsor", "c1_n32_cg_mg_bj23", "cg", "mg", "V", ("bsor", "rbsor"), ("mcbgs", "mcrbgs"), (1, 1), 1.2,
     None, 2),
    ("c1_n32_cg_mg_abgs", "c1_n32_cg_mg_bj23", "cg", "mg", "V", ("abgs", "rabgs"), ("mcabgs", "mcrabgs"), (2, 2),
     1.0, None, 2),
    ("tet8_k0_fcg_amg_nu", "tet8_k0_fcg_amg", "fcg", "aggregamg", "K", ("bgs", "rbgs"), ("mcbgs", "mcrbgs"), (1, 1),
     1.0, "+1", 1),
]


@pytest.mark.parametrize("case", MC_CASES, ids=[c[0] for c in MC_CASES])
def test_sor_alternating_and_level_smoothing_match_oracle_multicolour(ctx, case):
    """Block SOR (relaxation 1.2), alternating Gauss-Seidel and the per-level
    change of the smoothing iterations (Multigrid.h:224-242) on the GPU
    against the oracle run with the same multicolour ordering: iterations
    equal, the solution to 1e-10; the reference's lexicographic count is
    stated beside it."""
    name, sysname, solver, prec, cycle, sm, osm, nu, relax, nuc, nf = case
    print("case", name, flush=True)
    f, d = load(name), load(sysname)
    A = csr(d, "A")
    levels = gmg_levels(d) if prec == "mg" else None
    x, hist, st = run_gpu(ctx, d, nf, levels=levels, solver=solver, prec=prec, cycle=cycle, nu=nu, smoothers=sm,
                          relax=relax, nu_change=nuc)
    xo, ho, oo = O.solve(A, d["b"], solver, prec, nf=nf, cycle=cycle, nu=nu, smoothers=osm, omega=relax,
                         levels=levels, nu_change=nuc)
    assert st["iterations"] == oo.iterations
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    print("reference lexicographic iterations", int(f["info_iterations"]), "multicolour", st["iterations"])
    assert abs(st["iterations"] - int(f["info_iterations"])) <= 3


def test_amg_hierarchy_matches_reference_levels(ctx):
    d = load("tet8_k0_fcg_amg")
    A = csr(d, "A")
    M = G.Matrix.from_csr(ctx, A, 1, keep_host=True)
    s = G.Solver(ctx, solver="fcg", prec="aggregamg", cycle="K", smoothers=("bgs", "rbgs"))
    s.setup(M)
    sizes = [s.level_size(l) for l in range(int(d["nlevels"]))]
    ref = [A.shape[0]] + [csr(d, "level%d_A" % l).shape[0] for l in range(1, int(d["nlevels"]))]
    assert sizes == ref


@pytest.mark.parametrize("mg_prec,mg_store,coarse_store", [("f32", "f32", None), ("f32", "bf16", None),
                                                           ("f32", "f32", "bf16")])
def test_mixed_precision_fcg_reaches_fp64_tolerance(ctx, mg_prec, mg_store, coarse_store):
    """FP64 FCG with an FP32 (or BF16-stored, or BF16 on the coarse levels
    only) multigrid preconditioner converges to the same 1e-8 relative
    residual; its solution is compared with the FP64 oracle solve."""
    d = load("c1_n64_cg_mg_bj")
    A = csr(d, "A")
    x, hist, st = run_gpu(ctx, d, 2, levels=gmg_levels(d), solver="fcg", prec="mg", smoothers=("bj23", "bj23"),
                          mg_prec=mg_prec, mg_store=mg_store, coarse_store=coarse_store)
    assert st["converged"] == 1 and hist[-1] < 1e-8
    xo, _, oo = O.solve(A, d["b"], "fcg", "mg", nf=2, smoothers=("bj23", "bj23"), levels=gmg_levels(d))
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print("mixed precision", mg_prec, mg_store, coarse_store, "iterations", st["iterations"], "vs FP64", oo.iterations,
          "rel. error vs FP64 oracle", err)
    assert err < 1e-7
    assert st["iterations"] <= oo.iterations + 3


def test_exported_levels_match_reference(ctx, tmp_path):
    """-export mg: level<l>_A/P/R.dat of the GPU solver's hierarchy equal the
    reference's exported level matrices (geometric levels given from the
    reference; aggregation levels built here, bit-exact indices)."""
    import scipy.io as sio
    # geometric hierarchy (operators and prolongations are the reference's own)
    d = load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    M = G.Matrix.from_csr(ctx, A, 2, keep_host=True)
    s = G.Solver(ctx, solver="cg", prec="mg", smoothers=("bj23", "bj23"))
    keep = []
    for l, (Al, Pl) in enumerate(gmg_levels(d)):
        Ml = G.Matrix.from_csr(ctx, Al, 2) if l > 0 else None
        Pm = G.Matrix.from_csr(ctx, Pl, 2) if Pl is not None else None
        keep += [Ml, Pm]
        s.set_level(l, Ml, Pm)
    s.setup(M)
    s.export_levels(str(tmp_path))
    for l in range(int(d["nlevels"])):
        got = sio.mmread(str(tmp_path / ("level%d_A.dat" % l))).tocsr()
        ref = A if l == 0 else csr(d, "level%d_A" % l)
        assert abs(got - ref).max() == 0.0
        if ("level%d_P_shape" % l) in d:
            P = sio.mmread(str(tmp_path / ("level%d_P.dat" % l))).tocsr()
            R = sio.mmread(str(tmp_path / ("level%d_R.dat" % l))).tocsr()
            assert abs(P - csr(d, "level%d_P" % l)).max() == 0.0
            assert abs(R - csr(d, "level%d_P" % l).T).max() == 0.0
    # aggregation AMG (scalar): every coarse operator equals the reference's
    d = load("tet8_k0_fcg_amg")
    A = csr(d, "A")
    M = G.Matrix.from_csr(ctx, A, 1, keep_host=True)
    s = G.Solver(ctx, solver="fcg", prec="aggregamg", cycle="K", smoothers=("bgs", "rbgs"))
    s.setup(M)
    out = tmp_path / "amg"
    out.mkdir()
    s.export_levels(str(out))
    for l in range(1, int(d["nlevels"])):
        got = sio.mmread(str(out / ("level%d_A.dat" % l))).tocsr()
        ref = csr(d, "level%d_A" % l)
        assert got.shape == ref.shape
        assert abs(got - ref).max() <= 1e-12 * abs(ref).max()
        P = sio.mmread(str(out / ("level%d_P.dat" % (l - 1)))).tocsr()
        Af = sio.mmread(str(out / ("level%d_A.dat" % (l - 1)))).tocsr()
        assert abs(P.T @ Af @ P - got).max() <= 1e-12 * abs(got).max()


def test_stopping_rules_match_the_reference(ctx):
    """IterativeSolver::StoppingCriteriaReached (IterativeSolver.h:224-247):
    max_iter = 0 stops before the first iteration, leaving the zero initial
    guess; a zero right-hand side makes CG's first step 0/0 and the NaN
    residual is reported as an error, as the reference's FatalError ("the
    residual is NaN.").  The oracle behaves the same (test_oracle.py)."""
    d = load("c1_n32_cg_mg_bj23")
    M = G.Matrix.from_csr(ctx, csr(d, "A"), 2)
    s = G.Solver(ctx, solver="cg", prec="bj", tol=1e-8, max_iter=200)
    s.setup(M)
    b = dev(d["b"])
    x = torch.full_like(b, 3.0)
    s.set_stop(1e-8, 0)
    st = s.solve(b, x)
    ctx.sync()
    assert st["iterations"] == 0 and st["converged"] == 0
    assert not x.cpu().numpy().any()
    s.set_stop(1e-8, 200)
    with pytest.raises(G.FhError, match="NaN"):
        s.solve(torch.zeros_like(b), x)
    st = s.solve(b, x)  # the solver stays usable after the error
    assert st["converged"] == 1


def test_graph_recaptured_after_layout_change(ctx):
    """A multigrid cycle replayed from a HIP graph reads the sliced copy of
    the fine operator; dropping that copy (fh_mat_set_sliced(A, 0)) must make
    the next solve capture again on block CSR, with the same result."""
    from fhhos4_amd import gmg
    P = G.HhoProblem(dim=3, mesh="stetra", n=12, k=1, perturb=0.1, seed=3, source="pisine")
    hier = gmg.build_hierarchy(ctx, P)
    A = hier[0][1]
    s = G.Solver(ctx, solver="cg", prec="mg", nu=(1, 1), smoothers=("cheb2", "cheb2"), galerkin=True,
                 coarse_smoother="cheb1", max_iter=200)
    gmg.attach(s, hier)
    s.setup(A)
    b = dev(P.rhs())
    x = torch.zeros_like(b)
    st0 = s.solve(b, x)   # decides the layout, captures and replays the cycle
    x0 = x.cpu().numpy()
    assert A.layout() in (1, 2), A.layout()
    s.solve(b, x)
    assert not A.set_sliced(False)
    assert A.layout() == 0
    st1 = s.solve(b, x)   # same (r, z) pair: must not replay the stale graph
    x1 = x.cpu().numpy()
    assert st1["iterations"] == st0["iterations"]
    assert np.linalg.norm(x1 - x0) / np.linalg.norm(x0) < 1e-12
