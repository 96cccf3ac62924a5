"""Geometric multigrid on the library's HHO systems (3-D tetrahedra with
perturbed vertices, k = 2): the GPU solver against the CPU oracle run on the
same hierarchy (re-discretised or Galerkin coarse operators, the same
multicolour Gauss-Seidel ordering), and repeated solves through the captured
cycle."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
F = pytest.importorskip("fhhos4_amd")
G = F.fhgpu
from fhhos4_amd import gmg  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "no ROCm device visible"
    return G.Context(0)


@pytest.fixture(scope="module")
def system(ctx):
    P = G.HhoProblem(dim=3, mesh="stetra", n=8, k=2, perturb=0.2, seed=2026, source="pisine")
    H = gmg.build_hierarchy(ctx, P)
    hctx = G.Context(-1)
    Hh = gmg.build_hierarchy(hctx, P)
    gal = [Hh[0][1]]
    for l in range(len(Hh) - 1):
        gal.append(G.galerkin_rap(hctx, gal[-1], Hh[l][2]))
    return P, H, Hh, gal


CASES = [
    dict(galerkin=True, smoothers=("sbgs", "sbgs"), relax=1.0),
    dict(galerkin=True, smoothers=("sbgs", "sbgs"), relax=1.0, coarse_smoother="bj", coarse_relax=0.5),
    dict(galerkin=True, smoothers=("sbgs", "sbgs"), relax=1.0, coarse_smoother="cheb2"),
    dict(galerkin=True, smoothers=("cheb3", "cheb3"), relax=1.0),
    dict(galerkin=True, smoothers=("cheb2", "cheb2"), relax=1.0, coarse_smoother="cheb1"),
    dict(galerkin=False, smoothers=("bj", "bj"), relax=0.5),
    dict(galerkin=False, smoothers=("sbgs", "sbgs"), relax=1.0, cycle="W"),
    dict(galerkin=True, smoothers=("sbgs", "sbgs"), relax=1.0, solver="fcg", fcg_truncation=4),
]


@pytest.mark.parametrize("case", CASES, ids=["galerkin-sbgs", "galerkin-sbgs-coarse-bj", "galerkin-sbgs-coarse-cheb2",
                                             "galerkin-cheb3", "galerkin-cheb2-coarse-cheb1", "rediscr-bj", "rediscr-sbgs-W", "galerkin-sbgs-fcg4"])
def test_gmg_matches_oracle(ctx, system, case):
    P, H, Hh, gal = system
    cycle = case.get("cycle", "V")
    krylov = case.get("solver", "cg")
    m = case.get("fcg_truncation", 1)
    s = G.Solver(ctx, solver=krylov, prec="mg", cycle=cycle, nu=(1, 1), smoothers=case["smoothers"],
                 relax=case["relax"], tol=1e-8, max_iter=300, galerkin=case["galerkin"],
                 coarse_smoother=case.get("coarse_smoother"), coarse_relax=case.get("coarse_relax", 0.0),
                 fcg_truncation=m)
    gmg.attach(s, H)
    s.setup(H[0][1])
    b = torch.from_numpy(P.rhs()).cuda()
    x = torch.zeros_like(b)
    st = s.solve(b, x)
    ops = gal if case["galerkin"] else [h[1] for h in Hh]
    levels = [(ops[l].scipy(), Hh[l][2].scipy() if Hh[l][2] is not None else None) for l in range(len(Hh))]
    sm = tuple("mc" + k if "gs" in k else k for k in case["smoothers"])
    xo, hist, out = O.solve(levels[0][0], P.rhs(), krylov, "mg", nf=P.nf, cycle=cycle, smoothers=sm,
                            omega=case["relax"], nu=(1, 1), levels=levels, max_iter=300,
                            coarse_smoother=case.get("coarse_smoother"), coarse_omega=case.get("coarse_relax", 0.0),
                            fcg_truncation=m)
    print("%s: GPU %d iterations, oracle %d" % (case, st["iterations"], out.iterations))
    assert st["converged"] == 1
    assert st["iterations"] == out.iterations
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)
    # the first iterations agree to rounding; later ones drift as CG amplifies
    # rounding differences (different summation orders), while the count and
    # the solution above stay equal
    np.testing.assert_allclose(s.history()[:12], hist[:12], rtol=1e-6, atol=1e-14)


ASSEMBLY_CASES = [
    dict(dim=3, mesh="stetra", n=6, k=2, perturb=0.2, seed=5),
    dict(dim=3, mesh="stetra", n=4, k=0, source="sine"),
    dict(dim=3, mesh="stetra", n=4, k=3, kappa="lognormal", sigma=1.0),
    dict(dim=2, mesh="stri", n=8, k=2, perturb=0.2, kappa="lognormal", sigma=2.0, source="modes", source_seed=7),
    dict(dim=3, mesh="cart", n=4, k=2, kappa="subdomain", sigma=3.0, subdomain=2),
    dict(dim=2, mesh="cart", n=8, k=3),
    dict(dim=3, mesh="hex", n=4, k=3, perturb=0.15, kappa="lognormal", sigma=3.0, seed=11),
    dict(dim=3, mesh="hex", n=5, k=1, perturb=0.15, seed=12, source="modes", source_seed=3),
    dict(dim=3, mesh="hex", n=4, k=0, perturb=0.15, seed=13),
    dict(dim=2, mesh="quad", n=8, k=2, perturb=0.3, kappa="lognormal", sigma=2.0, seed=14),
]


@pytest.mark.parametrize("kw", ASSEMBLY_CASES, ids=["tet-k2-perturbed", "tet-k0", "tet-k3-lognormal",
                                                    "tri-k2-modes", "hex-k2-subdomains", "quad-k3",
                                                    "trilinear-hex-k3-lognormal", "trilinear-hex-k1-modes",
                                                    "trilinear-hex-k0", "bilinear-quad-k2-lognormal"])
def test_gpu_assembly_matches_host(ctx, kw):
    """Local operators, static condensation and assembly on the GPU give the
    host assembly's operator and right-hand side (to rounding: the GPU fuses
    multiply-adds)."""
    P = G.HhoProblem(**kw)
    A, b = P.assemble_gpu(ctx, keep_host=True)
    Ah, Ag = P.scipy(), A.scipy()
    assert Ag.shape == Ah.shape
    scale = abs(Ah).max()
    assert abs(Ag - Ah).max() <= 1e-12 * scale
    bh = P.rhs()
    assert np.linalg.norm(b.cpu().numpy() - bh) <= 1e-12 * np.linalg.norm(bh)
    # deferred assembly: the host holds the pattern only
    Q = G.HhoProblem(defer_assembly=True, **kw)
    A2, b2 = Q.assemble_gpu(ctx, keep_host=True)
    assert abs(A2.scipy() - Ag).max() == 0.0
    assert torch.equal(b2, b)
    with pytest.raises(G.FhError, match="deferred"):
        Q.rhs()


def test_captured_cycle_replays_identically(ctx, system):
    """Solves replay the captured multigrid cycle; every replay gives the
    same bits, with other device work between the solves."""
    P, H, _, _ = system
    s = G.Solver(ctx, solver="cg", prec="mg", nu=(1, 1), smoothers=("sbgs", "sbgs"), tol=1e-8, max_iter=300,
                 galerkin=True, coarse_smoother="bj", coarse_relax=0.5)
    gmg.attach(s, H)
    s.setup(H[0][1])
    b = torch.from_numpy(P.rhs()).cuda()
    ref = None
    for k in range(4):
        x = torch.full_like(b, float(k))  # solves start from 0 whatever x holds
        y = torch.ones_like(b)
        G.axpby(ctx, 2.0, b, 0.5, y)  # unrelated work on the library stream
        st = s.solve(b, x)
        ctx.sync()
        cur = (x.cpu().numpy().copy(), s.history().copy(), st["iterations"])
        if ref is None:
            ref = cur
        else:
            assert cur[2] == ref[2]
            assert np.array_equal(cur[0], ref[0])
            assert np.array_equal(cur[1], ref[1])


@pytest.mark.parametrize("store", ["f32", "bf16"])
def test_heterogeneous_mixed_precision_fcg(ctx, store):
    """Heterogeneous diffusion on hexahedra (k = 2): kappa piecewise constant
    per 2^3-cell subdomain, lognormal with sigma = 3 (contrast ~1e6), one
    subdomain in eight anisotropic diag(1, 1, 1e-2).  FP64 FCG with an FP32
    (or BF16-stored) geometric multigrid reaches 1e-8 like the FP64 MG-PCG;
    the solutions agree to the tolerance's order (error printed).  Coarse
    operators are re-discretised: on hexahedra the trace prolongation is
    rank-deficient, so its Galerkin product is singular."""
    P = G.HhoProblem(dim=3, mesh="cart", n=8, k=2, kappa="subdomain", sigma=3.0, subdomain=2, aniso=1e-2, seed=7)
    H = gmg.build_hierarchy(ctx, P)
    kw = dict(prec="mg", nu=(1, 1), smoothers=("sbgs", "sbgs"), tol=1e-8, max_iter=500, galerkin=False)
    b = torch.from_numpy(P.rhs()).cuda()
    s64 = G.Solver(ctx, solver="cg", **kw)
    gmg.attach(s64, H)
    s64.setup(H[0][1])
    x64 = torch.zeros_like(b)
    st64 = s64.solve(b, x64)
    s = G.Solver(ctx, solver="fcg", mg_prec="f32", mg_store=store, **kw)
    gmg.attach(s, H)
    s.setup(H[0][1])
    x = torch.zeros_like(b)
    st = s.solve(b, x)
    err = (torch.linalg.norm(x - x64) / torch.linalg.norm(x64)).item()
    l2_64, l2 = P.l2_error(x64.cpu().numpy()), P.l2_error(x.cpu().numpy())
    print("hetero %s: FP64 CG %d it, FCG %d it, rel. diff %.2e, L2 %.6e vs %.6e"
          % (store, st64["iterations"], st["iterations"], err, l2_64, l2))
    assert st64["converged"] == 1 and st["converged"] == 1
    assert err < 1e-6
    assert abs(l2 - l2_64) <= 1e-6 * l2_64


@pytest.mark.parametrize("kw", ASSEMBLY_CASES, ids=["tet-k2-perturbed", "tet-k0", "tet-k3-lognormal",
                                                    "tri-k2-modes", "hex-k2-subdomains", "quad-k3",
                                                    "trilinear-hex-k3-lognormal", "trilinear-hex-k1-modes",
                                                    "trilinear-hex-k0", "bilinear-quad-k2-lognormal"])
def test_gpu_l2_error_matches_host(ctx, kw):
    """Cell recovery, k+1 reconstruction and the quadrature of the L2 error on
    the GPU give the host's relative error, for the discrete solution and for
    an arbitrary face vector (the error is a fixed function of the face
    values)."""
    import scipy.sparse.linalg as spla
    P = G.HhoProblem(**kw)
    x = spla.spsolve(P.scipy().tocsc(), P.rhs())
    y = x + 0.1 * np.random.default_rng(3).standard_normal(x.shape) * np.abs(x).max()
    for v in (x, y):
        eh = P.l2_error(v)
        eg = P.l2_error_gpu(ctx, torch.from_numpy(np.ascontiguousarray(v)).to("cuda"))
        assert abs(eg - eh) <= 1e-10 * eh, (eg, eh)


@pytest.mark.parametrize("name,kw,digits", [
    ("c1_n64_cg_mg_bj", dict(dim=2, mesh="cart", n=64, k=1, source="pisine"), 9),
    ("tet8_k2_cg_bj", dict(dim=3, mesh="stetra", n=8, k=2, source="sine"), 3),
    ("stri8_k2_cg_bj", dict(dim=2, mesh="stri", n=8, k=2, source="pisine"), 7)])
def test_gpu_l2_error_matches_reference(ctx, name, kw, digits):
    """The GPU post-processing of the discrete solution gives the reference's
    L2 error to the digits the host restatement matches (the face bases of
    this package and the reference differ for k >= 1, so the solution is
    this package's own)."""
    import scipy.sparse.linalg as spla
    from conftest import load
    f = load(name)
    P = G.HhoProblem(**kw)
    x = spla.spsolve(P.scipy().tocsc(), P.rhs())
    e = P.l2_error_gpu(ctx, torch.from_numpy(np.ascontiguousarray(x)).to("cuda"))
    ref = float(f["info_l2"])
    assert float("%.*e" % (digits - 1, e)) == float("%.*e" % (digits - 1, ref)), (e, ref)


def test_gmg_chebyshev_on_sliced_and_chunked_levels_matches_oracle(ctx):
    """16^3 x 6 perturbed tetrahedra, k = 2: the fine level (47 616 block
    rows) runs on its sliced copy and the first Galerkin level (5 760 block
    rows of about 50 blocks) on its chunked copy, both with the fused
    Chebyshev step.  CG + V(1,1) (Chebyshev-2 fine, Chebyshev-1 coarse)
    needs the oracle's iteration count and gives its solution to 1e-10."""
    P = G.HhoProblem(dim=3, mesh="stetra", n=16, k=2, perturb=0.2, seed=2026, source="pisine")
    H = gmg.build_hierarchy(ctx, P)
    hctx = G.Context(-1)
    Hh = gmg.build_hierarchy(hctx, P)
    gal = [Hh[0][1]]
    for l in range(len(Hh) - 1):
        gal.append(G.galerkin_rap(hctx, gal[-1], Hh[l][2]))
    s = G.Solver(ctx, solver="cg", prec="mg", cycle="V", nu=(1, 1), smoothers=("cheb2", "cheb2"), relax=1.0,
                 tol=1e-8, max_iter=300, galerkin=True, coarse_smoother="cheb1")
    gmg.attach(s, H)
    s.setup(H[0][1])
    b = torch.from_numpy(P.rhs()).cuda()
    x = torch.zeros_like(b)
    st = s.solve(b, x)
    assert H[0][1].set_sliced(True), "fine level without an execution copy"
    levels = [(gal[l].scipy(), Hh[l][2].scipy() if Hh[l][2] is not None else None) for l in range(len(Hh))]
    xo, hist, out = O.solve(levels[0][0], P.rhs(), "cg", "mg", nf=P.nf, cycle="V", smoothers=("cheb2", "cheb2"),
                            omega=1.0, nu=(1, 1), levels=levels, max_iter=300, coarse_smoother="cheb1")
    print("GPU %d iterations, oracle %d" % (st["iterations"], out.iterations))
    assert st["converged"] == 1 and st["iterations"] == out.iterations
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.parametrize("kw", [dict(dim=3, mesh="hex", n=4, k=3, perturb=0.15, kappa="lognormal", sigma=3.0, seed=11),
                                dict(dim=2, mesh="quad", n=8, k=2, perturb=0.3, seed=14)],
                         ids=["trilinear-hex-k3", "bilinear-quad-k2"])
def test_gpu_l2_error_general_cells(ctx, kw):
    """Cell recovery, reconstruction and L2 quadrature of the general-cell
    kernel against the host on the same face solution."""
    P = G.HhoProblem(source="pisine", **kw)
    import scipy.sparse.linalg as spl
    x = spl.spsolve(P.scipy().tocsc(), P.rhs())
    xd = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()
    eh = P.l2_error(x)
    eg = P.l2_error_gpu(ctx, xd)
    assert abs(eg - eh) <= 1e-10 * eh
