"""The hexahedral k = 3 benchmark family on the GPU: trilinear hexahedra with
+-0.15 h vertex perturbation and a per-cell lognormal kappa (sigma = 3), the
p-then-h multigrid hierarchy (p-injection keeping the lowest face modes, then
Galerkin h-levels of the k = 0 problem with CellInterp_Trace prolongations),
FP64 MG-PCG against the oracle run on the same hierarchy, and the mixed
precision FCG against the FP64 solution."""
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


def problem(defer=False, n=8):
    return G.HhoProblem(dim=3, mesh="hex", n=n, k=3, perturb=0.15, seed=2026, kappa="lognormal", sigma=3.0,
                        source="pisine", defer_assembly=defer)


@pytest.mark.parametrize("p_modes", [(1,), (3, 1)], ids=["p10-1", "p10-3-1"])
def test_hp_mgpcg_matches_oracle(ctx, p_modes):
    P = problem()
    H = gmg.build_hp_hierarchy(ctx, P, p_modes=p_modes)
    s = G.Solver(ctx, solver="cg", prec="mg", nu=(1, 1), smoothers=("cheb2", "cheb2"), coarse_smoother="cheb1",
                 galerkin=True, tol=1e-8, max_iter=500)
    gmg.attach_hp(s, H)
    s.setup(H["A"])
    b = torch.from_numpy(P.rhs()).cuda()
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    st = s.solve(b, x)
    assert st["converged"]
    levels = gmg.hp_levels_host(H, P.scipy(), P.nf)
    assert [s.level_size(l) for l in range(len(levels))] == [L[0].shape[0] for L in levels]
    xo, ho, oo = O.solve(levels[0][0], P.rhs(), "cg", "mg", nf=P.nf, smoothers=("cheb2", "cheb2"),
                         coarse_smoother="cheb1", levels=levels, max_iter=500)
    xg = x.cpu().numpy()
    assert st["iterations"] == oo.iterations, (st["iterations"], oo.iterations)
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)


def test_gpu_assembled_hierarchy_and_mixed_precision(ctx):
    """The deferred (GPU-assembled) operator gives the same FP64 solve as the
    host-assembled one; FCG with the FP32 and the BF16-stored multigrid
    reaches the tolerance with the FP64 solution to the stated accuracy."""
    Ph = problem()
    Hh = gmg.build_hp_hierarchy(ctx, Ph, p_modes=(1,))
    Pd = problem(defer=True)
    Hd = gmg.build_hp_hierarchy(ctx, Pd, p_modes=(1,))
    b = Pd.b_device
    assert np.linalg.norm(b.cpu().numpy() - Ph.rhs()) <= 1e-12 * np.linalg.norm(Ph.rhs())
    xs = {}
    for key, H, kw in (("host", Hh, {}), ("gpu", Hd, {}), ("f32", Hd, dict(solver="fcg", mg_prec="f32")),
                       ("bf16", Hd, dict(solver="fcg", mg_prec="f32", mg_store="bf16"))):
        args = dict(solver="cg", prec="mg", nu=(1, 1), smoothers=("cheb2", "cheb2"), coarse_smoother="cheb1",
                    galerkin=True, tol=1e-8, max_iter=500)
        args.update(kw)
        s = G.Solver(ctx, **args)
        gmg.attach_hp(s, H)
        s.setup(H["A"])
        x = torch.zeros_like(b)
        st = s.solve(b, x)
        assert st["converged"], (key, st)
        xs[key] = x.cpu().numpy()
    ref = np.linalg.norm(xs["host"])
    assert np.linalg.norm(xs["gpu"] - xs["host"]) <= 1e-10 * ref
    # the FP64 outer iteration refines to the tolerance: the mixed solutions
    # agree with the FP64 one to about the 1e-8 relative residual times the
    # condition of the operator
    for key in ("f32", "bf16"):
        assert np.linalg.norm(xs[key] - xs["host"]) <= 1e-6 * ref, key
    l2h = Ph.l2_error(xs["host"])
    l2g = Pd.l2_error_gpu(ctx, torch.from_numpy(xs["gpu"]).cuda())
    assert abs(l2g - l2h) <= 1e-8 * l2h


@pytest.mark.parametrize("k,n,kl", [(3, 16, 2), (3, 16, 0), (1, 16, 2)], ids=["k3-K2", "k3-Kall", "k1-K2"])
def test_pfirst_amg_kcycle_matches_oracle(ctx, k, n, kl):
    """FCG + the p-first aggregation AMG (p-injection to the face means,
    double pairwise aggregation of the scalar level) with a K-cycle on the
    levels 1..kl, the benchmark solver, against the oracle's own hierarchy
    and cycle on the same scalar matrix: the same level sizes and iteration
    count, the solution to 1e-10."""
    P = G.HhoProblem(dim=3, mesh="hex", n=n, k=k, perturb=0.15, seed=2026, kappa="lognormal", sigma=3.0,
                     source="pisine")
    As = P.scipy()
    A = G.Matrix.from_csr(ctx, As, P.nf, keep_host=True)
    kw = dict(nu=(1, 1), smoothers=("cheb2", "cheb2"), coarse_smoother="cheb1", max_iter=400)
    s = G.Solver(ctx, solver="fcg", prec="aggregamg", cycle="K", kcycle_levels=kl, tol=1e-8, **kw)
    s.setup(A)
    b = torch.from_numpy(P.rhs()).cuda()
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    st = s.solve(b, x)
    assert st["converged"]
    xo, ho, oo = O.solve(As, P.rhs(), "fcg", "aggregamg", nf=P.nf, cycle="K", kcycle_levels=kl, **kw)
    assert [s.level_size(l) for l in range(oo.nlevels)] == list(oo.level_n[:oo.nlevels])
    assert oo.nlevels >= 4  # a p-level and at least two aggregation levels
    assert st["iterations"] == oo.iterations, (st["iterations"], oo.iterations)
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)
