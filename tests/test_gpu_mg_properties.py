"""Properties of the GPU multigrid preconditioner that hold by construction,
checked on a small perturbed-tetrahedra system (8^3 x 6, k = 2):

- a V(1,1) cycle with the same Chebyshev polynomial before and after the
  coarse correction, Galerkin coarse operators and R = P^T is a symmetric
  positive definite operator M (to rounding in FP64; to FP32 rounding with
  FP32 vectors and matrices);
- the error propagator of one cycle, I - M A, contracts the energy norm, so
  the stationary multigrid iteration lowers ||x - x*||_A at every step.
"""
import numpy as np
import pytest
import torch

F = pytest.importorskip("fhhos4_amd")
G = F.fhgpu
from fhhos4_amd import gmg  # noqa: E402

pytestmark = pytest.mark.gpu

KW = dict(prec="mg", cycle="V", nu=(1, 1), smoothers=("cheb2", "cheb2"), galerkin=True, coarse_smoother="cheb1")


@pytest.fixture(scope="module")
def ctx():
    return G.Context(0)


@pytest.fixture(scope="module")
def system(ctx):
    P = G.HhoProblem(dim=3, mesh="stetra", n=8, k=2, perturb=0.2, seed=2026, source="pisine")
    return P, gmg.build_hierarchy(ctx, P)


def _solver(ctx, H, **kw):
    s = G.Solver(ctx, **dict(KW, **kw))
    gmg.attach(s, H)
    s.setup(H[0][1])
    return s


@pytest.mark.parametrize("mg_prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_cycle_is_symmetric_positive_definite(ctx, system, mg_prec, tol):
    print("case", mg_prec, flush=True)
    P, H = system
    kw = dict(solver="fcg", mg_prec=mg_prec)
    if mg_prec != "f64":
        kw["mg_store"] = mg_prec
    s = _solver(ctx, H, **kw)
    rng = np.random.default_rng(3)
    U = [torch.from_numpy(rng.standard_normal(P.n_unknowns)).cuda() for _ in range(4)]
    MU = []
    for u in U:
        z = torch.zeros_like(u)
        s.apply_prec(u, z)
        MU.append(z)
    ctx.sync()
    uMu = [float(torch.dot(U[a], MU[a])) for a in range(4)]
    assert min(uMu) > 0
    for a in range(4):
        for b in range(a + 1, 4):
            x, y = float(torch.dot(U[a], MU[b])), float(torch.dot(U[b], MU[a]))
            assert abs(x - y) <= tol * np.sqrt(uMu[a] * uMu[b]), (a, b, x, y)


def test_stationary_multigrid_contracts_the_energy_norm(ctx, system):
    P, H = system
    A = P.scipy()
    b_host = P.rhs()
    b = torch.from_numpy(b_host).cuda()
    ref = _solver(ctx, H, solver="cg", tol=1e-13, max_iter=500)
    xs = torch.zeros_like(b)
    assert ref.solve(b, xs)["converged"] == 1
    xs = xs.cpu().numpy()

    def energy(x):
        e = x - xs
        return float(np.sqrt(e @ (A @ e)))

    s = _solver(ctx, H, solver="mg", tol=1e-30, max_iter=1)
    errs = [energy(np.zeros_like(xs))]
    for k in range(1, 7):
        s.set_stop(1e-30, k)
        x = torch.zeros_like(b)
        s.solve(b, x)
        ctx.sync()
        errs.append(energy(x.cpu().numpy()))
    print("energy norm of the error per stationary iteration", errs)
    ratios = np.array(errs[1:]) / np.array(errs[:-1])
    assert np.all(ratios < 1.0), ratios
    assert errs[-1] < 0.5 * errs[0]
