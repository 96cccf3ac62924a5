"""Uncondensed AMG on the GPU: FCG with a K-cycle over the U-AMG hierarchy
built from the uncondensed blocks, against the reference's own solve of the
same systems (golden fixtures: FCG + U-AMG K-cycle, block Jacobi 2/3 pre- and
post-smoothing, the reference's default U-AMG options).  Block-Jacobi cycles
must reproduce the reference's iteration count exactly and its FP64 solution
to 1e-10 relative.
"""
import numpy as np
import pytest

from conftest import csr, load

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import fhhos4_amd as F  # noqa: E402

G = F.fhgpu

CASES = [("uamg_tet8_k0", 1, 1), ("uamg_tet12_k0", 1, 1), ("uamg_tet4_k1", 4, 3), ("uamg_tet4_k2", 10, 6),
         ("uamg_stri16_k2", 6, 3)]


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "no ROCm device visible"
    return G.Context(0)


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    return t


@pytest.mark.parametrize("name,cbs,fbs", CASES)
def test_uamg_fcg_matches_reference(ctx, name, cbs, fbs):
    print("case", name, flush=True)
    d = load(name)
    M = G.Matrix.from_csr(ctx, csr(d, "A"), fbs, keep_host=True)
    s = G.Solver(ctx, solver="fcg", prec="uamg", cycle="K", smoothers=("bj23", "bj23"))
    s.setup_uncondensed(M, csr(d, "ATT"), csr(d, "ATF"), csr(d, "AFF"), cbs)
    b = dev(d["b"])
    x = torch.zeros_like(b)
    st = s.solve(b, x)
    ctx.sync()
    xs, hist = x.cpu().numpy(), s.history()
    assert st["levels"] == int(d["nlevels"])
    assert st["iterations"] == int(d["info_iterations"])
    assert np.linalg.norm(xs - d["x"]) <= 1e-10 * np.linalg.norm(d["x"])
    k = min(6, len(hist))
    np.testing.assert_allclose(hist[:k], d["hist"][:k], rtol=1e-8)


def test_uamg_on_own_assembly(ctx):
    """End to end on this package's assembly (tetrahedra 8^3 x 6, k = 1): the
    U-AMG-preconditioned FCG reaches 1e-8 and its solution equals the
    block-Jacobi-preconditioned CG one to the tolerance."""
    P = G.HhoProblem(dim=3, mesh="stetra", n=8, k=1, source="pisine")
    ATT, ATF, AFF, nc = P.uncondensed()
    M = P.matrix(ctx, keep_host=True)
    b = dev(P.rhs())
    s = G.Solver(ctx, solver="fcg", prec="uamg", cycle="K", smoothers=("bj23", "bj23"))
    s.setup_uncondensed(M, ATT, ATF, AFF, nc)
    x = torch.zeros_like(b)
    st = s.solve(b, x)
    assert st["converged"] and st["levels"] >= 2
    ref = G.Solver(ctx, solver="cg", prec="bj", max_iter=2000)
    ref.setup(M)
    y = torch.zeros_like(b)
    st2 = ref.solve(b, y)
    ctx.sync()
    print("uamg iterations", st["iterations"], "levels", st["levels"], "bj-CG iterations", st2["iterations"])
    assert st["iterations"] < st2["iterations"]
    xs, ys = x.cpu().numpy(), y.cpu().numpy()
    assert np.linalg.norm(xs - ys) <= 1e-6 * np.linalg.norm(ys)
