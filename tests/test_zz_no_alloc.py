"""The allocation-free block-Jacobi sweep (fh_bjac_sweep_noalloc) against
fh_bjac_sweep with a caller-provided work vector: the same kernels on the
same inputs give the same bits, on two calls in a row (the second on a
smaller system, so the shared scratch array holds the first call's data),
and a system larger than the scratch array is rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
F = pytest.importorskip("fhhos4_amd")
G = F.fhgpu


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "no ROCm device visible"
    return G.Context(0)


def _system(ctx, n, k):
    P = G.HhoProblem(dim=3, mesh="stetra", n=n, k=k, perturb=0.1, seed=7, source="pisine")
    A = P.matrix(ctx, keep_host=True)
    b = torch.from_numpy(P.rhs()).cuda()
    return P, A, b


@pytest.mark.parametrize("vprec", ["f64", "f32"])
def test_noalloc_sweep_equals_sweep_with_work(ctx, vprec):
    dt = torch.float64 if vprec == "f64" else torch.float32
    for n, k, sweeps, zero in ((6, 2, 3, True), (4, 1, 2, False)):  # second call: smaller system
        P, A, b64 = _system(ctx, n, k)
        assert A.info()["nbr"] < 4096  # both calls stay on the block-CSR kernels
        bj = G.BlockJacobi(ctx, A, omega=0.7, dprec="f64" if vprec == "f64" else "f32")
        Aw = A if vprec == "f64" else A.convert("f32")  # FP32 vectors run on an FP32-stored operator
        b = b64.to(dt)
        x0 = torch.linspace(-1, 1, b.numel(), dtype=dt, device="cuda")
        x1, x2 = x0.clone(), x0.clone()
        work = torch.empty_like(b)
        torch.cuda.synchronize()
        bj.sweep(b, x1, work, nsweeps=sweeps, x_is_zero=zero, A=Aw)
        bj.sweep_noalloc(b, x2, nsweeps=sweeps, x_is_zero=zero, A=Aw)
        torch.cuda.synchronize()
        assert torch.equal(x1, x2), (n, k, vprec)
        assert float(torch.linalg.norm(x1 - x0)) > 0


def test_noalloc_sweep_rejects_systems_beyond_the_scratch(ctx):
    import scipy.sparse as sp
    n = (1 << 21) + 1  # one FP64 entry more than the 16 MiB scratch holds
    A = G.Matrix.from_csr(ctx, sp.identity(n, format="csr"), 1, keep_host=False)
    bj = G.BlockJacobi(ctx, A, omega=1.0)
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    with pytest.raises(G.FhError, match="exceed the scratch"):
        bj.sweep_noalloc(b, x, nsweeps=2)
    # the FP32 form of the same system fits (2^22 FP32 entries)
    xf = torch.zeros(n, dtype=torch.float32, device="cuda")
    bjf = G.BlockJacobi(ctx, A, omega=1.0, dprec="f32")
    bjf.sweep_noalloc(b.float(), xf, nsweeps=2, A=A.convert("f32"))
    torch.cuda.synchronize()
    assert float(xf.max()) == 1.0 and float(xf.min()) == 1.0
