"""Parity of the gfx950 kernels with the CPU oracle, through the C ABI."""
import numpy as np
import pytest

from conftest import csr, load
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import fhhos4_amd as F  # noqa: E402

G = F.fhgpu

SYSTEMS = [("tet8_k0_fcg_amg", 1), ("c1_n64_cg_mg_bj", 2), ("tet4_k2", 6), ("cart8_k2_cg_mg_bj23", 6)]


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "no ROCm device visible"
    return G.Context(0)


def dev(a, dtype=torch.float64):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)
    torch.cuda.synchronize()
    return t


def host(ctx, t):
    ctx.sync()
    return t.double().cpu().numpy()


@pytest.mark.parametrize("name,nf", SYSTEMS)
def test_spmv_fp64_matches_oracle(ctx, name, nf):
    print("case", name, flush=True)
    A = csr(load(name), "A")
    M = G.Matrix.from_csr(ctx, A, nf)
    rng = np.random.default_rng(0)
    x, y0 = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    want = 1.5 * O.spmv(A, x) - 0.5 * y0
    xd, yd = dev(x), dev(y0)
    G.spmv(ctx, M, xd, yd, alpha=1.5, beta=-0.5)
    got = host(ctx, yd)
    assert np.linalg.norm(got - want) <= 1e-14 * np.linalg.norm(want) * 10
    rd = dev(np.zeros_like(x))
    G.residual(ctx, M, dev(y0), xd, rd)
    want_r = y0 - O.spmv(A, x)
    assert np.linalg.norm(host(ctx, rd) - want_r) <= 1e-13 * np.linalg.norm(want_r)


@pytest.mark.parametrize("prec,tol", [("f32", 1e-6), ("bf16", 2e-2)])
def test_spmv_reduced_storage_error_bound(ctx, prec, tol):
    A = csr(load("tet4_k2"), "A")
    M = G.Matrix.from_csr(ctx, A, 6).convert(prec)
    x = np.random.default_rng(1).standard_normal(A.shape[0])
    want = O.spmv(A, x)
    yd = dev(np.zeros_like(x), torch.float32)
    G.spmv(ctx, M, dev(x, torch.float32), yd)
    got = host(ctx, yd)
    assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want)


def test_dot_is_accurate_and_deterministic(ctx):
    rng = np.random.default_rng(2)
    for n in (1, 1000, 3_000_001):
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        xd, yd = dev(x), dev(y)
        a, b = G.dot(ctx, xd, yd), G.dot(ctx, xd, yd)
        assert a == b
        assert abs(a - x @ y) <= 1e-12 * (np.abs(x) @ np.abs(y))


def test_axpby(ctx):
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(12345), rng.standard_normal(12345)
    yd = dev(y)
    G.axpby(ctx, 0.25, dev(x), -2.0, yd)
    np.testing.assert_allclose(host(ctx, yd), 0.25 * x - 2.0 * y, rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("name,nf", SYSTEMS)
def test_block_jacobi_matches_oracle(ctx, name, nf):
    print("case", name, flush=True)
    A = csr(load(name), "A")
    M = G.Matrix.from_csr(ctx, A, nf)
    inv_ref = O.block_inverse(A, nf)
    for omega in (1.0, 2.0 / 3.0):
        bj = G.BlockJacobi(ctx, M, omega)
        inv = bj.inverses(A.shape[0] // nf, nf)
        np.testing.assert_allclose(inv, inv_ref, rtol=1e-10, atol=1e-10 * np.abs(inv_ref).max())
        rng = np.random.default_rng(4)
        b, x = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
        want = O.bj_sweep(A, nf, inv_ref, omega, b, x)
        want = O.bj_sweep(A, nf, inv_ref, omega, b, want)
        xd, wd = dev(x), dev(np.zeros_like(x))
        bj.sweep(dev(b), xd, wd, nsweeps=2)
        got = host(ctx, xd)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
        # from a zero guess: x = w D^-1 b
        want0 = O.bj_sweep(A, nf, inv_ref, omega, b, np.zeros_like(b))
        xd = dev(np.full_like(x, 7.0))
        bj.sweep(dev(b), xd, wd, nsweeps=1, x_is_zero=True)
        assert np.linalg.norm(host(ctx, xd) - want0) <= 1e-12 * np.linalg.norm(want0)


@pytest.mark.parametrize("name,nf", SYSTEMS)
def test_multicolour_gs_matches_oracle(ctx, name, nf):
    print("case", name, flush=True)
    A = csr(load(name), "A")
    M = G.Matrix.from_csr(ctx, A, nf)
    col = G.Colouring(ctx, M)
    nc, colour = col.get()
    cptr = np.zeros(nc + 1, dtype=np.int32)
    np.add.at(cptr, colour + 1, 1)
    cptr = np.cumsum(cptr).astype(np.int32)
    rows = np.argsort(colour, kind="stable").astype(np.int32)
    inv = O.block_inverse(A, nf)
    bj = G.BlockJacobi(ctx, M, 1.0)
    rng = np.random.default_rng(5)
    b, x = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    for direction in (0, 1, 2):
        if direction < 2:
            want = O.mcbgs_sweep(A, nf, inv, cptr, rows, direction, b, x)
        else:
            want = O.mcbgs_sweep(A, nf, inv, cptr, rows, 1, b, O.mcbgs_sweep(A, nf, inv, cptr, rows, 0, b, x))
        xd = dev(x)
        G.mcgs_sweep(ctx, M, bj, col, direction, dev(b), xd)
        got = host(ctx, xd)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), direction
