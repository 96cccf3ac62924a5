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


def _modes(ctx, M, x, y0, b, vdt, bj=None):
    """y = 1.5 A x - 0.5 y0, r = b - A x and two block-Jacobi sweeps on M."""
    yd = dev(y0, vdt)
    G.spmv(ctx, M, dev(x, vdt), yd, alpha=1.5, beta=-0.5)
    rd = dev(np.zeros_like(x), vdt)
    G.residual(ctx, M, dev(b, vdt), dev(x, vdt), rd)
    out = [host(ctx, yd), host(ctx, rd)]
    if bj is not None:
        xd, wd = dev(x, vdt), dev(np.zeros_like(x), vdt)
        bj.sweep(dev(b, vdt), xd, wd, nsweeps=2, A=M)
        out.append(host(ctx, xd))
    return out


@pytest.mark.parametrize("name,nf", SYSTEMS)
def test_sliced_layout_matches_oracle_and_block_csr(ctx, name, nf):
    """The sliced execution layout gives the block-CSR results (FP64: to
    summation-order roundoff; FP32 / BF16-stored: to FP32 roundoff) and the
    oracle's, in the three modes of the block-row kernel."""
    print("case", name, flush=True)
    A = csr(load(name), "A")
    rng = np.random.default_rng(6)
    x, y0, b = (rng.standard_normal(A.shape[0]) for _ in range(3))
    inv = O.block_inverse(A, nf)
    want = [1.5 * O.spmv(A, x) - 0.5 * y0, b - O.spmv(A, x),
            O.bj_sweep(A, nf, inv, 0.7, b, O.bj_sweep(A, nf, inv, 0.7, b, x))]
    for prec, vdt, tol in (("f64", torch.float64, 1e-13), ("f32", torch.float32, 2e-6), ("bf16", torch.float32, 2e-6)):
        res = {}
        for sliced in (False, True):
            M64 = G.Matrix.from_csr(ctx, A, nf)
            M = M64 if prec == "f64" else M64.convert(prec)
            assert M.set_sliced(sliced) == sliced
            bj = G.BlockJacobi(ctx, M64, 0.7, dprec="f64" if prec == "f64" else "f32")
            res[sliced] = _modes(ctx, M, x, y0, b, vdt, bj)
        for k, (u, v) in enumerate(zip(res[False], res[True])):
            assert np.linalg.norm(u - v) <= tol * np.linalg.norm(u), (prec, k)
        if prec == "f64":
            for k, (u, w) in enumerate(zip(res[True], want)):
                assert np.linalg.norm(u - w) <= 1e-12 * np.linalg.norm(w), k


def _random_blocks(rng, nbr, nbc, nf, wmax):
    lens = rng.integers(0, wmax + 1, size=nbr)
    lens[rng.integers(0, nbr)] = wmax
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(nbc, size=k, replace=False)) for k in lens]).astype(np.int32)
    v = rng.standard_normal((int(rp[-1]), nf, nf))
    return rp, ci, v


@pytest.mark.parametrize("nf", [1, 2, 3, 4, 6, 10])
def test_sliced_layout_irregular_rows(ctx, nf):
    """Rows of 0..W blocks (W = the widest that fits one wavefront), a
    non-square block pattern and a row count that is not a multiple of the
    rows per wavefront: sliced products equal scipy's.  Patterns whose rows
    do not fit a wavefront take the chunked wide-row layout (several chunks
    per row, empty rows, a row of exactly one chunk) with the same result."""
    import scipy.sparse as sp
    rng = np.random.default_rng(nf)
    cw = 2 if nf % 2 == 0 else 1
    wfit = 64 // (nf // cw)
    for nbr, wmax in ((4099, min(wfit, 40)), (517, 3), (300, wfit + 1), (4101, 3 * wfit + 5), (64, wfit)):
        nbc = nbr + 37
        rp, ci, v = _random_blocks(rng, nbr, nbc, nf, wmax)
        S = sp.bsr_matrix((v, ci, rp), shape=(nbr * nf, nbc * nf)).tocsr()
        M = G.Matrix.from_blocks(ctx, nbc, nf, rp, ci, v)
        assert M.set_sliced(True)
        x = rng.standard_normal(nbc * nf)
        y0 = rng.standard_normal(nbr * nf)
        yd = dev(y0)
        G.spmv(ctx, M, dev(x), yd, alpha=-2.0, beta=0.25)
        want = -2.0 * (S @ x) + 0.25 * y0
        assert np.linalg.norm(host(ctx, yd) - want) <= 1e-13 * np.linalg.norm(want), (nbr, wmax)


@pytest.mark.parametrize("nf", [1, 2, 3, 4, 6, 10])
def test_wide_row_layout_modes(ctx, nf):
    """The chunked layout of wide rows (the Galerkin coarse operators) in the
    three modes of the block-row kernel, FP64 against the block-CSR kernel
    and the oracle, FP32 and BF16-stored against the block-CSR kernel."""
    import scipy.sparse as sp
    rng = np.random.default_rng(40 + nf)
    cw = 2 if nf % 2 == 0 else 1
    wfit = 64 // (nf // cw)
    nbr = 4500
    lens = rng.integers(wfit + 1, 3 * wfit, size=nbr)
    lens[7] = 0
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)

    def row_cols(i, k):  # the diagonal and k - 1 other columns, sorted
        if k == 0:
            return np.zeros(0, dtype=np.int64)
        c = rng.choice(nbr - 1, size=k - 1, replace=False)
        return np.sort(np.concatenate([[i], c + (c >= i)]))

    ci = np.concatenate([row_cols(i, k) for i, k in enumerate(lens)]).astype(np.int32)
    v = rng.standard_normal((int(rp[-1]), nf, nf))
    for i in range(nbr):  # dominant diagonal blocks: block Jacobi is well defined
        for p in range(rp[i], rp[i + 1]):
            if ci[p] == i:
                v[p] += 4.0 * nf * lens[i] * np.eye(nf)
    A = sp.bsr_matrix((v, ci, rp), shape=(nbr * nf, nbr * nf)).tocsr()
    A = (A + sp.kron(sp.diags((lens == 0).astype(float)), np.eye(nf))).tocsr()  # the empty row: identity block
    x, y0, b = (rng.standard_normal(nbr * nf) for _ in range(3))
    inv = O.block_inverse(A, nf)
    want = [1.5 * O.spmv(A, x) - 0.5 * y0, b - O.spmv(A, x),
            O.bj_sweep(A, nf, inv, 0.7, b, O.bj_sweep(A, nf, inv, 0.7, b, x))]
    for prec, vdt, tol in (("f64", torch.float64, 1e-13), ("f32", torch.float32, 2e-6), ("bf16", torch.float32, 2e-6)):
        res = {}
        for layout in (False, True):
            M64 = G.Matrix.from_csr(ctx, A, nf)
            M = M64 if prec == "f64" else M64.convert(prec)
            assert M.set_sliced(layout) == layout
            bj = G.BlockJacobi(ctx, M64, 0.7, dprec="f64" if prec == "f64" else "f32")
            res[layout] = _modes(ctx, M, x, y0, b, vdt, bj)
        for k, (u, w) in enumerate(zip(res[False], res[True])):
            assert np.linalg.norm(u - w) <= tol * np.linalg.norm(u), (prec, k)
        if prec == "f64":
            for k, (u, w) in enumerate(zip(res[True], want)):
                assert np.linalg.norm(u - w) <= 1e-12 * np.linalg.norm(w), k


@pytest.mark.parametrize("nf", [1, 2, 3, 6, 10])
def test_sliced_two_part_layout_modes(ctx, nf):
    """Rows of two length classes (1-3 blocks and nearly the widest that fits
    a wavefront, like a prolongation's uneven rows) pad a one-width sliced
    copy by far more than 25%: the first product splits the rows into two
    parts, each with its own width and a row map.  The three modes equal the
    block-CSR kernel's results (FP64, FP32, BF16-stored) and the oracle's."""
    import scipy.sparse as sp
    rng = np.random.default_rng(70 + nf)
    cw = 2 if nf % 2 == 0 else 1
    wfit = 64 // (nf // cw)
    nbr = 33001  # the split layout is used from 32768 block rows on
    lens = np.where(rng.random(nbr) < 0.5, rng.integers(1, 4, size=nbr), rng.integers(max(wfit - 1, 4), wfit + 1, size=nbr))
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    offsets = np.arange(1, 200)

    def row_cols(i, k):  # the diagonal and k - 1 distinct nearby columns, sorted
        c = (i + rng.choice(offsets, size=k - 1, replace=False) * rng.choice([-1, 1], size=k - 1)) % nbr
        c = np.unique(c[c != i])
        while len(c) < k - 1:  # rare collisions of +o and -o' modulo nbr
            c = np.unique(np.concatenate([c, [(i + 200 + len(c)) % nbr]]))
        return np.sort(np.concatenate([[i], c[:k - 1]]))

    ci = np.concatenate([row_cols(i, k) for i, k in enumerate(lens)]).astype(np.int32)
    v = rng.standard_normal((int(rp[-1]), nf, nf))
    for i in range(nbr):
        for p in range(rp[i], rp[i + 1]):
            if ci[p] == i:
                v[p] += 4.0 * nf * lens[i] * np.eye(nf)
    A = sp.bsr_matrix((v, ci, rp), shape=(nbr * nf, nbr * nf)).tocsr()
    x, y0, b = (rng.standard_normal(nbr * nf) for _ in range(3))
    inv = O.block_inverse(A, nf)
    want = [1.5 * O.spmv(A, x) - 0.5 * y0, b - O.spmv(A, x),
            O.bj_sweep(A, nf, inv, 0.7, b, O.bj_sweep(A, nf, inv, 0.7, b, x))]
    for prec, vdt, tol in (("f64", torch.float64, 1e-13), ("f32", torch.float32, 2e-6), ("bf16", torch.float32, 2e-6)):
        res = {}
        for auto in (False, True):
            M64 = G.Matrix.from_csr(ctx, A, nf)
            M = M64 if prec == "f64" else M64.convert(prec)
            if not auto:
                assert not M.set_sliced(False)
            bj = G.BlockJacobi(ctx, M64, 0.7, dprec="f64" if prec == "f64" else "f32")
            res[auto] = _modes(ctx, M, x, y0, b, vdt, bj)
            assert M.layout() == (2 if auto else 0), (prec, M.layout())
        for k, (u, w) in enumerate(zip(res[False], res[True])):
            assert np.linalg.norm(u - w) <= tol * np.linalg.norm(u), (prec, k)
        if prec == "f64":
            for k, (u, w) in enumerate(zip(res[True], want)):
                assert np.linalg.norm(u - w) <= 1e-12 * np.linalg.norm(w), k


def test_sell_layout_modes_for_scalar_levels(ctx):
    """Scalar square operators with rows of 5-11 entries (the aggregation
    levels' shape) get the SELL-64 copy (layout 4) at their first product:
    the product, residual and block-Jacobi modes equal the block-CSR kernel's
    (FP64 to summation-order roundoff, FP32 / BF16-stored to FP32 roundoff)
    and the oracle's; a Chebyshev-smoothed FCG + AMG solve on such a level
    hierarchy is checked by the K-cycle parity tests."""
    import scipy.sparse as sp
    rng = np.random.default_rng(90)
    n = 20011  # a partial last slice
    lens = rng.integers(5, 12, size=n)
    rows, cols = [], []
    for i, k in enumerate(lens):
        c = np.unique((i + rng.integers(-300, 301, size=k)) % n)
        c = np.unique(np.concatenate([[i], c]))
        rows += [i] * len(c)
        cols += list(c)
    v = rng.standard_normal(len(rows))
    A = sp.csr_matrix((v, (rows, cols)), shape=(n, n))
    A = (A + sp.diags(4.0 * np.abs(A).sum(axis=1).A.ravel() + 1.0)).tocsr()
    A.sort_indices()
    x, y0, b = (rng.standard_normal(n) for _ in range(3))
    inv = O.block_inverse(A, 1)
    want = [1.5 * O.spmv(A, x) - 0.5 * y0, b - O.spmv(A, x),
            O.bj_sweep(A, 1, inv, 0.7, b, O.bj_sweep(A, 1, inv, 0.7, b, x))]
    for prec, vdt, tol in (("f64", torch.float64, 1e-13), ("f32", torch.float32, 2e-6), ("bf16", torch.float32, 2e-6)):
        res = {}
        for auto in (False, True):
            M64 = G.Matrix.from_csr(ctx, A, 1)
            M = M64 if prec == "f64" else M64.convert(prec)
            if not auto:
                assert not M.set_sliced(False)
            bj = G.BlockJacobi(ctx, M64, 0.7, dprec="f64" if prec == "f64" else "f32")
            res[auto] = _modes(ctx, M, x, y0, b, vdt, bj)
            assert M.layout() == (4 if auto else 0), (prec, M.layout())
        for k, (u, w) in enumerate(zip(res[False], res[True])):
            assert np.linalg.norm(u - w) <= tol * np.linalg.norm(u), (prec, k)
        if prec == "f64":
            for k, (u, w) in enumerate(zip(res[True], want)):
                assert np.linalg.norm(u - w) <= 1e-12 * np.linalg.norm(w), k
