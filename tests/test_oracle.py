"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

Every oracle function is checked against the reference's golden fixtures
(tests/golden/*.npz, written by the reference's own solver) and/or against a
closed form or brute force that a plausible mistake would fail.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import csr, gmg_levels, load
from oracle import oracle as O

GMG_CASES = [
    # fixture, system fixture, solver, prec, cycle, smoothers, nf, max_iter
    ("c1_n64_cg_mg_bj", "c1_n64_cg_mg_bj", "cg", "mg", "V", ("bj", "bj"), 2, 200),
    ("c1_n64_cg_mg_bj23", "c1_n64_cg_mg_bj", "cg", "mg", "V", ("bj23", "bj23"), 2, 200),
    ("c1_n32_cg_mg_bj23", "c1_n32_cg_mg_bj23", "cg", "mg", "V", ("bj23", "bj23"), 2, 200),
    ("c1_n32_cg_mg_bj", "c1_n32_cg_mg_bj23", "cg", "mg", "V", ("bj", "bj"), 2, 15),
    ("c1_n32_cg_mg_bgs", "c1_n32_cg_mg_bj23", "cg", "mg", "V", ("bgs", "rbgs"), 2, 200),
    ("c1_n32_cg_bj", "c1_n32_cg_mg_bj23", "cg", "bj", "V", ("bj", "bj"), 2, 200),
    ("c1_n32_cg_none", "c1_n32_cg_mg_bj23", "cg", "none", "V", ("bj", "bj"), 2, 200),
    ("c1_n32_fcg_mg_bj23", "c1_n32_cg_mg_bj23", "fcg", "mg", "V", ("bj23", "bj23"), 2, 200),
    ("c1_n32_cg_mg_w_bj23", "c1_n32_cg_mg_bj23", "cg", "mg", "W", ("bj23", "bj23"), 2, 200),
    ("c1_n32_fcg_mg_k_bgs", "c1_n32_cg_mg_bj23", "fcg", "mg", "K", ("bgs", "rbgs"), 2, 200),
    ("cart8_k2_cg_mg_bj23", "cart8_k2_cg_mg_bj23", "cg", "mg", "V", ("bj23", "bj23"), 6, 200),
    ("tet4_k2_cg_mg_bj23", "tet4_k2_cg_mg_bj23", "cg", "mg", "V", ("bj23", "bj23"), 6, 200),
]


@pytest.mark.parametrize("case", GMG_CASES, ids=[c[0] for c in GMG_CASES])
def test_oracle_gmg_matches_reference(case):
    name, sysname, solver, prec, cycle, sm, nf, maxit = case
    f = load(name)
    d = load(sysname)
    A = csr(d, "A")
    x, hist, out = O.solve(A, d["b"], solver, prec, nf=nf, cycle=cycle, smoothers=sm,
                           levels=gmg_levels(d), max_iter=maxit)
    assert out.iterations == int(f["info_iterations"])
    assert np.linalg.norm(x - f["x"]) <= 1e-10 * np.linalg.norm(f["x"])
    k = min(6, len(hist))
    np.testing.assert_allclose(hist[:k], f["hist"][:k], rtol=1e-9)


AMG_CASES = [
    ("tet8_k0_fcg_amg", "fcg", "K", 200),
    ("tet8_k0_cg_amg_v", "cg", "V", 25),
]


@pytest.mark.parametrize("case", AMG_CASES, ids=[c[0] for c in AMG_CASES])
def test_oracle_amg_matches_reference(case):
    name, solver, cycle, maxit = case
    f = load(name)
    d = load("tet8_k0_fcg_amg")
    A = csr(d, "A")
    x, hist, out = O.solve(A, d["b"], solver, "aggregamg", nf=1, cycle=cycle, smoothers=("bgs", "rbgs"),
                           max_iter=maxit)
    assert out.iterations == int(f["info_iterations"])
    assert np.linalg.norm(x - f["x"]) <= 1e-10 * np.linalg.norm(f["x"])
    np.testing.assert_allclose(hist[:8], f["hist"][:8], rtol=1e-9)


def test_oracle_aggregation_bit_exact_scalar():
    d = load("tet8_k0_fcg_amg")
    na, agg, coarsest = O.pairwise_aggregate(csr(d, "A"), 1)
    assert np.array_equal(agg, d["agg1"])
    assert na == d["agg1"].max() + 1 and not coarsest


def test_oracle_aggregation_bit_exact_blocks():
    d = load("tet4_k2")
    na, agg, _ = O.pairwise_aggregate(csr(d, "A"), 6)
    assert np.array_equal(agg, d["agg1"])


def test_oracle_amg_levels_match_reference_galerkin():
    d = load("tet8_k0_fcg_amg")
    sizes, _ = O.amg_hierarchy(csr(d, "A"), 1)
    ref = [csr(d, "A").shape[0]] + [csr(d, "level%d_A" % l).shape[0] for l in range(1, int(d["nlevels"]))]
    assert sizes == ref


def test_aggregation_invariants():
    d = load("tet4_k2")
    A = csr(d, "A")
    na, agg, _ = O.pairwise_aggregate(A, 6)
    counts = np.bincount(agg, minlength=na)
    assert counts.min() >= 1 and counts.max() <= 2  # pairs or singletons
    assert agg.min() == 0 and agg.max() == na - 1


def test_spmv_matches_scipy():
    d = load("tet4_k2")
    A = csr(d, "A")
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    np.testing.assert_allclose(O.spmv(A, x), A @ x, rtol=1e-13, atol=1e-13 * np.abs(A @ x).max())


def _blocks(A, nf):
    nb = A.shape[0] // nf
    Ad = A.toarray()
    return [Ad[i * nf:(i + 1) * nf, i * nf:(i + 1) * nf] for i in range(nb)], Ad


def test_block_inverse_is_inverse():
    d = load("tet4_k2")
    A = csr(d, "A")
    nf = 6
    inv = O.block_inverse(A, nf).reshape(-1, nf, nf)
    D, _ = _blocks(A, nf)
    for i in range(0, len(D), 37):
        np.testing.assert_allclose(D[i] @ inv[i], np.eye(nf), atol=1e-10)


def test_bj_sweep_residual_form():
    """Reference form w b + (1-w) D x - w (L+U) x, then D^-1, equals the
    residual form x + w D^-1 (b - A x)."""
    d = load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    nf, w = 2, 2.0 / 3.0
    rng = np.random.default_rng(1)
    b, x = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    inv = O.block_inverse(A, nf)
    got = O.bj_sweep(A, nf, inv, w, b, x)
    Dinv = sp.block_diag(list(inv.reshape(-1, nf, nf)))
    want = x + w * (Dinv @ (b - A @ x))
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)


def test_bgs_forward_solves_block_triangular_system():
    d = load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    nf = 2
    rng = np.random.default_rng(2)
    b, x = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    inv = O.block_inverse(A, nf)
    Ad = A.toarray()
    nb = A.shape[0] // nf
    mask = np.kron(np.tril(np.ones((nb, nb))), np.ones((nf, nf)))
    LD, U = Ad * mask, Ad * (1 - mask)
    got = O.bgs_sweep(A, nf, inv, 0, b, x)
    np.testing.assert_allclose(LD @ got, b - U @ x, rtol=1e-9, atol=1e-9)
    got = O.bgs_sweep(A, nf, inv, 1, b, x)
    maskU = np.kron(np.triu(np.ones((nb, nb))), np.ones((nf, nf)))
    np.testing.assert_allclose((Ad * maskU) @ got, b - (Ad * (1 - maskU)) @ x, rtol=1e-9, atol=1e-9)


def _block_graph(A, nf):
    B = (abs(A) > 0).astype(np.int8)
    nb = A.shape[0] // nf
    R = sp.kron(sp.eye(nb), np.ones((1, nf))).tocsr()
    G = (R @ B @ R.T).tocsr()
    G.sort_indices()
    return G.indptr.astype(np.int32), G.indices.astype(np.int32)


def test_greedy_colouring_is_valid():
    d = load("tet4_k2")
    brp, bci = _block_graph(csr(d, "A"), 6)
    nc, col = O.greedy_colouring(brp, bci)
    for i in range(len(brp) - 1):
        for j in bci[brp[i]:brp[i + 1]]:
            assert j == i or col[j] != col[i]
    assert nc == col.max() + 1


def test_multicolour_gs_reduces_to_lexicographic():
    """One colour per block row, visited in row order = lexicographic GS."""
    d = load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    nf = 2
    nb = A.shape[0] // nf
    rng = np.random.default_rng(3)
    b, x = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    inv = O.block_inverse(A, nf)
    cp = np.arange(nb + 1, dtype=np.int32)
    rows = np.arange(nb, dtype=np.int32)
    np.testing.assert_allclose(O.mcbgs_sweep(A, nf, inv, cp, rows, 0, b, x), O.bgs_sweep(A, nf, inv, 0, b, x),
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.mcbgs_sweep(A, nf, inv, cp, rows, 1, b, x), O.bgs_sweep(A, nf, inv, 1, b, x),
                               rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("m,restart", [(1, False), (3, False), (4, True)])
def test_fcg_equals_cg_with_fixed_spd_preconditioner(m, restart):
    """With a fixed SPD preconditioner FCG(m) reproduces CG for any
    truncation or restart (Notay): the extra A-orthogonalisations remove
    components that are zero in exact arithmetic."""
    d = load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    x1, h1, o1 = O.solve(A, d["b"], "cg", "mg", nf=2, smoothers=("bj23", "bj23"), levels=gmg_levels(d))
    x2, h2, o2 = O.solve(A, d["b"], "fcg", "mg", nf=2, smoothers=("bj23", "bj23"), levels=gmg_levels(d),
                         fcg_truncation=m, fcg_restart=restart)
    assert o1.iterations == o2.iterations
    np.testing.assert_allclose(h1, h2, rtol=1e-6)


def test_cg_solution_matches_direct_solve():
    d = load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    x, _, _ = O.solve(A, d["b"], "cg", "none", nf=2, tol=1e-12, max_iter=1000)
    xd = spla.spsolve(A.tocsc(), d["b"])
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)


def _mpnn_case(n=4, seed=3):
    """Small 2-D perturbed-triangle k=2 operator with lognormal kappa and its
    face features (built by the library's host-only assembly)."""
    G = pytest.importorskip("fhhos4_amd").fhgpu
    P = G.HhoProblem(dim=2, mesh="stri", n=n, k=2, perturb=0.2, seed=seed, kappa="lognormal", sigma=1.0)
    return P.scipy(), P.nf, G.face_features(P)


@pytest.mark.parametrize("layers", [1, 2])
def test_mpnn_oracle_gradient_matches_finite_differences(layers):
    """The oracle's backward pass against central differences of its own
    loss, on random parameters (every parameter block is probed)."""
    A, nf, feat = _mpnn_case()
    rng = np.random.default_rng(0)
    npar = O.mpnn_num_params(nf, layers)
    params = 0.3 * rng.standard_normal(npar)
    e1 = rng.standard_normal(A.shape[0])
    r1 = A @ e1
    loss, g = O.mpnn_loss_grad(A, nf, feat, params, r1, e1, layers=layers)
    assert loss > 0
    idx = np.unique(np.concatenate([rng.choice(npar, 60, replace=False), [0, npar - 1, npar - 1 - nf]]))
    worst = 0.0
    for k in idx:
        h = 1e-6 * max(1.0, abs(params[k]))
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (O.mpnn_loss_grad(A, nf, feat, pp, r1, e1, layers, False)[0] -
              O.mpnn_loss_grad(A, nf, feat, pm, r1, e1, layers, False)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[k]))
    assert worst <= 1e-7 * np.abs(g).max()


def test_mpnn_oracle_structure():
    """Zero output layer: the correction is the block-Jacobi step D^-1 r1 and
    the loss is its relative energy error; the correction is homogeneous of
    degree 1 in r1; a perfect correction (delta = e1) has loss 0."""
    A, nf, feat = _mpnn_case()
    rng = np.random.default_rng(1)
    npar = O.mpnn_num_params(nf, 3)
    params = 0.3 * rng.standard_normal(npar)
    wo = npar - nf - 32 * nf
    params[wo:] = 0.0
    r1 = rng.standard_normal(A.shape[0])
    d = O.mpnn_forward(A, nf, feat, params, r1)
    Ad = A.toarray()
    bj = np.concatenate([np.linalg.solve(Ad[i:i + nf, i:i + nf], r1[i:i + nf]) for i in range(0, len(r1), nf)])
    np.testing.assert_allclose(d, bj, rtol=1e-12, atol=1e-12 * np.abs(bj).max())
    params = 0.3 * rng.standard_normal(npar)
    d1 = O.mpnn_forward(A, nf, feat, params, r1)
    d2 = O.mpnn_forward(A, nf, feat, params, 3.0 * r1)
    np.testing.assert_allclose(d2, 3.0 * d1, rtol=1e-12, atol=1e-12 * np.abs(d2).max())
    e1 = spla.spsolve(A.tocsc(), r1)
    loss, _ = O.mpnn_loss_grad(A, nf, feat, params, r1, e1, 3, False)
    diff = e1 - d1
    assert abs(loss - diff @ (A @ diff) / (e1 @ r1)) <= 1e-12 * loss


@pytest.mark.parametrize("smoother,deg", [("cheb1", 1), ("cheb2", 2), ("cheb3", 3), ("cheb4", 4)])
def test_chebyshev_two_grid_matches_closed_form(smoother, deg):
    """One two-grid cycle from x = 0 (V(1,1), Chebyshev pre- and
    post-smoothing, exact Galerkin coarse solve) against its closed form
    x = x* + E (I - P Ac^-1 P^T A) E (0 - x*), where the Chebyshev error
    propagator is E = p(D^-1 A), p(t) = T_d((theta - t)/delta) / T_d(theta/delta)
    on [0.25, 1.1] lambda_max, evaluated through the eigenvectors of
    D^-1/2 A D^-1/2 and numpy's Chebyshev series (not a recurrence).
    lambda_max(D^-1 A) is well separated, so the oracle's power-iteration
    estimate equals it to rounding."""
    import scipy.sparse as sp
    from numpy.polynomial import chebyshev as C
    n = 24
    rng = np.random.default_rng(5)
    L = sp.diags([-np.ones(n - 1), 2.5 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).toarray()
    v = 1.0 + rng.random(n)
    A = L + 40.0 * np.outer(v, v) / n
    P = np.zeros((n, n // 4))
    for j in range(n // 4):
        P[4 * j:4 * j + 4, j] = 1.0
    Ac = P.T @ A @ P
    b = rng.standard_normal(n)
    levels = [(sp.csr_matrix(A), sp.csr_matrix(P)), (sp.csr_matrix(Ac), None)]
    x, _, out = O.solve(sp.csr_matrix(A), b, "none", "mg", nf=1, smoothers=(smoother, smoother), nu=(1, 1),
                        levels=levels, tol=1e-30, max_iter=1)
    assert out.iterations == 1
    d = np.diag(A)
    lam, U = np.linalg.eigh(A / np.sqrt(np.outer(d, d)))
    assert lam[-2] < 0.2 * lam[-1]
    hi, lo = 1.1 * lam[-1], 0.25 * lam[-1]
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    coef = np.zeros(deg + 1)
    coef[deg] = 1.0
    p = C.chebval((theta - lam) / delta, coef) / C.chebval(theta / delta, coef)
    E = (U * p) @ U.T
    E = E / np.sqrt(d)[:, None] * np.sqrt(d)[None, :]  # D^-1/2 (U p U^T) D^1/2
    xs = np.linalg.solve(A, b)
    K = np.eye(n) - P @ np.linalg.solve(Ac, P.T @ A)
    want = xs + E @ K @ E @ (-xs)
    np.testing.assert_allclose(x, want, rtol=1e-9, atol=1e-12 * np.abs(xs).max())


def test_oracle_stopping_rules():
    """IterativeSolver::StoppingCriteriaReached (IterativeSolver.h:224-247):
    max_iter = 0 stops before the first iteration (x stays the zero initial
    guess); a zero right-hand side gives a NaN residual in CG's first step,
    which the reference reports as a fatal error."""
    d = load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    x, hist, out = O.solve(A, d["b"], "cg", "none", nf=2, max_iter=0)
    assert out.iterations == 0 and not x.any() and len(hist) == 1
    with pytest.raises(RuntimeError):
        O.solve(A, np.zeros(A.shape[0]), "cg", "none", nf=2)


def test_oracle_p_transfer_is_the_injection_galerkin_closed_form():
    """The oracle's p-transfer (truncation to the first nfc modes, injection
    back) is the explicit p-injection P = [I_nfc; 0] per block, and its
    Galerkin operator P^T A P is the leading nfc x nfc sub-block of every
    block: a two-level solve through the p-transfer equals the solve through
    the explicit matrix P with the coarse operator formed as P^T A P."""
    import scipy.sparse as sp
    from conftest import csr, load
    d = load("cart8_k2_cg_mg_bj23")
    A = csr(d, "A")
    nf, nfc = 6, 1
    nb = A.shape[0] // nf
    rows = (np.arange(nb)[:, None] * nf + np.arange(nfc)[None, :]).ravel()
    cols = (np.arange(nb)[:, None] * nfc + np.arange(nfc)[None, :]).ravel()
    P = sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(nb * nf, nb * nfc))
    Ac = (P.T @ A @ P).tocsr()
    Ab = sp.bsr_matrix(A, blocksize=(nf, nf))
    sub = sp.bsr_matrix((Ab.data[:, :nfc, :nfc], Ab.indices, Ab.indptr), shape=(nb * nfc, nb * nfc)).tocsr()
    assert abs(Ac - sub).max() == 0.0
    b = d["b"]
    kw = dict(nf=nf, smoothers=("bj23", "bj23"), max_iter=100)
    x1, h1, o1 = O.solve(A, b, "cg", "mg", levels=[(A, ("p", nfc), nf), (sub, None, nfc)], **kw)
    x2, h2, o2 = O.solve(A, b, "cg", "mg", levels=[(A, P, nf), (Ac, None, nfc)], **kw)
    assert o1.iterations == o2.iterations
    np.testing.assert_allclose(h1, h2, rtol=1e-10, atol=1e-14)
    assert np.linalg.norm(x1 - x2) <= 1e-10 * np.linalg.norm(x2)


def test_oracle_block_storage_path_equals_scalar_path():
    """The oracle's block-stored system matrix (the CPU baseline of the
    26-million-unknown hexahedral system, whose scalar entries overflow int32)
    runs the same p-first aggregation AMG K-cycle FCG as the scalar path,
    which the reference's fixtures pin: the same level sizes and residual
    history, the same solution to rounding."""
    import fhhos4_amd.fhgpu as G
    P = G.HhoProblem(dim=3, mesh="hex", n=10, k=2, perturb=0.15, seed=2026, kappa="lognormal", sigma=3.0,
                     source="pisine")
    rp, ci, v = P.blocks()
    kw = dict(nu=(1, 1), smoothers=("cheb2", "cheb2"), coarse_smoother="cheb1", max_iter=400)
    x1, h1, o1 = O.solve(P.scipy(), P.rhs(), "fcg", "aggregamg", nf=P.nf, cycle="K", kcycle_levels=2, **kw)
    x2, h2, o2 = O.solve_bcsr(rp, ci, v, P.nf, P.rhs(), "fcg", "aggregamg", cycle="K", kcycle_levels=2, **kw)
    assert o1.iterations == o2.iterations and o1.nlevels == o2.nlevels >= 3
    assert list(o1.level_n[:o1.nlevels]) == list(o2.level_n[:o2.nlevels])
    np.testing.assert_allclose(h1, h2, rtol=1e-9, atol=1e-15)
    assert np.linalg.norm(x1 - x2) <= 1e-12 * np.linalg.norm(x1)


@pytest.mark.parametrize("solver,prec", [("cg", "aggregamg"), ("fcg", "aggregamg"), ("none", "aggregamg")])
def test_oracle_initial_guess_is_a_shift(solver, prec):
    """From an initial guess x0 the iterates are x0 plus those of the
    problem A e = b - A x0 from zero (the first residual is b - A x0,
    ConjugateGradient.h:36-45): same iteration count, the same solution,
    and histories that differ only by the norms of the two right-hand sides.
    An explicit zero guess reproduces the default run exactly."""
    d = load("tet8_k0_fcg_amg")
    A, b = csr(d, "A"), d["b"]
    kw = dict(nf=1, cycle="K", smoothers=("bgs", "rbgs"), max_iter=12, tol=1e-30)  # 12 iterations each
    x_ref, h_ref, o_ref = O.solve(A, b, solver, prec, **kw)
    x_zero, h_zero, o_zero = O.solve(A, b, solver, prec, x0=np.zeros_like(b), **kw)
    assert o_zero.iterations == o_ref.iterations and np.array_equal(h_zero, h_ref) and np.array_equal(x_zero, x_ref)
    x0 = np.ones_like(b)
    x1, h1, o1 = O.solve(A, b, solver, prec, x0=x0, **kw)
    r0 = b - A @ x0
    e, he, oe = O.solve(A, r0, solver, prec, **kw)
    nb, nr = np.linalg.norm(b), np.linalg.norm(r0)
    assert o1.iterations == oe.iterations == 12
    assert h1[0] == pytest.approx(nr / nb, rel=1e-14)
    # same Krylov space: histories scaled by ||r0|| / ||b||
    k = min(len(h1), len(he))
    assert np.allclose(h1[:k], he[:k] * nr / nb, rtol=1e-8, atol=0)
    assert np.linalg.norm(x1 - (x0 + e)) <= 1e-12 * np.linalg.norm(x1)


SOR_CASES = [
    # fixture (reference run), smoothers, nu, omega
    ("c1_n32_cg_mg_bsor", ("bsor", "rbsor"), (1, 1), 1.2),
    ("c1_n32_cg_mg_abgs", ("abgs", "rabgs"), (2, 2), 1.0),
]


@pytest.mark.parametrize("case", SOR_CASES, ids=[c[0] for c in SOR_CASES])
def test_oracle_sor_and_alternating_match_reference(case):
    """Block SOR (BlockSOR.h:193-223, relaxation 1.2) and alternating block
    Gauss-Seidel (abgs / rabgs: forward and backward by smoothing iteration,
    BlockGaussSeidel.h:105-120) against the reference's own CG + GMG runs:
    the same iteration count and residual history, the solution to 1e-10."""
    name, sm, nu, omega = case
    f, d = load(name), load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    x, hist, out = O.solve(A, d["b"], "cg", "mg", nf=2, smoothers=sm, nu=nu, omega=omega, levels=gmg_levels(d))
    assert out.iterations == int(f["info_iterations"])
    np.testing.assert_allclose(hist, f["hist"], rtol=1e-8, atol=1e-15)
    assert np.linalg.norm(x - f["x"]) <= 1e-10 * np.linalg.norm(f["x"])


def test_oracle_level_smoothing_change_matches_reference():
    """The per-level change of the smoothing iterations (Multigrid.h:224-242,
    '+1': one more smoothing iteration on every coarser level) in the
    reference's FCG + aggregation AMG K-cycle: the same iteration count and
    residual history (17 iterations, against 18 without the change), the
    solution to 1e-10."""
    f, d = load("tet8_k0_fcg_amg_nu"), load("tet8_k0_fcg_amg")
    A = csr(d, "A")
    kw = dict(nf=1, cycle="K", smoothers=("bgs", "rbgs"))
    x, hist, out = O.solve(A, d["b"], "fcg", "aggregamg", nu_change="+1", **kw)
    assert out.iterations == int(f["info_iterations"]) != int(d["info_iterations"])
    np.testing.assert_allclose(hist, f["hist"], rtol=1e-9, atol=1e-15)
    assert np.linalg.norm(x - f["x"]) <= 1e-10 * np.linalg.norm(f["x"])


def test_given_coarse_smoothing_counts_reproduce_the_reference_nu_change():
    """The coarse-level smoothing counts given directly (not in the
    reference): on this three-level hierarchy level 1 is the only smoothed
    coarse level, so coarse counts (2, 2) are the reference's "+1" change
    of (1, 1), and the solve reproduces its fixture; counts (1, 1) give the
    unchanged solve."""
    f, d = load("tet8_k0_fcg_amg_nu"), load("tet8_k0_fcg_amg")
    A = csr(d, "A")
    kw = dict(nf=1, cycle="K", smoothers=("bgs", "rbgs"))
    x, hist, out = O.solve(A, d["b"], "fcg", "aggregamg", coarse_nu=(2, 2), **kw)
    assert out.nlevels == 3
    assert out.iterations == int(f["info_iterations"])
    np.testing.assert_allclose(hist, f["hist"], rtol=1e-9, atol=1e-15)
    x, hist, out = O.solve(A, d["b"], "fcg", "aggregamg", coarse_nu=(1, 1), **kw)
    assert out.iterations == int(d["info_iterations"])
    np.testing.assert_allclose(hist, d["hist"], rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("omega", [0.7, 1.3])
def test_bsor_sweep_solves_the_relaxed_triangular_system(omega):
    """One forward block SOR sweep solves (D + omega L) x' = omega b -
    (omega U + (omega - 1) D) x; backward swaps L and U (L, U: strictly block
    lower / upper parts, D: diagonal blocks)."""
    d = load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    nf = 2
    nb = A.shape[0] // nf
    rng = np.random.default_rng(5)
    b, x = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    inv = O.block_inverse(A, nf)
    Ad = A.toarray()
    blk = np.kron(np.eye(nb), np.ones((nf, nf)))
    low = np.kron(np.tril(np.ones((nb, nb)), -1), np.ones((nf, nf)))
    D, L = Ad * blk, Ad * low
    U = Ad - D - L
    got = O.bsor_sweep(A, nf, inv, 0, omega, b, x)
    np.testing.assert_allclose((D + omega * L) @ got, omega * b - (omega * U + (omega - 1) * D) @ x,
                               rtol=1e-9, atol=1e-9)
    got = O.bsor_sweep(A, nf, inv, 1, omega, b, x)
    np.testing.assert_allclose((D + omega * U) @ got, omega * b - (omega * L + (omega - 1) * D) @ x,
                               rtol=1e-9, atol=1e-9)
    # omega = 1 is Gauss-Seidel
    np.testing.assert_allclose(O.bsor_sweep(A, nf, inv, 0, 1.0, b, x), O.bgs_sweep(A, nf, inv, 0, b, x),
                               rtol=0, atol=0)


def test_multicolour_sor_reduces_to_lexicographic():
    """One colour per block row, visited in row order = lexicographic SOR."""
    d = load("c1_n32_cg_mg_bj23")
    A = csr(d, "A")
    nf = 2
    nb = A.shape[0] // nf
    rng = np.random.default_rng(6)
    b, x = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    inv = O.block_inverse(A, nf)
    cp = np.arange(nb + 1, dtype=np.int32)
    rows = np.arange(nb, dtype=np.int32)
    for direction in (0, 1):
        np.testing.assert_allclose(O.mcsor_sweep(A, nf, inv, cp, rows, direction, 1.25, b, x),
                                   O.bsor_sweep(A, nf, inv, direction, 1.25, b, x), rtol=1e-12, atol=1e-12)
