"""Element layout of the GPU-assembled hexahedral operators (condensed cell
matrices kept beside the block rows): products, residuals and the fused
Chebyshev step against the block-row kernels on the same assembly, and
solves with it against the oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fhhos4_amd import fhgpu as G  # noqa: E402
from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return G.Context(0)


def _problem(n, k, seed=2026):
    return G.HhoProblem(dim=3, mesh="hex", n=n, k=k, perturb=0.15, seed=seed, kappa="lognormal", sigma=3.0,
                        source="pisine", defer_assembly=True)


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_elem_products_match_block_rows(ctx, k):
    """y = A x, y = 2 A x + 0.5 y and r = b - A x with the element layout
    against the block-row kernels on the same GPU assembly (FP64: the
    layout's cell matrices are the upper triangles of the values the block
    rows sum, so the results agree to rounding), and against the host
    product of the block values; FP32 values with FP32 vectors against the
    FP32 block rows, and with FP64 vectors (FCG's FP32-stored product)
    against the FP64 product."""
    print("case hex n=6 k=%d" % k, flush=True)
    P = _problem(6, k)
    A, b = P.assemble_gpu(ctx, keep_host=True)
    assert A.layout() == 5  # used from the first product on
    n = P.n_unknowns
    rng = np.random.default_rng(k)
    xh = rng.standard_normal(n)
    x = torch.from_numpy(xh).cuda()
    y0 = torch.from_numpy(rng.standard_normal(n)).cuda()
    out = {}
    for on in (True, False):
        assert A.set_elem(on) == on
        y = torch.empty_like(x)
        G.spmv(ctx, A, x, y)
        y2 = y0.clone()
        G.spmv(ctx, A, x, y2, alpha=2.0, beta=0.5)
        r = torch.empty_like(x)
        G.residual(ctx, A, b, x, r)
        torch.cuda.synchronize()
        out[on] = [t.cpu().numpy() for t in (y, y2, r)]
    scale = np.linalg.norm(out[False][0])
    for a, c in zip(out[True], out[False]):
        assert np.linalg.norm(a - c) <= 1e-13 * scale
    yh = A.scipy() @ xh
    assert np.linalg.norm(out[True][0] - yh) <= 1e-13 * scale
    # FP32 copy: the element layout follows the conversion
    A.set_elem(True)
    A32 = A.convert("f32")
    assert A32.layout() == 5
    x32 = x.float()
    res = {}
    for on in (True, False):
        A32.set_elem(on)
        y32 = torch.empty_like(x32)
        G.spmv(ctx, A32, x32, y32)
        torch.cuda.synchronize()
        res[on] = y32.double().cpu().numpy()
    assert np.linalg.norm(res[True] - res[False]) <= 1e-5 * scale
    A32.set_elem(True)
    y64 = torch.empty_like(x)
    G.spmv(ctx, A32, x, y64)
    # every cell matrix is rounded to FP32 on its own (the block rows round
    # the sum of a face's two cells), so the two FP32 operators differ by
    # ~1e-8; both are within FP32 rounding of the FP64 one
    assert np.linalg.norm(y64.cpu().numpy() - yh) <= 1e-6 * scale


@pytest.mark.parametrize("k,nu", [(3, (1, 1)), (3, (1, 0)), (2, (1, 1))])
def test_elem_solve_matches_block_rows_and_oracle(ctx, k, nu):
    """FCG + the p-first aggregation AMG K-cycle with Chebyshev smoothing on
    the 16^3 member of the headline family, the finest level's products and
    Chebyshev steps on the element layout: the same iteration count as the
    oracle on the assembled matrix and as the block-row kernels, solutions
    within 1e-9; then the FP32 multigrid with FCG's own product on the FP32
    element layout (FP64 residual replacement): FP64 residual below the
    tolerance, solution within 1e-6 of the FP64 one."""
    print("case hex n=16 k=%d nu=%s" % (k, nu), flush=True)
    P = _problem(16, k)
    A, b = P.assemble_gpu(ctx, keep_host=True)
    As = A.scipy()
    bh = b.cpu().numpy()
    kw = dict(solver="fcg", prec="aggregamg", cycle="K", kcycle_levels=2, nu=nu, smoothers=("cheb1", "cheb1"),
              coarse_smoother="cheb4", theta=0.1, max_coarse=2500, tol=1e-8, max_iter=400, coarse_nu=(1, 1))
    xs, its = {}, {}
    for on in (True, False):
        assert A.set_elem(on) == on
        s = G.Solver(ctx, **kw)
        s.setup(A)
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        st = s.solve(b, x)
        assert st["converged"] == 1
        xs[on], its[on] = x.cpu().numpy(), st["iterations"]
        del s
    kwo = {k2: v for k2, v in kw.items() if k2 != "solver"}
    xo, ho, oo = O.solve(As, bh, "fcg", nf=P.nf, **kwo)
    print("iterations: element layout %d, block rows %d, oracle %d" % (its[True], its[False], oo.iterations))
    assert its[True] == its[False] == oo.iterations
    ref = np.linalg.norm(xo)
    assert np.linalg.norm(xs[True] - xo) <= 1e-9 * ref
    assert np.linalg.norm(xs[False] - xo) <= 1e-9 * ref
    A.set_elem(True)
    s = G.Solver(ctx, **kw, mg_prec="f32", mg_store="f32", outer_prec="f32", replace_every=8)
    s.setup(A)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    st = s.solve(b, x)
    assert st["converged"] == 1
    xm = x.cpu().numpy()
    assert np.linalg.norm(bh - As @ xm) < 1e-8 * np.linalg.norm(bh)
    assert np.linalg.norm(xm - xo) <= 1e-6 * ref
    print("FP32 multigrid + FP32 element-layout outer product: %d iterations" % st["iterations"])


def test_learned_post_smoother_with_fp32_products(ctx):
    """The learned post-smoother inside FCG(8) with the FP32 multigrid: with
    outer_prec FP32 its two products per application (r1 = r - A z0 and
    A delta) run on the FP32 element layout with FP64 vectors, and FCG
    replaces its residual by the FP64 b - A x: the solve reaches the
    tolerance for the FP64 residual, with the solution of the FP64-product
    solve to 1e-6 and an iteration count within two of it.  The untrained
    network (its correction is the block-Jacobi step) is used, so every
    application is accepted or falls back deterministically."""
    print("case hex n=12 k=3 learned, FP32 products", flush=True)
    P = _problem(12, 3)
    A, b = P.assemble_gpu(ctx, keep_host=True)
    As = A.scipy()
    bh = b.cpu().numpy()
    nn = G.Mpnn(ctx, P.nf, layers=3, seed=1)
    graph = G.NnGraph(ctx, A, G.face_features(P))
    kw = dict(solver="fcg", prec="aggregamg", cycle="K", kcycle_levels=2, nu=(1, 0), coarse_nu=(1, 1),
              smoothers=("cheb1", "cheb1"), coarse_smoother="cheb4", theta=0.1, max_coarse=2500, tol=1e-8,
              max_iter=400, fcg_truncation=8, mg_prec="f32", mg_store="f32")
    xs, its = {}, {}
    for outer in ("f64", "f32"):
        s = G.Solver(ctx, **kw, outer_prec=outer, replace_every=8)
        s.set_learned_smoother(nn, graph, 1.0)
        s.setup(A)
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        st = s.solve(b, x)
        assert st["converged"] == 1, outer
        assert st["nn_used"] + st["nn_fallback"] == st["iterations"]
        xs[outer], its[outer] = x.cpu().numpy(), st["iterations"]
        del s
    print("iterations: FP64 products %d, FP32 products %d" % (its["f64"], its["f32"]))
    assert np.linalg.norm(bh - As @ xs["f32"]) < 1e-8 * np.linalg.norm(bh)
    assert np.linalg.norm(xs["f32"] - xs["f64"]) <= 1e-6 * np.linalg.norm(xs["f64"])
    assert abs(its["f32"] - its["f64"]) <= 2
