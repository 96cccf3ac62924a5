"""Learned post-smoother on the GPU: forward and gradient of the HIP kernels
against the FP64 oracle (itself pinned by finite differences in
test_oracle.py), BF16 inference, training, and the gain of the trained
network inside FCG over the same solver without it."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

F = pytest.importorskip("fhhos4_amd")
G = F.fhgpu
pytestmark = pytest.mark.gpu

SOLVER = dict(prec="aggregamg", cycle="V", nu=(1, 1), smoothers=("sbgs", "sbgs"), fcg_truncation=8)


@pytest.fixture(scope="module")
def ctx():
    return G.Context(0)


def _problem(n=6, seed=3, sigma=1.0, source="pisine", source_seed=1):
    return G.HhoProblem(dim=2, mesh="stri", n=n, k=2, perturb=0.2, seed=seed, kappa="lognormal", sigma=sigma,
                        source=source, source_seed=source_seed)


def _random_params(nn, scale=0.1, seed=1):
    p = nn.get_params()
    rng = np.random.default_rng(seed)
    return (p + scale * rng.standard_normal(p.size)).astype(np.float32)


@pytest.mark.parametrize("layers", [1, 3])
def test_forward_matches_oracle(ctx, layers):
    P = _problem()
    A = P.matrix(ctx, keep_host=True)
    feat = G.face_features(P)
    g = G.NnGraph(ctx, A, feat)
    nn = G.Mpnn(ctx, P.nf, layers=layers, seed=7)
    params = _random_params(nn)
    nn.set_params(params)
    r1 = np.random.default_rng(2).standard_normal(P.n_unknowns)
    ref = O.mpnn_forward(P.scipy(), P.nf, feat, params.astype(np.float64), r1, layers=layers)
    d = torch.zeros(P.n_unknowns, dtype=torch.float64, device="cuda")
    nn.forward(g, torch.from_numpy(r1).cuda(), d, bf16=False)
    err32 = np.linalg.norm(d.cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err32 < 1e-5, err32
    nn.forward(g, torch.from_numpy(r1).cuda(), d, bf16=True)
    err16 = np.linalg.norm(d.cpu().numpy() - ref) / np.linalg.norm(ref)
    print("layers", layers, "FP32 error", err32, "BF16 error", err16)
    assert err16 < 3e-2, err16
    # positive homogeneity: the correction scales with the residual (sigma > 0)
    nn.forward(g, torch.from_numpy(2.5 * r1).cuda(), d, bf16=False)
    assert np.linalg.norm(d.cpu().numpy() - 2.5 * ref) <= 1e-5 * 2.5 * np.linalg.norm(ref)


def test_untrained_network_is_block_jacobi(ctx):
    P = _problem()
    A = P.matrix(ctx, keep_host=True)
    g = G.NnGraph(ctx, A, G.face_features(P))
    nn = G.Mpnn(ctx, P.nf, layers=3, seed=7)
    r1 = np.random.default_rng(4).standard_normal(P.n_unknowns)
    d = torch.zeros(P.n_unknowns, dtype=torch.float64, device="cuda")
    nn.forward(g, torch.from_numpy(r1).cuda(), d, bf16=False)
    S = P.scipy().toarray()
    nf = P.nf
    bj = np.zeros_like(r1)
    for i in range(P.nb):
        s = slice(i * nf, (i + 1) * nf)
        bj[s] = np.linalg.solve(S[s, s], r1[s])
    assert np.linalg.norm(d.cpu().numpy() - bj) <= 1e-5 * np.linalg.norm(bj)


@pytest.mark.parametrize("layers", [1, 3])
def test_gradient_matches_oracle(ctx, layers):
    P = _problem(n=5)
    A = P.matrix(ctx, keep_host=True)
    feat = G.face_features(P)
    g = G.NnGraph(ctx, A, feat)
    nn = G.Mpnn(ctx, P.nf, layers=layers, seed=11)
    params = _random_params(nn, 0.2, seed=5)
    nn.set_params(params)
    rng = np.random.default_rng(6)
    e1 = rng.standard_normal(P.n_unknowns)
    S = P.scipy()
    r1 = S @ e1
    loss_ref, grad_ref = O.mpnn_loss_grad(S, P.nf, feat, params.astype(np.float64), r1, e1, layers=layers)
    loss, grad = nn.loss_grad(g, torch.from_numpy(r1).cuda(), torch.from_numpy(e1).cuda())
    assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref)
    scale = np.abs(grad_ref).max()
    assert np.abs(grad - grad_ref).max() <= 2e-4 * scale, np.abs(grad - grad_ref).max() / scale
    # deterministic: a second evaluation gives the same bits
    loss2, grad2 = nn.loss_grad(g, torch.from_numpy(r1).cuda(), torch.from_numpy(e1).cuda())
    assert loss2 == loss and np.array_equal(grad2, grad)


@pytest.fixture(scope="module")
def trained(ctx):
    """A network trained on pairs of two operators (seeded lognormal kappa,
    sigma = 2) and 8 random-mode right-hand sides each."""
    from fhhos4_amd import learned as Lr
    systems, pairs = [], []
    for ks in (11, 12):
        sy = Lr.System(ctx, _problem(n=24, seed=ks, sigma=2.0), SOLVER)
        rhs = [_problem(n=24, seed=ks, sigma=2.0, source="modes", source_seed=100 * ks + j).rhs() for j in range(8)]
        systems.append(sy)
        pairs += Lr.make_pairs(ctx, sy, rhs, iters=(1, 3, 6))
    nn = G.Mpnn(ctx, 3, layers=3, seed=1)
    before = Lr.mean_loss(nn, pairs)
    hist = Lr.train(nn, pairs, epochs=15, lr=3e-3)
    return nn, pairs, before, hist, systems


def test_training_lowers_the_energy_error(trained):
    nn, pairs, before, hist, _ = trained
    print("relative energy error of the corrected error: untrained %.4f, epochs %s" % (before, np.round(hist, 4)))
    assert len(pairs) >= 40
    assert hist[-1] < before - 0.03
    assert hist[-1] < hist[0]


def test_learned_fcg_gain(ctx, trained):
    """What the trained network gains on an unseen operator (kappa seed 13)
    and unseen right-hand sides.

    Per application, the relative energy error left after the correction on
    held-out pairs is lower than with the untrained network (whose correction
    is the block-Jacobi step).  Inside FCG(8) + AMG the iteration counts of the
    plain solver, the block-Jacobi post-step and the learned post-smoother
    (energy-optimal step and unit step) are reported; every variant reaches
    1e-8 with the plain solver's solution.  The counts are a measurement, not
    a claim: on this problem the per-application gain does not shorten the
    flexible iteration."""
    from fhhos4_amd import learned as Lr
    nn, _, _, _, _ = trained
    sy = Lr.System(ctx, _problem(n=24, seed=13, sigma=2.0), SOLVER)
    untrained = G.Mpnn(ctx, 3, layers=3, seed=1)
    held = Lr.make_pairs(ctx, sy, [_problem(n=24, seed=13, sigma=2.0, source="modes", source_seed=5000 + j).rhs()
                                   for j in range(3)], iters=(1, 3, 6))
    l_trained, l_untrained = Lr.mean_loss(nn, held), Lr.mean_loss(untrained, held)
    print("held-out relative energy error: untrained %.4f, trained %.4f" % (l_untrained, l_trained))
    assert l_trained < l_untrained - 0.02
    b = _problem(n=24, seed=13, sigma=2.0, source="modes", source_seed=9999).rhs()
    st0, x0, _ = Lr.solve_with(ctx, sy, b, None, solver_kw=SOLVER)
    st1, x1, _ = Lr.solve_with(ctx, sy, b, untrained, solver_kw=SOLVER)
    st2, x2, _ = Lr.solve_with(ctx, sy, b, nn, solver_kw=SOLVER)
    st3, x3, _ = Lr.solve_with(ctx, sy, b, nn, solver_kw=SOLVER, unit_step=True)
    print("FCG(8) iterations: plain %d, block-Jacobi post-step %d, learned (optimal step) %d [%d accepted, %d "
          "fallback], learned (unit step) %d [%d accepted, %d fallback]"
          % (st0["iterations"], st1["iterations"], st2["iterations"], st2["nn_used"], st2["nn_fallback"],
             st3["iterations"], st3["nn_used"], st3["nn_fallback"]))
    for st in (st0, st1, st2, st3):
        assert st["converged"] == 1
    for st in (st2, st3):
        assert st["nn_used"] + st["nn_fallback"] == st["iterations"]
    xr = x0.cpu().numpy()
    for x in (x1, x2, x3):
        assert np.linalg.norm(x.cpu().numpy() - xr) <= 1e-6 * np.linalg.norm(xr)


def test_learned_gain_with_geometric_multigrid(ctx):
    """The regime where the trained network shortens the flexible iteration:
    2-D perturbed triangles (HHO k = 2, per-cell lognormal kappa, sigma = 2)
    with FCG(8) and the Galerkin geometric multigrid (V(1,1), degree-2
    Chebyshev), whose coarse space leaves error the network learns to remove.
    Trained on 90 pairs of the 64^2 mesh, evaluated on the 128^2 mesh with
    unseen kappa and right-hand-side seeds: the learned post-smoother needs
    fewer iterations than the plain solver and than the block-Jacobi
    post-step (the untrained network), with the same solution.  (With the
    p-first aggregation AMG K-cycle of the hexahedral benchmark the plain
    solver converges in 42 iterations on 2-D quadrilaterals and the network
    gains nothing; there it is measured, not asserted.)"""
    from fhhos4_amd import gmg
    from fhhos4_amd import learned as Lr
    solver = dict(prec="mg", cycle="V", nu=(1, 1), smoothers=("cheb2", "cheb2"), galerkin=True, fcg_truncation=8)
    base = dict(dim=2, mesh="stri", k=2, perturb=0.2, kappa="lognormal", sigma=2.0)

    def system(problem):
        return Lr.System(ctx, problem, solver, tol=1e-12, max_iter=600, hierarchy=gmg.build_hierarchy(ctx, problem))

    train = system(G.HhoProblem(n=64, seed=11, source="pisine", **base))
    rhs = [G.HhoProblem(n=64, seed=11, source="modes", source_seed=1000 + j, **base).rhs() for j in range(30)]
    pairs = Lr.make_pairs(ctx, train, rhs, iters=(1, 3, 6))
    nn = G.Mpnn(ctx, 3, layers=3, seed=1)
    Lr.train(nn, pairs, epochs=8, lr=2e-3)
    P = G.HhoProblem(n=128, seed=12, source="modes", source_seed=5000, **base)
    ev = system(P)
    b = P.rhs()
    st0, x0, _ = Lr.solve_with(ctx, ev, b, None, solver_kw=solver, max_iter=600)
    st1, x1, _ = Lr.solve_with(ctx, ev, b, G.Mpnn(ctx, 3, layers=3, seed=1), solver_kw=solver, max_iter=600)
    st2, x2, _ = Lr.solve_with(ctx, ev, b, nn, solver_kw=solver, max_iter=600)
    print("FCG(8) + Galerkin GMG iterations: plain %d, block-Jacobi post-step %d, learned %d"
          % (st0["iterations"], st1["iterations"], st2["iterations"]))
    for st in (st0, st1, st2):
        assert st["converged"] == 1
    assert st2["iterations"] <= 0.92 * st0["iterations"]
    assert st2["iterations"] < st1["iterations"]
    xr = x0.cpu().numpy()
    assert np.linalg.norm(x2.cpu().numpy() - xr) <= 1e-6 * np.linalg.norm(xr)
