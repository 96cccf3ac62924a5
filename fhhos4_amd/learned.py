"""Training data and training loop of the learned post-smoother.

Pairs come from the plain solver.  For a system A x = b with converged
solution x*, the FCG iterate x_k has residual r_k = b - A x_k and error
e_k = x* - x_k.  The multigrid preconditioner maps r_k to z0 = M r_k, and the
network sees what the cycle leaves: r1 = r_k - A z0 with error
e1 = e_k - z0 (A e1 = r1).  Training minimises the relative energy norm
(e1 - delta)^T A (e1 - delta) / e1^T A e1 of the corrected error with Adam,
one pair per step, with the library's forward and backward kernels.
"""
import time

import numpy as np
import torch

from . import fhgpu as G
from . import gmg


class System:
    """One operator on the device with its solver, network graph and
    features.  With `hierarchy` (gmg.build_hierarchy of the problem) the
    solver is a geometric multigrid one (solver_kw with prec="mg")."""

    def __init__(self, ctx, problem, solver_kw, tol=1e-12, max_iter=400, hierarchy=None):
        self.problem = problem
        self.hierarchy = hierarchy
        self.A = hierarchy[0][1] if hierarchy else problem.matrix(ctx, keep_host=True)
        self.features = G.face_features(problem)
        self.graph = G.NnGraph(ctx, self.A, self.features)
        self.solver = G.Solver(ctx, solver="fcg", tol=tol, max_iter=max_iter, **solver_kw)
        if hierarchy:
            gmg.attach(self.solver, hierarchy)
        self.solver.setup(self.A)
        self.n = problem.n_unknowns
        self.tol, self.max_iter = tol, max_iter


def make_pairs(ctx, system, rhs, iters=(1, 3, 6), log=None):
    """(r1, e1) device pairs of the right-hand sides `rhs` (host arrays) at
    the FCG iterations `iters`; `log(message)` every 100 right-hand sides."""
    n = system.n
    s = system.solver
    xs = torch.zeros(n, dtype=torch.float64, device="cuda")
    xk = torch.zeros_like(xs)
    rk = torch.zeros_like(xs)
    z0 = torch.zeros_like(xs)
    pairs = []
    for j, b_host in enumerate(rhs):
        if log and j and j % 100 == 0:
            log("pairs: %d right-hand sides solved, %d pairs" % (j, len(pairs)))
        b = torch.from_numpy(np.ascontiguousarray(b_host)).cuda()
        s.set_stop(system.tol, system.max_iter)
        s.solve(b, xs)
        for k in iters:
            s.set_stop(system.tol, k)
            s.solve(b, xk)
            G.residual(ctx, system.A, b, xk, rk)
            s.apply_prec(rk, z0)
            r1 = torch.zeros_like(xs)
            G.residual(ctx, system.A, rk, z0, r1)
            e1 = xs - xk - z0
            ctx.sync()
            if float(torch.dot(e1, r1)) > 0:
                pairs.append((system, r1, e1))
        s.set_stop(system.tol, system.max_iter)
    return pairs


def train(nn, pairs, epochs=3, lr=2e-3, seed=0, log=None):
    """Adam over the pairs (one pair per step, shuffled per epoch); returns
    the mean loss of every epoch."""
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        tot = 0.0
        t0 = time.perf_counter()
        for i in order:
            sy, r1, e1 = pairs[i]
            tot += nn.train_step(sy.graph, r1, e1, lr)
        hist.append(tot / len(pairs))
        if log:
            log("epoch %d: mean relative energy error %.4f (%d pairs, %.1fs)" % (ep, hist[-1], len(pairs),
                                                                                time.perf_counter() - t0))
    return hist


def mean_loss(nn, pairs):
    return float(np.mean([nn.loss_grad(sy.graph, r1, e1, want_grad=False)[0] for sy, r1, e1 in pairs]))


def solve_with(ctx, system, b_host, nn=None, omega=1.0, solver_kw=None, tol=1e-8, max_iter=400, unit_step=False):
    """FCG solve of one right-hand side with (nn) or without the learned
    post-smoother; returns (stats, x, seconds)."""
    s = G.Solver(ctx, solver="fcg", tol=tol, max_iter=max_iter, **(solver_kw or {}))
    if nn is not None:
        s.set_learned_smoother(nn, system.graph, omega, unit_step)
    if getattr(system, "hierarchy", None):
        gmg.attach(s, system.hierarchy)
    s.setup(system.A)
    b = torch.from_numpy(np.ascontiguousarray(b_host)).cuda()
    x = torch.zeros_like(b)
    s.solve(b, x)  # warm-up (captures the cycle)
    ctx.sync()
    t0 = time.perf_counter()
    st = s.solve(b, x)
    ctx.sync()
    return st, x, time.perf_counter() - t0
