#!/usr/bin/env python3
"""Time to 1e-8 relative residual of the condensed HHO face system on one GPU.

Default workload (one "step" = one solve from x = 0 to ||r||/||b|| < 1e-8
with the operator, the multigrid hierarchy and b already resident in HBM):
3-D diffusion -div(kappa grad u) = f on the unit cube, 96^3 hexahedra with
interior vertices perturbed by a seeded uniform +-0.15 h (trilinear cells,
warped faces), HHO k = 3 (10 unknowns per face, 2.63 M interior faces,
26.3 M unknowns), kappa lognormal per cell with sigma = 3, f = 3 pi^2 prod
sin(pi x_i).  The plain method is FP64 FCG(1) preconditioned by one cycle of
the p-first aggregation AMG: p-injection of every face block to its first
mode (the k = 0 face means), then double pairwise aggregation of that
scalar level down to at most 2500 unknowns (dense coarse solve), K-cycle on
aggregation levels 1-2 and V-cycle below, degree-1 Chebyshev smoothing of
D^-1 A on the finest level (D = block diagonal) and degree 4 below.  FCG,
not CG, because the K-cycle makes the preconditioner nonlinear.  Beside it:
the same FCG with an FP32 multigrid (FP32 vectors and level matrices, the
outer residual and dot products in FP64) and with the learned post-smoother.

--config c2 runs the tetrahedral workload (64^3 x 6 perturbed tetrahedra,
k = 2, kappa = 1), hetero and learned2d the other families.  Prints one JSON
line; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "s to 1e-8 rel. residual & GB/s: FP64 MG-PCG vs mixed/learned FCG, 1 GPU"
HBM_PEAK_GBS = 8000.0
T0 = time.time()


def log(*a):
    print("[bench %.1fs]" % (time.time() - T0), *a, file=sys.stderr, flush=True)


def spmv_bytes(nb, nf, nnzb, sv=8, sx=8):
    """Algorithmic bytes of one block-CSR SpMV: values, block column indices,
    row pointers, x read once and y written once."""
    return nnzb * nf * nf * sv + 4 * nnzb + 4 * (nb + 1) + 2 * nb * nf * sx


def elem_spmv_bytes(ncell, nb, nf, sv=8, sx=8, nfc=6):
    """Algorithmic bytes of one product on the element layout: every cell's
    matrix by its m/2 + 1 cyclic diagonals ((m/2 + 1) m values, m = nfc nf),
    its nfc block rows, x read once and y written once.  The two-colour
    scheme also reads x of every face twice and y once more (its actual
    traffic is counted by rocprofv3, not here)."""
    m = nfc * nf
    return ncell * (m // 2 + 1) * m * sv + 4 * ncell * nfc + 2 * nb * nf * sx


# rocprofv3 FETCH_SIZE / WRITE_SIZE summaries of the fine-level SpMV, per layout
TRAFFIC_SUMMARY = {"sliced": os.path.join(REPO, "profiles", "sliced_spmv_traffic_tet64_k2.md"),
                   "bcsr": os.path.join(REPO, "profiles", "bcsr_spmv_traffic_tet64_k2.md"),
                   "c5": os.path.join(REPO, "profiles", "c5_hex96_k3_fp64_traffic.md"),
                   "c5elem": os.path.join(REPO, "profiles", "c5_hex96_k3_elem_traffic.md")}


def measured_traffic(workload_key, summary, field="f64_bytes"):
    """HBM bytes per launch of a fine-level kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the same workload and
    kernel (scaled by the counters' reading for a streaming kernel of known
    size), or None."""
    import re
    try:
        text = open(TRAFFIC_SUMMARY[summary]).read()
    except OSError:
        return None
    m = re.search(r"<!-- traffic: (\{.*?\}) -->", text)
    if not m:
        return None
    d = json.loads(m.group(1))
    return d.get(field) if d.get("workload") == workload_key else None


def timed_solves(ctx, solver, b, x, steps, warmup):
    for _ in range(warmup):
        st = solver.solve(b, x)
    ctx.sync()
    ts = []
    for _ in range(steps):
        ctx.sync()
        t0 = time.perf_counter()
        st = solver.solve(b, x)
        ctx.sync()
        ts.append(time.perf_counter() - t0)
    return ts, st


class Variant:
    def __init__(self, prec="gmg", galerkin=True, smoother="bj", relax=0.5, nu=2, cycle="V", coarse=None,
                 coarse_relax=0.0):
        self.prec, self.galerkin, self.smoother, self.relax, self.nu, self.cycle = \
            prec, galerkin, smoother, relax, nu, cycle
        self.coarse, self.coarse_relax = coarse, coarse_relax

    def label(self):
        s = "%s%s %s(%d,%d) %s%s" % (self.prec, "-galerkin" if self.prec == "gmg" and self.galerkin else "",
                                     self.cycle, self.nu, self.nu, self.smoother,
                                     " w=%g" % self.relax if self.smoother == "bj" else "")
        if self.coarse:
            s += ", coarse levels %s%s" % (self.coarse, " w=%g" % self.coarse_relax if self.coarse == "bj" else "")
        return s


def make_solver(ctx, G, hier, A, v, solver="cg", mg_prec="f64", mg_store=None, max_iter=1000, coarse_store=None):
    from fhhos4_amd import gmg
    s = G.Solver(ctx, solver=solver, prec="mg" if v.prec == "gmg" else "aggregamg", cycle=v.cycle, nu=(v.nu, v.nu),
                 smoothers=(v.smoother, v.smoother), relax=v.relax, max_iter=max_iter, tol=1e-8, mg_prec=mg_prec,
                 mg_store=mg_store, galerkin=v.galerkin, coarse_smoother=v.coarse, coarse_relax=v.coarse_relax,
                 coarse_store=coarse_store, kcycle_levels=2 if v.prec == "amg" else 0,
                 theta=0.1 if v.prec == "amg" else 0.25, max_coarse=2500 if v.prec == "amg" else 1000)
    if v.prec == "gmg":
        gmg.attach(s, hier)
    t = time.perf_counter()
    s.setup(A)
    return s, time.perf_counter() - t


def cpu_baseline(args, G, O, v):
    """The oracle (CPU restatement of the same algorithm, OpenMP) on a
    bounded sample of the same workload family (smaller mesh)."""
    from fhhos4_amd import gmg
    n = args.cpu_n
    hctx = G.Context(-1)
    P = G.HhoProblem(dim=3, mesh="stetra", n=n, k=args.k, perturb=args.perturb, seed=args.seed, source="pisine")
    H = gmg.build_hierarchy(hctx, P)
    As = [H[0][1]]
    if v.galerkin:
        for l in range(len(H) - 1):
            As.append(G.galerkin_rap(hctx, As[-1], H[l][2]))
    else:
        As = [h[1] for h in H]
    levels = [(As[l].scipy(), H[l][2].scipy() if H[l][2] is not None else None) for l in range(len(H))]
    mc = lambda k: "mc" + k if k and "gs" in k else k  # the GPU's multicolour ordering
    sm = mc(v.smoother)
    x, hist, out = O.solve(levels[0][0], P.rhs(), "cg", "mg", nf=P.nf, cycle=v.cycle, smoothers=(sm, sm),
                           omega=v.relax, nu=(v.nu, v.nu), levels=levels, max_iter=args.max_iter,
                           coarse_smoother=mc(v.coarse), coarse_omega=v.coarse_relax)
    setup_s, solve_s = O.last_times()
    N = levels[0][0].shape[0]
    return {"N": N, "iterations": out.iterations, "solve_s": solve_s, "setup_s": setup_s,
            "dof_it_per_s": N * out.iterations / solve_s, "n": n}


def kernel_bench(args, G, np, torch):
    """Block-row SpMV of the fine operator alone (FP64, FP32, BF16-stored),
    timed with HIP events on the library stream, plus a device copy of known
    size (bandwidth reference and calibration of the HBM byte counters)."""
    P = G.HhoProblem(dim=3, mesh="stetra", n=args.n, k=args.k, perturb=args.perturb, seed=args.seed,
                     source="pisine", defer_assembly=True)
    nb, nf, nnzb = P.nb, P.nf, P.nnzb
    N = nb * nf
    ctx = G.Context(0)
    t = time.perf_counter()
    A64, _ = P.assemble_gpu(ctx, keep_host=False)
    ctx.sync()
    log("GPU assembly of %d cells: %.3f s" % (P.info["ncells"], time.perf_counter() - t))
    out = {"kernel": "bcsr_spmv", "n": args.n, "unknowns": N, "blocks": nnzb, "reps": args.reps}
    x = torch.rand(N, dtype=torch.float64, device="cuda")
    for name, A, xv, sv, sx in (("f64", A64, x, 8, 8), ("f32", A64.convert("f32"), x.float(), 4, 4),
                                ("bf16", A64.convert("bf16"), x.float(), 2, 4)):
        y = torch.zeros_like(xv)
        torch.cuda.synchronize()
        for _ in range(3):
            G.spmv(ctx, A, xv, y)
        ctx.sync()
        G.profile_begin(ctx, A, 0)
        for _ in range(args.reps):
            G.spmv(ctx, A, xv, y)
        nl, ms = G.profile_end(ctx)
        byt = spmv_bytes(nb, nf, nnzb, sv, sx)
        out[name] = {"mean_ms": ms, "launches": nl, "bytes": byt, "GBs": byt / (ms * 1e-3) / 1e9}
        log("SpMV %s: %.3f ms, %.0f GB/s" % (name, ms, out[name]["GBs"]))
        if A is not A64:
            A.close()
    # device-to-device copy of 4 GiB on torch's stream (the events below are on that stream)
    src = torch.empty(1 << 29, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.reps
    cb = 2 * src.numel() * 8
    out["copy"] = {"bytes": cb, "mean_ms": ms, "GBs": cb / (ms * 1e-3) / 1e9}
    log("copy 2 x %d MiB: %.3f ms, %.0f GB/s" % (src.numel() * 8 >> 20, ms, out["copy"]["GBs"]))
    print(json.dumps(out), flush=True)


def condense_bench(args, G, np, torch):
    """Batched static condensation (fh_condense_batched) of synthetic element
    matrices of the headline cell shape: hexahedra at k = 3, 20 cell unknowns
    and 6 faces x 10 face unknowns (local matrices 80 x 80), on an n^3 grid
    of cells (--cond-n) with the structured face numbering (boundary faces
    eliminated).  Local matrices: symmetric, R + R^T over 2 plus 80 I with R
    standard normal (seeded), so A_TT is positive definite.  Per cell the
    Schur product A_TF^T (A_TT^-1 A_TF) is 60 x 20 x 60 (144 kflop, on MFMA),
    the Cholesky and triangular solves about 52 kflop; 51 KB of local matrix
    are read and 36 blocks of 10 x 10 added."""
    n = args.cond_n
    nc, nfc, nf = 20, 6, 10
    nl = nc + nfc * nf
    ncells = n ** 3
    # faces: x-normal (n+1) n n, then y-normal, then z-normal; interior ones numbered
    idx = np.arange(ncells)
    cx, cy, cz = idx % n, (idx // n) % n, idx // (n * n)
    nx = (n + 1) * n * n
    xf = lambda i, y, z: (z * n + y) * (n + 1) + i
    yf = lambda x, j, z: nx + (z * (n + 1) + j) * n + x
    zf = lambda x, y, k: 2 * nx + (k * n + y) * n + x
    raw = np.stack([xf(cx, cy, cz), xf(cx + 1, cy, cz), yf(cx, cy, cz), yf(cx, cy + 1, cz), zf(cx, cy, cz),
                    zf(cx, cy, cz + 1)], axis=1)
    bnd = np.stack([cx == 0, cx == n - 1, cy == 0, cy == n - 1, cz == 0, cz == n - 1], axis=1)
    used = np.unique(raw[~bnd])
    renum = -np.ones(3 * nx, dtype=np.int64)
    renum[used] = np.arange(len(used))
    faces = np.where(bnd, -1, renum[raw]).astype(np.int32)
    nfaces = len(used)
    gen = torch.Generator(device="cuda").manual_seed(args.seed)
    local = torch.randn(ncells, nl, nl, dtype=torch.float64, device="cuda", generator=gen)
    local = 0.5 * (local + local.transpose(1, 2).contiguous())
    local += nl * torch.eye(nl, dtype=torch.float64, device="cuda")
    bt = torch.randn(ncells * nc, dtype=torch.float64, device="cuda", generator=gen)
    torch.cuda.synchronize()
    ctx = G.Context(0)
    G.condense_batched(ctx, faces[:64], local[:64], nc, nf, nfaces, B_T=bt[:64 * nc])  # warm-up (code objects)
    ts = []
    for _ in range(args.steps):
        t = time.perf_counter()
        M, b = G.condense_batched(ctx, faces, local, nc, nf, nfaces, B_T=bt)
        ctx.sync()
        ts.append(time.perf_counter() - t)
        del M, b
    flop = ncells * (2.0 * (nfc * nf) ** 2 * nc + nc ** 3 / 3.0 + 2.0 * nc * nc * (nfc * nf + 1))
    byt = ncells * (nl * nl * 8 + nc * 8 + nfc * nfc * nf * nf * 8 * 2)
    out = {"kernel": "condense_batched", "cells": ncells, "faces": nfaces, "nc": nc, "face_unknowns": nfc * nf,
           "call_s": float(np.median(ts)), "flop": flop, "bytes": byt,
           "note": "call time includes the host block pattern; the kernel's own duration is in the rocprof summary"}
    log("condensation of %d cells (%d faces): %.3f s per call (pattern on the host, kernel, right-hand side)"
        % (ncells, nfaces, out["call_s"]))
    print(json.dumps(out), flush=True)


def hetero_bench(args, G, np, torch):
    """Heterogeneous diffusion on Cartesian hexahedra (n^3 cells, HHO k): kappa
    piecewise constant per 8^3-cell subdomain, lognormal with sigma = 3
    (seeded, contrast ~1e6), one subdomain in eight anisotropic
    diag(1, 1, 1e-2).  FP64 CG with the re-discretised geometric multigrid
    against FP64 FCG with the FP32 / BF16-stored / BF16-coarse multigrid; the
    error of each against the FP64 solution.  (Galerkin operators need a
    full-rank prolongation, which CellInterp_Trace is not on hexahedra.)"""
    from fhhos4_amd import gmg
    P = G.HhoProblem(dim=3, mesh="cart", n=args.n, k=args.k, kappa="subdomain", sigma=3.0, subdomain=8, aniso=1e-2,
                     seed=args.seed, source="pisine", defer_assembly=True)
    N = P.n_unknowns
    ctx = G.Context(0)
    t = time.perf_counter()
    hier = gmg.build_hierarchy(ctx, P)
    log("hexahedral hierarchy", [h[0].n_unknowns for h in hier], "%.1fs" % (time.perf_counter() - t))
    A, b = hier[0][1], P.b_device
    x = torch.zeros_like(b)
    # --prec amg: the p-first aggregation AMG (K-cycle on two levels with
    # --cycle K, the coarse smoother of --coarse-smoother)
    amg = args.prec == "amg"
    v = Variant(args.prec, False, args.smoother, args.relax, args.nu, args.cycle if amg else "V",
                args.coarse_smoother if amg else None)
    krylov = "fcg" if v.cycle == "K" else "cg"
    s64, su = make_solver(ctx, G, hier, A, v, solver=krylov, max_iter=args.max_iter)
    ts, st = timed_solves(ctx, s64, b, x, args.steps, args.warmup)
    t64 = float(np.mean(ts))
    x64 = x.cpu().numpy()
    out = {"workload": "3-D heterogeneous diffusion, %d^3 hexahedra, HHO k=%d (%d unknowns), kappa lognormal per 8^3 "
                       "subdomain (sigma 3), 1/8 anisotropic diag(1,1,1e-2)" % (args.n, args.k, N),
           "solver": v.label(), "fp64_" + krylov: {"s": t64, "iterations": st["iterations"], "setup_s": su,
                                                  "converged": st["converged"]}}
    log("FP64 %s + %s: %d iterations, %.3f s" % (krylov.upper(), v.label(), st["iterations"], t64))
    del s64
    for store, cstore in (("f32", None), ("bf16", None), ("f32", "bf16")):
        key = store if cstore is None else "%s_coarse_%s" % (store, cstore)
        s, su = make_solver(ctx, G, hier, A, v, solver="fcg", mg_prec="f32", mg_store=store, coarse_store=cstore,
                            max_iter=args.max_iter)
        ts, st = timed_solves(ctx, s, b, x, args.steps, args.warmup)
        xm = x.cpu().numpy()
        out["fcg_" + key] = {"s": float(np.mean(ts)), "iterations": st["iterations"], "converged": st["converged"],
                             "rel_err_vs_fp64": float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)),
                             "speedup_vs_fp64": t64 / float(np.mean(ts))}
        log("FCG + %s MG: %d iterations, %.3f s, rel. error vs FP64 %.2e" % (key, st["iterations"], out["fcg_" + key]["s"],
                                                                            out["fcg_" + key]["rel_err_vs_fp64"]))
        del s
    print(json.dumps(out), flush=True)


def learned2d_bench(args, G, np, torch):
    """Learned post-smoother on 2-D perturbed quadrilaterals (--mesh2d quad;
    bilinear cells, HHO k = 2, per-cell lognormal kappa, sigma = 2; "stri":
    the perturbed triangles): trained on the --train-n2d mesh with random
    sinusoid right-hand sides (pairs from plain FCG solves), evaluated on the
    --n mesh with unseen kappa and right-hand-side seeds.  FCG with the
    multigrid without a post-step, with the block-Jacobi post-step (the
    untrained network) and with the trained network; energy-norm gain on
    held-out pairs of the evaluation mesh."""
    from fhhos4_amd import gmg
    from fhhos4_amd import learned as Lr
    ctx = G.Context(0)
    if args.prec2d == "aggregamg":
        # the p-first aggregation AMG with a K-cycle on the two top levels
        # (the hexahedral benchmark's preconditioner)
        solver = dict(prec="aggregamg", cycle=args.cycle2d, kcycle_levels=2, nu=(1, 1),
                      smoothers=("cheb2", "cheb2"), coarse_smoother="cheb1", fcg_truncation=args.learned_fcg_m)
        label = "FCG(%d) + p-first aggregation AMG %s(1,1)%s, cheb2 / cheb1" % (
            args.learned_fcg_m, args.cycle2d, " on levels 1-2" if args.cycle2d == "K" else "")
    else:
        solver = dict(prec="mg", cycle="V", nu=(1, 1), smoothers=("cheb2", "cheb2"), galerkin=True,
                      fcg_truncation=8)
        label = "FCG(8) + geometric MG V(1,1) cheb2, Galerkin coarse operators"

    def system(problem, tol=1e-12):
        hier = gmg.build_hierarchy(ctx, problem) if args.prec2d == "gmg" else None
        return Lr.System(ctx, problem, solver, tol=tol, max_iter=args.max_iter, hierarchy=hier)

    base = dict(dim=2, mesh=args.mesh2d, k=2, perturb=0.2, kappa="lognormal", sigma=2.0)
    t0 = time.perf_counter()
    train_sys = system(G.HhoProblem(n=args.train_n2d, seed=11, source="pisine", **base))
    rhs = []
    for j in range(args.train_rhs):
        rhs.append(G.HhoProblem(n=args.train_n2d, seed=11, source="modes", source_seed=1000 + j, **base).rhs())
        if j and j % 200 == 0:
            log("training right-hand sides: %d" % j)
    pairs = Lr.make_pairs(ctx, train_sys, rhs, iters=(1, 3, 6), log=log)
    t_pairs = time.perf_counter() - t0
    nf = train_sys.problem.nf
    nn = G.Mpnn(ctx, nf, layers=3, seed=1)
    before = Lr.mean_loss(nn, pairs)
    t1 = time.perf_counter()
    hist = Lr.train(nn, pairs, epochs=args.train_epochs, lr=2e-3, log=log)
    t_train = time.perf_counter() - t1
    log("trained on the %d^2 %s mesh: %d pairs (%.1fs), %.1fs, energy error %.3f -> %.3f"
        % (args.train_n2d, "quadrilateral" if args.mesh2d == "quad" else "triangle", len(pairs), t_pairs, t_train,
           before, hist[-1]))
    P = G.HhoProblem(n=args.n, seed=12, source="modes", source_seed=5000, **base)
    ev = system(P)
    held = Lr.make_pairs(ctx, ev, [G.HhoProblem(n=args.n, seed=12, source="modes", source_seed=6000 + j,
                                                **base).rhs() for j in range(2)], iters=(3,))
    untrained = G.Mpnn(ctx, nf, layers=3, seed=1)
    gain = {"untrained": Lr.mean_loss(untrained, held), "trained": Lr.mean_loss(nn, held), "pairs": len(held)}
    b = P.rhs()
    cells = "quadrilaterals" if args.mesh2d == "quad" else "triangles (2 per square)"
    out = {"workload": "2-D perturbed %s, %d^2 squares, HHO k=2 (%d unknowns), lognormal kappa sigma 2; "
                       "trained on %d^2 (%d pairs, %d epochs)" % (cells, args.n, P.n_unknowns, args.train_n2d,
                                                                  len(pairs), args.train_epochs),
           "solver": label,
           "train": {"pairs_s": t_pairs, "train_s": t_train, "energy_error_before": before, "after": hist[-1]},
           "held_out_energy_error": gain}
    xs = {}
    for key, net, unit in (("plain", None, False), ("block_jacobi_post", untrained, False),
                           ("learned", nn, False), ("learned_unit_step", nn, True)):
        st, x, sec = Lr.solve_with(ctx, ev, b, net, solver_kw=solver, max_iter=args.max_iter, unit_step=unit)
        xs[key] = x.cpu().numpy()
        out[key] = {"s": sec, "iterations": st["iterations"], "converged": st["converged"],
                    "accepted": st["nn_used"], "fallback": st["nn_fallback"]}
        log("%s: %d iterations, %.3f s" % (key, st["iterations"], sec))
    for key in xs:
        out[key]["rel_diff_vs_plain"] = float(np.linalg.norm(xs[key] - xs["plain"]) / np.linalg.norm(xs["plain"]))
    print(json.dumps(out), flush=True)


def learned_leg(args, G, ctx, hier, A, P, b, x, x64, t64, v):
    """FCG with the FP32 multigrid and the learned post-smoother.  The
    network is trained on the GPU (library kernels) on pairs from plain
    FCG solves of the same family on a smaller mesh (n = --train-n, random
    sinusoid right-hand sides), then runs on the full system."""
    import numpy as np
    import torch
    from fhhos4_amd import gmg
    from fhhos4_amd import learned as Lr

    t0 = time.perf_counter()
    base = dict(dim=3, mesh="stetra", n=args.train_n, k=args.k, perturb=args.perturb, seed=args.seed)
    Pt = G.HhoProblem(source="pisine", **base)
    Ht = gmg.build_hierarchy(ctx, Pt)
    kw = dict(prec="mg", cycle=v.cycle, nu=(v.nu, v.nu), smoothers=(v.smoother, v.smoother), relax=v.relax,
              galerkin=v.galerkin, coarse_smoother=v.coarse, coarse_relax=v.coarse_relax)

    class Sys(Lr.System):
        def __init__(self):
            self.problem, self.A, self.n = Pt, Ht[0][1], Pt.n_unknowns
            self.features = G.face_features(Pt)
            self.graph = G.NnGraph(ctx, self.A, self.features)
            self.solver = G.Solver(ctx, solver="fcg", tol=1e-12, max_iter=args.max_iter, **kw)
            gmg.attach(self.solver, Ht)
            self.solver.setup(self.A)
            self.tol, self.max_iter = 1e-12, args.max_iter

    sy = Sys()
    rhs = [G.HhoProblem(source="modes", source_seed=1000 + j, **base).rhs() for j in range(args.train_rhs)]
    pairs = Lr.make_pairs(ctx, sy, rhs, iters=(1, 3, 6))
    t_pairs = time.perf_counter() - t0
    nn = G.Mpnn(ctx, P.nf, layers=3, seed=1)
    before = Lr.mean_loss(nn, pairs)
    t1 = time.perf_counter()
    hist = Lr.train(nn, pairs, epochs=args.train_epochs, lr=2e-3, log=log)
    t_train = time.perf_counter() - t1
    log("learned smoother: %d pairs (%.1fs), training %.1fs, energy error %.3f -> %.3f"
        % (len(pairs), t_pairs, t_train, before, hist[-1]))
    feats = G.face_features(P)
    graph = G.NnGraph(ctx, A, feats)
    # FCG(8): the network makes the preconditioner vary from one application
    # to the next, and eight kept directions hold the A-orthogonality that
    # FCG(1) loses under it
    s = G.Solver(ctx, solver="fcg", tol=1e-8, max_iter=args.max_iter, mg_prec="f32", mg_store="f32",
                 fcg_truncation=args.learned_fcg_m, **kw)
    gmg.attach(s, hier)
    s.set_learned_smoother(nn, graph, 1.0)
    t2 = time.perf_counter()
    s.setup(A)
    su = time.perf_counter() - t2
    ts, st = timed_solves(ctx, s, b, x, args.steps, args.warmup)
    xm = x.cpu().numpy()
    out = {"s": float(np.mean(ts)), "iterations": st["iterations"], "setup_s": su, "mg": "f32",
           "accepted": st["nn_used"], "fallback": st["nn_fallback"], "params": nn.num_params(),
           "train": {"n": args.train_n, "pairs": len(pairs), "pairs_s": t_pairs, "train_s": t_train,
                     "epochs": args.train_epochs, "energy_error_untrained": before, "energy_error_trained": hist[-1]},
           "rel_err_vs_fp64": float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)),
           "l2": P.l2_error_gpu(ctx, x), "speedup_vs_fp64": t64 / float(np.mean(ts))}
    log("FCG + f32 MG + learned post-smoother: %d iterations (%d accepted, %d fallback), %.3f s"
        % (st["iterations"], st["nn_used"], st["nn_fallback"], out["s"]))
    del s, graph
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def c5_cpu_baseline(args, A, b, nf, gpu_iterations, base, krylov):
    """The oracle (CPU restatement of the same solver: p-first aggregation
    AMG, the same cycle and smoothers, OpenMP) on the full system: its own
    hierarchy set-up, then --cpu-iters iterations timed.  The oracle is
    iteration-exact with the GPU solver, so the time to tolerance is the
    measured seconds per iteration times the GPU's iteration count."""
    from oracle import oracle as O
    t = time.perf_counter()
    rp, ci, v = A.blocks()
    dl = time.perf_counter() - t
    log("CPU baseline: fine operator on the host (%.1f GB, %.1f s); oracle set-up and %d iterations"
        % (v.nbytes / 1e9, dl, args.cpu_iters))
    x, hist, out = O.solve_bcsr(rp, ci, v, nf, b, krylov, base["prec"], cycle=base["cycle"], nu=base["nu"],
                                smoothers=base["smoothers"], coarse_smoother=base["coarse_smoother"],
                                kcycle_levels=base["kcycle_levels"], theta=base["theta"],
                                max_coarse=base["max_coarse"], fcg_truncation=base["fcg_truncation"],
                                coarse_nu=base["coarse_nu"], max_iter=args.cpu_iters)
    setup_s, solve_s = O.last_times()
    del rp, ci, v
    per_it = solve_s / max(1, out.iterations)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))
    res = {"value": per_it * gpu_iterations, "unit": "s", "cores": cores, "kind": "port",
           "sample": "oracle %s + the same AMG on the full %d-unknown system, %s: set-up %.1f s, %d iterations "
                     "in %.2f s (%.3f s/iteration, residual %.3e after them); value = s/iteration x the %d "
                     "iterations of the GPU FP64 solve" % (krylov.upper(), len(b), cpu_model(), setup_s,
                                                            out.iterations, solve_s, per_it, hist[-1],
                                                            gpu_iterations)}
    log("CPU baseline:", res["sample"])
    return res


def c5_learned_leg(args, G, ctx, P, A, b, x, x64, t64, base, mixed_ref, l2_64):
    """FCG with the FP32 multigrid and the learned post-smoother on the
    hexahedral system.  The network (n_f = 10) is trained on the GPU by the
    library's kernels on (residual, error) pairs of plain FCG solves of the
    same family on a smaller mesh with another kappa field (--train-n, random
    sinusoid right-hand sides), then runs on the full system."""
    import numpy as np
    from fhhos4_amd import learned as Lr
    t0 = time.perf_counter()
    fam = dict(dim=3, mesh="hex", n=args.train_n, k=args.hex_k, perturb=args.hex_perturb, seed=args.seed + 1,
               kappa="lognormal", sigma=args.hex_sigma)
    train_kw = dict(base)
    train_kw.pop("tol", None)
    train_kw.pop("max_iter", None)
    sy = Lr.System(ctx, G.HhoProblem(source="pisine", **fam), train_kw, tol=1e-12, max_iter=args.max_iter)
    rhs = [G.HhoProblem(source="modes", source_seed=1000 + j, **fam).rhs() for j in range(args.train_rhs)]
    pairs = Lr.make_pairs(ctx, sy, rhs, iters=(1, 3, 6))
    t_pairs = time.perf_counter() - t0
    nn = G.Mpnn(ctx, P.nf, layers=3, seed=1)
    before = Lr.mean_loss(nn, pairs)
    t1 = time.perf_counter()
    hist = Lr.train(nn, pairs, epochs=args.train_epochs, lr=2e-3, log=log)
    t_train = time.perf_counter() - t1
    log("learned smoother (n_f %d, %d parameters): %d pairs (%.1fs), training %.1fs, energy error %.3f -> %.3f"
        % (P.nf, nn.num_params(), len(pairs), t_pairs, t_train, before, hist[-1]))
    graph = G.NnGraph(ctx, A, G.face_features(P))
    # per-application gain on the full system: held-out pairs at FCG
    # iterations 1 and 3 of the benchmark's own right-hand side (unseen
    # kappa field and right-hand side)
    held_sys = type("HeldOut", (), {})()
    held_sys.solver = G.Solver(ctx, **dict(base, solver="fcg", tol=1e-12, max_iter=args.max_iter))
    held_sys.solver.setup(A)
    held_sys.A, held_sys.graph, held_sys.n = A, graph, P.n_unknowns
    held_sys.tol, held_sys.max_iter = 1e-12, args.max_iter
    held = Lr.make_pairs(ctx, held_sys, [b.cpu().numpy()], iters=(1, 3))
    untrained = G.Mpnn(ctx, P.nf, layers=3, seed=1)
    held_gain = {"untrained": Lr.mean_loss(untrained, held), "trained": Lr.mean_loss(nn, held), "pairs": len(held)}
    log("held-out energy error on the full system: untrained (block-Jacobi step) %.3f, trained %.3f"
        % (held_gain["untrained"], held_gain["trained"]))
    del held, held_sys, untrained
    out = {}
    for m in sorted({args.fcg_m, args.learned_fcg_m}):
        s = G.Solver(ctx, **dict(base, solver="fcg", mg_prec="f32", mg_store="f32", fcg_truncation=m,
                                 outer_prec="f32", replace_every=args.replace_every))
        s.set_learned_smoother(nn, graph, 1.0)
        t2 = time.perf_counter()
        s.setup(A)
        su = time.perf_counter() - t2
        ts, st = timed_solves(ctx, s, b, x, args.leg_steps, 1)
        xm = x.cpu().numpy()
        out["fcg%d" % m] = {"s": float(np.mean(ts)), "iterations": st["iterations"], "setup_s": su,
                            "converged": st["converged"], "accepted": st["nn_used"], "fallback": st["nn_fallback"],
                            "rel_err_vs_fp64": float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)),
                            "speedup_vs_fp64": t64 / float(np.mean(ts)),
                            "vs_mixed_without": float(np.mean(ts)) / mixed_ref if mixed_ref else None}
        out["fcg%d" % m]["l2_rel_diff_vs_fp64"] = abs(P.l2_error_gpu(ctx, x) - l2_64) / l2_64
        log("FCG(%d) + f32 MG + learned: %d iterations (%d accepted, %d fallback), %.3f s"
            % (m, st["iterations"], st["nn_used"], st["nn_fallback"], out["fcg%d" % m]["s"]))
        del s
    out["train"] = {"n": args.train_n, "pairs": len(pairs), "pairs_s": t_pairs, "train_s": t_train,
                    "epochs": args.train_epochs, "energy_error_untrained": before, "energy_error_trained": hist[-1],
                    "params": nn.num_params(), "held_out_full_system": held_gain}
    del graph
    return out


def c5_bench(args, G, np, torch):
    """The hexahedral k = 3 workload (see the module docstring)."""
    from fhhos4_amd import gmg
    pm = tuple(int(t) for t in args.p_modes.split(","))
    log("mesh and pattern: %d^3 hexahedra, k=%d, perturb %.2f, lognormal kappa sigma %.1f"
        % (args.hex_n, args.hex_k, args.hex_perturb, args.hex_sigma))
    P = G.HhoProblem(dim=3, mesh="hex", n=args.hex_n, k=args.hex_k, perturb=args.hex_perturb, seed=args.seed,
                     kappa="lognormal", sigma=args.hex_sigma, source="pisine", defer_assembly=True)
    nb, nf, nnzb = P.nb, P.nf, P.nnzb
    N = nb * nf
    log("system: %d faces x %d = %d unknowns, %d blocks; cell quality >= %.3f (%d redrawn draws)"
        % (nb, nf, N, nnzb, P.info["min_quality"], P.info["redraws"]))
    ctx = G.Context(0)
    t = time.perf_counter()
    if args.c5_prec == "aggregamg" and not args.sweep:
        # the aggregation hierarchy needs only the fine operator
        H = {"A": gmg._operator(ctx, P, keep_host=False), "h": []}
        hier_s = time.perf_counter() - t
        log("system assembled on the GPU in %.1f s" % hier_s)
    else:
        H = gmg.build_hp_hierarchy(ctx, P, p_modes=pm)
        hier_s = time.perf_counter() - t
        log("hierarchy (GPU assembly, host k=0 prolongations) %.1f s; p-levels keep %s modes, h-levels %s"
            % (hier_s, pm, [h[0].n_unknowns for h in H["h"]]))
    A, b = H["A"], P.b_device
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    amg = args.c5_prec == "aggregamg"
    cnu = tuple(int(t) for t in args.c5_coarse_nu.split(","))
    base = dict(prec=args.c5_prec, nu=(args.nu, args.c5_nu_post), smoothers=(args.c5_smoother, args.c5_smoother),
                coarse_smoother=args.c5_coarse_smoother, galerkin=not amg, tol=1e-8, max_iter=args.max_iter,
                cycle=args.c5_cycle, kcycle_levels=args.kcycle_levels, theta=args.c5_theta,
                max_coarse=args.max_coarse, fcg_truncation=args.fcg_m, coarse_nu=cnu)
    if amg:
        label = ("p-first aggregation AMG (p-injection to the face means, double pairwise aggregation) %s-cycle%s, "
                 "finest level (%d,%d) %s, coarse levels (%d,%d) %s"
                 % (args.c5_cycle, " on levels 1-%d" % args.kcycle_levels if args.c5_cycle == "K" else "",
                    args.nu, args.c5_nu_post, args.c5_smoother, cnu[0], cnu[1], args.c5_coarse_smoother))
        label += ", strong-coupling threshold %g" % args.c5_theta
    else:
        label = "hp-MG %s-cycle p%s+Galerkin h(k=0), finest level (%d,%d) %s, coarse levels (%d,%d) %s" % (
            args.c5_cycle, args.p_modes, args.nu, args.c5_nu_post, args.c5_smoother, cnu[0], cnu[1],
            args.c5_coarse_smoother)
    krylov = "fcg" if args.c5_cycle == "K" else "cg"

    def solver(**kw):
        a = dict(base)
        a.update(kw)
        s = G.Solver(ctx, **a)
        if not amg:
            gmg.attach_hp(s, H)
        t0 = time.perf_counter()
        s.setup(A)
        return s, time.perf_counter() - t0

    if args.sweep:
        # FP64 CG with hierarchy / cycle / smoother variants: "key=value,..."
        # with p (modes kept, "3:1"), cycle, sm, csm, prec (mg | aggregamg)
        for spec in args.sweep_list.split(";"):
            kv = dict(t.split("=") for t in spec.split(",") if t)
            pmv = tuple(int(t) for t in kv.get("p", args.p_modes).replace(":", ",").split(","))
            prec = kv.get("prec", "mg")
            a = dict(base, prec=prec, cycle=kv.get("cycle", "V"), smoothers=(kv.get("sm", args.c5_smoother),) * 2,
                     coarse_smoother=kv.get("csm", args.c5_coarse_smoother), solver=kv.get("solver", "cg"))
            if "nu" in kv:
                a["nu"] = (int(kv["nu"]), int(kv["nu"]))
            if "cnu" in kv:  # coarse levels' (pre, post) counts, e.g. cnu=1:1
                a["coarse_nu"] = tuple(int(t) for t in kv["cnu"].split(":"))
            if "nu1" in kv or "nu2" in kv:
                a["nu"] = (int(kv.get("nu1", a["nu"][0])), int(kv.get("nu2", a["nu"][1])))
            if "outer" in kv:  # FCG's own product on the FP32 copy, residual replacement every re
                a.update(outer_prec=kv["outer"], replace_every=int(kv.get("re", args.replace_every)))
            if "m" in kv:
                a["fcg_truncation"] = int(kv["m"])
            if "kl" in kv:
                a["kcycle_levels"] = int(kv["kl"])
            if "mc" in kv:
                a["max_coarse"] = int(kv["mc"])
            if "theta" in kv:
                a["theta"] = float(kv["theta"])
            if "store" in kv:
                a.update(mg_prec="f32", mg_store=kv["store"])
            if "cstore" in kv:
                a.update(mg_prec="f32", coarse_store=kv["cstore"])
            A.set_elem(kv.get("elem", "1") != "0")  # elem=0: block-row kernels on the finest level
            try:
                Hv = H if pmv == pm or prec != "mg" else gmg.build_hp_hierarchy(ctx, P, p_modes=pmv)
                a["galerkin"] = prec == "mg"  # hp-MG h-levels are Galerkin products of the given P
                s = G.Solver(ctx, **a)
                if prec == "mg":
                    gmg.attach_hp(s, Hv)
                t0 = time.perf_counter()
                s.setup(A)
                su = time.perf_counter() - t0
                ts, st = timed_solves(ctx, s, b, x, 1, 1)
                out = {"variant": spec, "s": ts[0], "iterations": st["iterations"], "converged": st["converged"],
                       "setup_s": su, "ms_per_it": 1e3 * ts[0] / max(1, st["iterations"]),
                       "levels": [s.level_size(l) for l in range(s64_levels(s))]}
                del s
            except Exception as e:  # report and continue with the next variant
                out = {"variant": spec, "error": str(e)}
            log(json.dumps(out))
            print(json.dumps(out), flush=True)
        return
    log("set-up FP64 MG-PCG:", label)
    s64, setup64 = solver(solver=krylov)
    levels = [s64.level_size(l) for l in range(s64_levels(s64))]
    log("levels", levels, "setup %.1fs" % setup64)
    ts64, st64 = timed_solves(ctx, s64, b, x, args.steps, args.warmup)
    t64 = float(np.mean(ts64))
    x64 = x.cpu().numpy()
    t = time.perf_counter()
    l2_64 = P.l2_error_gpu(ctx, x)
    post_s = time.perf_counter() - t
    log("FP64 MG-PCG: %d iterations, %.3f s (%.1f ms/it), res %.2e, L2 %.4e, cycle %s"
        % (st64["iterations"], t64, 1e3 * t64 / max(1, st64["iterations"]), st64["final_res"], l2_64,
           {0: "eager", 1: "eager", 2: "graph replay"}[st64["graph_mode"]]))
    # in-situ durations of the fine-level kernels: SpMV (CG product and
    # residual kernel family) and the fused Chebyshev step
    prof = {}
    for mode, name in ((0, "spmv"), (1, "residual"), (3, "cheb_step")):
        G.profile_begin(ctx, A, mode)
        s64.solve(b, x)
        nl, ms = G.profile_end(ctx)
        prof[name] = (nl, ms)
    layout = {0: "bcsr", 1: "sliced", 2: "sliced2", 3: "chunked", 4: "sell", 5: "elem"}[A.layout()]
    del s64
    bytes_spmv = spmv_bytes(nb, nf, nnzb) if layout != "elem" else elem_spmv_bytes(P.info["ncells"], nb, nf)
    # fused Chebyshev step: the product's bytes + D^-1 (nb nf^2 doubles) +
    # b read + direction read and written + new iterate written
    bytes_cheb = bytes_spmv + nb * nf * nf * 8 + 4 * N * 8
    kern = {}
    for name, byt in (("spmv", bytes_spmv), ("residual", bytes_spmv + N * 8), ("cheb_step", bytes_cheb)):
        nl, ms = prof[name]
        if nl:
            kern[name] = {"launches_per_solve": nl, "mean_ms": ms, "bytes": byt, "GBs": byt / (ms * 1e-3) / 1e9,
                          "ms_per_solve": nl * ms}
            log("fine %s (%s): %d launches, %.3f ms, %.0f GB/s" % (name, layout, nl, ms, kern[name]["GBs"]))
    dom = max(kern, key=lambda k: kern[k]["ms_per_solve"])
    mixed = {}
    if not args.no_mixed:
        for spec in args.c5_mixed.split(","):
            # "f32", "f32:bf16" (coarse levels BF16), a trailing "o": FCG's own
            # product with the FP32 copy of the operator (residual replacement)
            outer = "f32" if spec.endswith("o") else "f64"
            spec = spec.rstrip("o")
            store, cstore = tuple(spec.split(":")) if ":" in spec else (spec, None)
            key = (store if cstore is None else "%s_coarse_%s" % (store, cstore)) + ("_outer_f32" if outer == "f32" else "")
            log("set-up FCG + %s multigrid" % key)
            s, su = solver(solver="fcg", mg_prec="f32", mg_store=store, coarse_store=cstore, outer_prec=outer,
                           replace_every=args.replace_every)
            ts, st = timed_solves(ctx, s, b, x, args.leg_steps, 1)
            xm = x.cpu().numpy()
            mixed[key] = {"s": float(np.mean(ts)), "iterations": st["iterations"], "setup_s": su,
                          "converged": st["converged"],
                          "rel_err_vs_fp64": float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)),
                          "l2_dist_to_sine": P.l2_error_gpu(ctx, x), "speedup_vs_fp64": t64 / float(np.mean(ts))}
            mixed[key]["l2_rel_diff_vs_fp64"] = abs(mixed[key]["l2_dist_to_sine"] - l2_64) / l2_64
            log("FCG + %s MG: %d iterations, %.3f s (x%.2f), rel. err vs FP64 %.1e"
                % (key, st["iterations"], mixed[key]["s"], t64 / mixed[key]["s"], mixed[key]["rel_err_vs_fp64"]))
            del s
    best = min(mixed.values(), key=lambda v: v["s"]) if mixed else None
    learned = None
    if not args.no_learned:
        learned = c5_learned_leg(args, G, ctx, P, A, b, x, x64, t64, base, best["s"] if best else None, l2_64)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = c5_cpu_baseline(args, A, b.cpu().numpy(), nf, st64["iterations"], base, krylov)
    # algorithmic bytes per iteration of the FP64 solve: the fine-level
    # kernels as launched (in-situ counts), the zero-guess block-Jacobi step
    # (D^-1, b, two outputs) and about 16 FP64 vectors of Krylov work; the
    # coarse levels (under 10% of the bytes) are not counted
    its = max(1, st64["iterations"])
    gb_iter = sum(v["launches_per_solve"] * v["bytes"] for v in kern.values()) / its
    gb_iter += nb * nf * nf * 8 + 3 * N * 8 + 16 * N * 8
    line = {
        "metric": METRIC, "value": t64, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t64 * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3-D diffusion, %d^3 trilinear hexahedra perturbed +-%.2fh, HHO k=%d (%d unknowns), "
                               "per-cell lognormal kappa sigma %.1f; %s + %s, tol 1e-8"
                               % (args.hex_n, args.hex_perturb, args.hex_k, N, args.hex_sigma,
                                  "FCG(%d)" % args.fcg_m if krylov == "fcg" else "CG", label),
                   "unknowns": N, "faces": nb, "blocks": nnzb, "levels": levels},
        "iterations": st64["iterations"], "setup_s": setup64, "assembly_s": hier_s,
        "cycle_execution": {0: "eager", 1: "eager", 2: "graph replay"}[st64["graph_mode"]],
        # no closed-form solution exists for the lognormal kappa field: the
        # discrete L2 distance of the reconstruction to prod sin(pi x_i) (the
        # kappa = 1 solution) is compared between the legs
        "l2_dist_to_sine": l2_64,
        "postprocess_s": post_s, "dof_it_per_s": N * st64["iterations"] / t64,
        "fine_kernels": kern, "fine_layout": layout,
        "mixed_fcg": mixed,
        "learned_fcg": learned,
        "best_mixed_vs_fp64": (best["s"] / t64) if best else None,
        "whole_solve_GBs": None if gb_iter is None else gb_iter * st64["iterations"] / t64 / 1e9,
        "roofline": {"bound": "hbm", "achieved": kern[dom]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kern[dom]["GBs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic("hex n=%d k=%d perturb %g sigma %g seed %d" % (
                         args.hex_n, args.hex_k, args.hex_perturb, args.hex_sigma, args.seed),
                         {"sliced": "c5", "elem": "c5elem"}[layout], dom)
                     if layout in ("sliced", "elem") else None,
                     "kernel": "fine-level %s (%s layout, n_f = %d)" % (dom, layout, nf),
                     "bytes_per_launch": kern[dom]["bytes"], "mean_ms": kern[dom]["mean_ms"],
                     "launches": kern[dom]["launches_per_solve"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--sweep-mixed", action="store_true", help="with --sweep: FCG + FP32 / BF16-coarse multigrid")
    ap.add_argument("--config", default="c5", choices=["c5", "c2", "hetero", "learned2d"],
                    help="hetero: the heterogeneous hexahedral case (64^3 hexahedra, k=2, subdomain lognormal kappa); "
                         "learned2d: the learned post-smoother on 2-D triangles (use --n 512)")
    ap.add_argument("--train-n2d", type=int, default=128)
    ap.add_argument("--learned-fcg-m", type=int, default=8, help="FCG truncation of the learned leg")
    ap.add_argument("--prec2d", default="aggregamg", choices=["gmg", "aggregamg"])
    ap.add_argument("--mesh2d", default="quad", choices=["quad", "stri"])
    ap.add_argument("--cycle2d", default="K", choices=["V", "K"])
    ap.add_argument("--k", type=int, default=2)
    ap.add_argument("--perturb", type=float, default=0.2)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--prec", default="gmg", choices=["gmg", "amg"])
    ap.add_argument("--no-galerkin", action="store_true")
    ap.add_argument("--smoother", default="cheb2")
    ap.add_argument("--relax", type=float, default=1.0)
    ap.add_argument("--nu", type=int, default=1)
    ap.add_argument("--cycle", default="V")
    ap.add_argument("--coarse-smoother", default="cheb1", help="smoother of the levels below the finest")
    ap.add_argument("--coarse-relax", type=float, default=0.0)
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--cpu-n", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--train-n", type=int, default=16, help="mesh of the learned smoother's training problems")
    ap.add_argument("--train-rhs", type=int, default=100)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--sweep", action="store_true", help="time a list of solver variants")
    ap.add_argument("--sweep-list", default="cycle=V;cycle=W;p=3:1;p=3:1,cycle=W;sm=cheb3;cycle=W,sm=cheb3;"
                                            "prec=aggregamg,cycle=K,solver=fcg;prec=aggregamg,cycle=V",
                    help="c5 --sweep: variants separated by ';'")
    ap.add_argument("--kernel", choices=["spmv", "condense"], help="time one kernel family alone")
    ap.add_argument("--cond-n", type=int, default=48, help="--kernel condense: cells per direction")
    ap.add_argument("--hex-n", type=int, default=96, help="c5: hexahedra per direction")
    ap.add_argument("--hex-k", type=int, default=3)
    ap.add_argument("--hex-perturb", type=float, default=0.15)
    ap.add_argument("--hex-sigma", type=float, default=3.0)
    ap.add_argument("--p-modes", default="1", help="c5: modes kept by the p-levels, e.g. 1 or 3,1")
    ap.add_argument("--leg-steps", type=int, default=2, help="timed solves of each mixed / learned leg")
    ap.add_argument("--cpu-iters", type=int, default=3, help="c5: oracle iterations timed on the full system")
    ap.add_argument("--c5-prec", default="aggregamg", choices=["aggregamg", "mg"],
                    help="c5 multigrid: p-first aggregation AMG, or the geometric p-then-h hierarchy")
    ap.add_argument("--c5-cycle", default="K", choices=["V", "W", "K"])
    ap.add_argument("--c5-smoother", default="cheb1", help="c5: smoother of the finest level")
    ap.add_argument("--c5-coarse-smoother", default="cheb4", help="c5: smoother of the levels below the finest")
    ap.add_argument("--c5-theta", type=float, default=0.1, help="c5: strong-coupling threshold of the aggregation")
    ap.add_argument("--c5-nu-post", type=int, default=0,
                    help="c5: post-smoothing steps of the finest level (pre-smoothing: --nu)")
    ap.add_argument("--c5-coarse-nu", default="1,1", help="c5: pre,post smoothing steps below the finest level")
    ap.add_argument("--fcg-m", type=int, default=4,
                    help="c5: FCG truncation (directions kept) of the FP64 and mixed legs")
    ap.add_argument("--max-coarse", type=int, default=2500, help="c5: largest coarsest level (dense solve)")
    ap.add_argument("--replace-every", type=int, default=12,
                    help="c5 FP32 outer product: iterations between residual replacements")
    ap.add_argument("--kcycle-levels", type=int, default=2, help="K-cycle on levels 1..this one, V below")
    ap.add_argument("--c5-mixed", default="f32,f32o",
                    help="c5 mixed legs: fine[:coarse] storage, comma-separated; a trailing o: FP32 outer product")
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import fhhos4_amd as F
    from fhhos4_amd import gmg

    G = F.fhgpu
    assert torch.cuda.is_available(), "no ROCm device"
    if args.config == "hetero":
        hetero_bench(args, G, np, torch)
        return
    if args.config == "learned2d":
        learned2d_bench(args, G, np, torch)
        return
    if args.kernel == "spmv":
        return kernel_bench(args, G, np, torch)
    if args.kernel == "condense":
        return condense_bench(args, G, np, torch)
    if args.config == "c5":
        return c5_bench(args, G, np, torch)
    log("mesh and pattern of the HHO system n=%d k=%d (perturb %.2f)" % (args.n, args.k, args.perturb))
    P = G.HhoProblem(dim=3, mesh="stetra", n=args.n, k=args.k, perturb=args.perturb, seed=args.seed,
                     source="pisine", defer_assembly=True)
    nb, nf, nnzb = P.nb, P.nf, P.nnzb
    N = nb * nf
    log("system: %d faces x %d = %d unknowns, %d blocks" % (nb, nf, N, nnzb))
    ctx = G.Context(0)
    t = time.perf_counter()
    hier = gmg.build_hierarchy(ctx, P)  # GPU local operators + condensation on every level
    hier_s = time.perf_counter() - t
    A = hier[0][1]
    log("geometric hierarchy", [h[0].n_unknowns for h in hier], "%.1fs (GPU assembly, host transfers)" % hier_s)
    b = P.b_device
    x = torch.zeros_like(b)
    torch.cuda.synchronize()

    if args.sweep:
        variants = [Variant("gmg", True, "cheb2", 1.0, 1), Variant("gmg", True, "cheb3", 1.0, 1),
                    Variant("gmg", False, "cheb2", 1.0, 1), Variant("gmg", False, "cheb3", 1.0, 1),
                    Variant("gmg", False, "cheb4", 1.0, 1), Variant("gmg", False, "cheb2", 1.0, 2),
                    Variant("gmg", False, "cheb2", 1.0, 1, "W")]
        # FP64 CG, or (--sweep-mixed) FCG with the FP32 / BF16-coarse multigrid
        mk = (dict(solver="fcg", mg_prec="f32", mg_store="f32", coarse_store="bf16") if args.sweep_mixed else {})
        for v in variants:
            log("variant", v.label(), "mixed" if mk else "fp64")
            try:
                s, su = make_solver(ctx, G, hier, A, v, max_iter=args.max_iter, **mk)
                log("  set-up %.1fs" % su)
                ts, st = timed_solves(ctx, s, b, x, args.steps, args.warmup)
                print(json.dumps({"variant": v.label(), "s": float(np.mean(ts)), "iterations": st["iterations"],
                                  "converged": st["converged"], "setup_s": su, "levels": st["levels"],
                                  "op_complexity": st["op_complexity"]}), flush=True)
                del s
            except Exception as e:  # report and continue with the next variant
                print(json.dumps({"variant": v.label(), "error": str(e)}), flush=True)
        return

    v = Variant(args.prec, not args.no_galerkin, args.smoother, args.relax, args.nu, args.cycle, args.coarse_smoother,
                args.coarse_relax)
    log("set-up FP64 MG-PCG:", v.label())
    # the K-cycle is a nonlinear preconditioner: flexible CG
    s64, setup64 = make_solver(ctx, G, hier, A, v, solver="fcg" if v.cycle == "K" else "cg", max_iter=args.max_iter)
    levels = [s64.level_size(l) for l in range(s64_levels(s64))]
    log("levels", levels, "setup %.1fs" % setup64)
    ts64, st64 = timed_solves(ctx, s64, b, x, args.steps, args.warmup)
    t64 = float(np.mean(ts64))
    x64 = x.cpu().numpy()
    t = time.perf_counter()
    l2_64 = P.l2_error_gpu(ctx, x)  # cell recovery + reconstruction + quadrature on the GPU
    post_s = time.perf_counter() - t
    log("FP64 MG-PCG: %d iterations, %.3f s, res %.2e, L2 %.4e" % (st64["iterations"], t64, st64["final_res"], l2_64))

    # in-situ duration of the fine-level SpMV (dominant kernel family)
    G.profile_begin(ctx, A, 0)
    s64.solve(b, x)
    nl, spmv_ms = G.profile_end(ctx)
    bytes_fine = spmv_bytes(nb, nf, nnzb)
    # the fine operator runs on its sliced copy unless FHGPU_SLICED=0
    sliced = os.environ.get("FHGPU_SLICED", "1") != "0" and A.set_sliced(True)
    kernel = ("k_sliced<6,double,double,double,OP_SPMV>" if sliced else
              "k_bcsr_grp<6,double,double,double,OP_SPMV>") + " (fine level)"
    achieved = bytes_fine / (spmv_ms * 1e-3) / 1e9
    log("fine SpMV: %d launches, %.3f ms, %.0f GB/s" % (nl, spmv_ms, achieved))
    G.profile_begin(ctx, A, 3)  # the fused Chebyshev step of the fine level
    s64.solve(b, x)
    nls, smooth_ms = G.profile_end(ctx)
    del s64

    mixed = {}
    if not args.no_mixed:
        # storage of the fine level / of the coarse levels
        for store, cstore in (("f32", None), ("bf16", None), ("f32", "bf16")):
            key = store if cstore is None else "%s_coarse_%s" % (store, cstore)
            log("set-up FCG + %s multigrid" % key)
            s, su = make_solver(ctx, G, hier, A, v, solver="fcg", mg_prec="f32", mg_store=store,
                                max_iter=args.max_iter, coarse_store=cstore)
            ts, st = timed_solves(ctx, s, b, x, args.steps, args.warmup)
            xm = x.cpu().numpy()
            mixed[key] = {"s": float(np.mean(ts)), "iterations": st["iterations"], "setup_s": su,
                          "rel_err_vs_fp64": float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)),
                          "l2": P.l2_error_gpu(ctx, x), "speedup_vs_fp64": t64 / float(np.mean(ts))}
            log("FCG + %s MG: %d iterations, %.3f s" % (key, st["iterations"], mixed[key]["s"]))
            del s

    learned = None
    if not args.no_learned:
        learned = learned_leg(args, G, ctx, hier, A, P, b, x, x64, t64, v)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        log("CPU baseline (oracle, n=%d)" % args.cpu_n)
        c = cpu_baseline(args, G, O, v)
        est = N * st64["iterations"] / c["dof_it_per_s"]
        cpu = {"value": est, "unit": "s", "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
               "kind": "port",
               "sample": "oracle CG + GMG %s on %d^3x6 perturbed tets (N=%d, %d it, %.2f s solve, %.3g DOF.it/s); "
                         "value = full workload N=%d x %d it at that DOF.it/s"
                         % (v.label(), c["n"], c["N"], c["iterations"], c["solve_s"], c["dof_it_per_s"], N,
                            st64["iterations"])}
        log("CPU baseline:", cpu["sample"])

    line = {
        "metric": METRIC, "value": t64, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t64 * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3-D Poisson, %d^3 cubes x 6 tets perturbed +-%.2fh, HHO k=%d (%d unknowns), "
                               "%s + %s, tol 1e-8" % (args.n, args.perturb, args.k, N,
                                                      "FCG" if v.cycle == "K" else "CG", v.label()),
                   "unknowns": N, "faces": nb, "blocks": nnzb, "levels": levels},
        "iterations": st64["iterations"], "setup_s": setup64, "hierarchy_s": hier_s, "l2_error": l2_64,
        "postprocess_s": post_s,
        "dof_it_per_s": N * st64["iterations"] / t64,
        "mixed_fcg": mixed,
        "learned_fcg": learned,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("stetra n=%d k=%d perturb %g seed %d" % (args.n, args.k, args.perturb,
                                                                                         args.seed),
                                                 "sliced" if sliced else "bcsr"),
                     "kernel": kernel,
                     "bytes_per_launch": bytes_fine, "mean_ms": spmv_ms, "launches": nl,
                     "cheb_step_mean_ms": smooth_ms, "cheb_step_launches": nls},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def s64_levels(s):
    n = 0
    while True:
        try:
            s.level_size(n)
            n += 1
        except Exception:
            return n


if __name__ == "__main__":
    main()
