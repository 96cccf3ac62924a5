"""Geometric multigrid hierarchy of an HHO problem on nested structured meshes
(standard coarsening n -> n/2, MultigridForHHO with re-discretised coarse
operators and CellInterp_Trace prolongations)."""
from . import fhgpu as G


def build_hierarchy(ctx, problem, max_coarse=1000, keep=None):
    """Returns [(problem_l, A_l, P_l)] from fine to coarse; P_l maps level
    l+1 to level l (None on the coarsest).  Coarsens while the level has more
    than max_coarse unknowns and n is even."""
    levels = [(problem, _operator(ctx, problem))]
    while levels[-1][0].n_unknowns > max_coarse and levels[-1][0].args["n"] % 2 == 0 and levels[-1][0].args["n"] > 1:
        coarse = levels[-1][0].coarsened()
        levels.append((coarse, _operator(ctx, coarse)))
    out = []
    for l, (p, A) in enumerate(levels):
        P = p.prolongation(ctx, levels[l + 1][0]) if l + 1 < len(levels) else None
        out.append((p, A, P))
    return out


def _operator(ctx, problem, keep_host=True):
    """Level operator: assembled on the GPU for problems created with
    defer_assembly (the right-hand side is kept as problem.b_device), else
    uploaded from the host assembly.  keep_host keeps a host copy of the
    values (needed by host set-up steps that read the operator)."""
    if problem.args.get("defer_assembly"):
        A, b = problem.assemble_gpu(ctx, keep_host=keep_host)
        problem.b_device = b
        return A
    return problem.matrix(ctx, keep_host=True)


def attach(solver, hierarchy):
    """Gives the hierarchy to a solver created with prec="mg"."""
    for l, (p, A, P) in enumerate(hierarchy):
        solver.set_level(l, A if l > 0 else None, P)


def build_hp_hierarchy(ctx, problem, p_modes=(1,), h_degree=0, max_coarse=1000):
    """p-then-h hierarchy of an HHO problem (the reference's "-hp-cs p_h"
    strategy, Multigrid.h:353-372): the fine operator, p-levels on the same
    mesh keeping p_modes[i] modes of every face (natural injection of the
    hierarchical face basis, P_MultigridForHHO.h:62-86; their Galerkin
    operators, the leading sub-blocks, are formed by the solver), then
    h-levels of the degree-h_degree problem on the coarsened meshes with
    CellInterp_Trace prolongations (LevelForHHO.h:182-213), whose operators
    the solver forms as Galerkin products.

    Returns a dict: "A" the fine operator (GPU assembly for problems created
    with defer_assembly; its right-hand side is kept as problem.b_device),
    "p_modes", and "h": [(problem_l, P_l)] from the finest mesh of the
    h-levels down, P_l mapping level l+1 to level l (None on the coarsest)."""
    args = dict(problem.args)
    nf_h = (h_degree + 1) if args["dim"] == 2 else (h_degree + 1) * (h_degree + 2) // 2
    if not p_modes or p_modes[-1] != nf_h:
        raise ValueError("the last p-level must keep the %d modes of the degree-%d h-levels" % (nf_h, h_degree))
    A = _operator(ctx, problem, keep_host=False)  # the p-levels gather their sub-blocks on the device
    kw = dict(args)
    kw.update(k=h_degree, defer_assembly=True, ref_n=args.get("ref_n") or args["n"])
    chain = [G.HhoProblem(**kw)]
    while chain[-1].n_unknowns > max_coarse and chain[-1].args["n"] % 2 == 0 and chain[-1].args["n"] > 1:
        chain.append(chain[-1].coarsened())
    h = []
    for l, p in enumerate(chain):
        P = p.prolongation(ctx, chain[l + 1]) if l + 1 < len(chain) else None
        h.append((p, P))
    return {"A": A, "p_modes": tuple(p_modes), "h": h}


def attach_hp(solver, H):
    """Gives a p-then-h hierarchy to a solver created with prec="mg" and
    galerkin=True."""
    npl = len(H["p_modes"])
    for l, m in enumerate(H["p_modes"]):
        solver.set_level_p(l, m)
    for j, (p, P) in enumerate(H["h"]):
        solver.set_level(npl + j, None, P)


def hp_levels_host(H, A_fine_host, nf):
    """The same hierarchy as host scipy matrices for the oracle: [(A_l, P_l
    or ("p", nfc), nf_l)], with the Galerkin operators formed on the host
    (library's host Galerkin product on a host-only context)."""
    import numpy as np
    import scipy.sparse as sp
    levels = []
    A = A_fine_host.tocsr()
    cur_nf = nf
    for m in H["p_modes"]:
        nb = A.shape[0] // cur_nf
        Ab = sp.bsr_matrix(A, blocksize=(cur_nf, cur_nf))
        Ab.sort_indices()
        sub = sp.bsr_matrix((np.ascontiguousarray(Ab.data[:, :m, :m]), Ab.indices, Ab.indptr),
                            shape=(nb * m, nb * m)).tocsr()
        levels.append((A, ("p", m), cur_nf))
        A, cur_nf = sub, m
    hctx = G.Context(-1)
    for j, (p, P) in enumerate(H["h"]):
        if P is None:
            levels.append((A, None, cur_nf))
            break
        Ps = P.scipy()
        levels.append((A, Ps, cur_nf))
        Am = G.Matrix.from_csr(hctx, A, cur_nf, keep_host=True)
        Pm = G.Matrix.from_csr(hctx, Ps, cur_nf, keep_host=True)
        A = G.galerkin_rap(hctx, Am, Pm).scipy()
    return levels
