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


def _operator(ctx, problem):
    """Level operator: assembled on the GPU for problems created with
    defer_assembly (the right-hand side is kept as problem.b_device), else
    uploaded from the host assembly."""
    if problem.args.get("defer_assembly"):
        A, b = problem.assemble_gpu(ctx, keep_host=True)
        problem.b_device = b
        return A
    return problem.matrix(ctx, keep_host=True)


def attach(solver, hierarchy):
    """Gives the hierarchy to a solver created with prec="mg"."""
    for l, (p, A, P) in enumerate(hierarchy):
        solver.set_level(l, A if l > 0 else None, P)
