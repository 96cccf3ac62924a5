"""General (bilinear / trilinear) cells of the HHO producer: the "hex" and
"quad" meshes, which carry the hexahedral k = 3 benchmark family.

Pinned by:
* the reference's own condensed operators on unperturbed Cartesian meshes
  (its Cube_CartesianMesh / Square_CartesianMesh with the Legendre face
  basis), including a per-cell lognormal kappa given to the reference as one
  physical part per cell: entries to 1e-13 relative, the same block pattern,
  the right-hand side to its quadrature accuracy and the L2 error of the
  reference's own solution;
* on perturbed (warped-face) meshes, which the reference cannot build: the
  optimal L2 convergence order k + 2 of the method on meshes whose
  distortion does not shrink with h, symmetry, and the quality floor of the
  vertex perturbation.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from conftest import csr, load

import fhhos4_amd as F

G = F.fhgpu


def _compare(P, d, nf):
    A = csr(d, "A")
    M = P.scipy()
    assert M.shape == A.shape
    assert abs(M - A).max() <= 1e-13 * abs(A).max()
    Ab = sp.bsr_matrix(A, blocksize=(nf, nf))
    Ab.eliminate_zeros()
    rp, ci, _ = P.blocks()
    assert np.array_equal(Ab.indptr, rp) and np.array_equal(Ab.indices, ci)
    b = d["b"]
    assert np.abs(P.rhs() - b).max() <= 1e-8 * np.abs(b).max()  # source quadrature only
    l2 = P.l2_error(d["x"])
    assert abs(l2 - float(d["info_l2"])) <= 1e-6 * float(d["info_l2"])


def test_hex_k2_equals_reference_cartesian_operator():
    d = load("cart8_k2_cg_mg_bj23")
    P = G.HhoProblem(dim=3, mesh="hex", n=8, k=2, source="pisine")
    _compare(P, d, 6)


@pytest.mark.parametrize("name,dim,mesh,n,k,sigma,seed", [
    ("hex4_k3_lognormal", 3, "hex", 4, 3, 3.0, 7),
    ("quad8_k2_lognormal", 2, "quad", 8, 2, 2.0, 5),
])
def test_lognormal_kappa_equals_reference(name, dim, mesh, n, k, sigma, seed):
    d = load(name)
    P = G.HhoProblem(dim=dim, mesh=mesh, n=n, k=k, kappa="lognormal", sigma=sigma, seed=seed, source="pisine")
    assert np.array_equal(P.cell_kappa(), d["kappa"])
    _compare(P, d, P.nf)


@pytest.mark.parametrize("dim,mesh,k,ns", [(3, "hex", 1, (4, 8)), (3, "hex", 2, (4, 8)), (2, "quad", 2, (8, 16, 32)),
                                           (2, "quad", 3, (8, 16))])
def test_perturbed_cells_converge_at_order_k_plus_2(dim, mesh, k, ns):
    errs = []
    for n in ns:
        P = G.HhoProblem(dim=dim, mesh=mesh, n=n, k=k, perturb=0.15, seed=3, source="pisine")
        A = P.scipy()
        assert abs(A - A.T).max() <= 1e-13 * abs(A).max()
        errs.append(P.l2_error(spl.spsolve(A.tocsc(), P.rhs())))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates.min() > k + 1.6, (errs, rates)


@pytest.mark.parametrize("mesh,dim,n,perturb", [("stetra", 3, 32, 0.2), ("hex", 3, 24, 0.15), ("stri", 2, 64, 0.2),
                                                ("quad", 2, 64, 0.3)])
def test_perturbation_keeps_cells_above_the_quality_floor(mesh, dim, n, perturb):
    P = G.HhoProblem(dim=dim, mesh=mesh, n=n, k=0, perturb=perturb, seed=2026, defer_assembly=True)
    assert P.info["min_quality"] >= 0.1
    assert P.info["kept_vertices"] == 0
    if mesh == "stetra":
        assert P.info["redraws"] > 0  # +-0.2h per coordinate draws some inverted tetrahedra


@pytest.mark.parametrize("k", [0, 2])
def test_prolongation_reproduces_constants_inside(k):
    """CellInterp_Trace on warped cells: coarse face values 1 (mode 0 of the
    Legendre face basis is the constant) make every coarse cell whose faces
    are all interior reconstruct the constant 1, whose trace on the fine
    faces is mode 0 = 1, higher modes 0 (weights of a shared face sum to 1)."""
    n = 8
    fine = G.HhoProblem(dim=3, mesh="hex", n=n, k=k, perturb=0.15, seed=4, source="pisine")
    coarse = fine.coarsened()
    assert coarse.args["ref_n"] == n and coarse.args["n"] == n // 2
    Pm = fine.prolongation(G.Context(-1), coarse).scipy()
    nf = fine.nf
    vc = np.zeros(coarse.n_unknowns)
    vc[0::coarse.nf] = 1.0
    vf = (Pm @ vc).reshape(-1, nf)
    fc = fine.face_cells()
    nc = n // 2

    def inner(cell):  # coarse cell of a fine cell, away from the boundary
        ix, iy, iz = cell % n, (cell // n) % n, cell // (n * n)
        return all(1 <= c // 2 <= nc - 2 for c in (ix, iy, iz))

    rows = [r for r in range(fine.nb) if inner(fc[r, 0]) and inner(fc[r, 1])]
    assert len(rows) > 100
    assert np.abs(vf[rows, 0] - 1).max() < 1e-12
    if nf > 1:
        assert np.abs(vf[rows, 1:]).max() < 1e-12


@pytest.mark.parametrize("name,kw,nf", [
    ("c1_n64_cg_mg_bj", dict(dim=2, mesh="quad", n=64, k=1, source="pisine"), 2),
    ("cart8_k2_cg_mg_bj23", dict(dim=3, mesh="hex", n=8, k=2, source="pisine"), 6),
])
def test_geometric_hierarchy_equals_reference_levels(name, kw, nf):
    """The CellInterp_Trace prolongation (LevelForHHO.h:182-213, 743-800,
    899-954) and the re-discretised coarse operators of the general-cell
    producer against the reference's own exported levels (level<l>_P,
    level<l>_A of its -export mg) on unperturbed Cartesian meshes: entries to
    1e-13 relative on every level, and the oracle's CG + V(1,1) block-Jacobi
    solve on this hierarchy takes the reference's iteration count."""
    from oracle import oracle as O
    d = load(name)
    probs = [G.HhoProblem(**kw)]
    levels = []
    for l in range(int(d["nlevels"]) - 1):
        coarse = probs[-1].coarsened()
        Pm = probs[-1].prolongation(G.Context(-1), coarse).scipy()
        Pr = csr(d, "level%d_P" % l)
        assert Pm.shape == Pr.shape
        assert abs(Pm - Pr).max() <= 1e-13 * abs(Pr).max(), l
        Ac, Ar = coarse.scipy(), csr(d, "level%d_A" % (l + 1))
        assert abs(Ac - Ar).max() <= 1e-13 * abs(Ar).max(), l
        levels.append((probs[-1].scipy(), Pm))
        probs.append(coarse)
    levels.append((probs[-1].scipy(), None))
    sm = "bj23" if "bj23" in name else "bj"
    x, hist, out = O.solve(probs[0].scipy(), d["b"], "cg", "mg", nf=nf, smoothers=(sm, sm), levels=levels)
    assert out.iterations == int(d["info_iterations"])
    assert np.linalg.norm(x - d["x"]) <= 1e-8 * np.linalg.norm(d["x"])
