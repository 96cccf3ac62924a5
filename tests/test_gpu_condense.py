"""Batched static condensation and cell recovery on the GPU, through the
C-ABI (fh_condense_batched, fh_cell_recover_batched), fed with the
reference's own element matrices (Diff_HHOElement::A dumped by the
reference driver): the condensed operator and right-hand side equal the
reference's A and b (Diffusion_HHO.h:624-637, :1026-1032), and the
recovered cell unknowns satisfy every cell's equations
A_TT x_T + A_TF x_F = B_T (SolveCellUnknowns, Diffusion_HHO.h:960-966)."""
import numpy as np
import pytest

from conftest import csr, load
from test_condense import CASES, elements

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
F = pytest.importorskip("fhhos4_amd")
G = F.fhgpu


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "no ROCm device visible"
    return G.Context(0)


@pytest.mark.parametrize("name", CASES)
def test_condense_and_recover_match_reference(ctx, name):
    import scipy.sparse.linalg as spla
    d = load(name)
    faces, local, BT, nc, nf, nfaces = elements(d)
    print("case %s: %d cells, %d faces per cell, nc %d, nf %d, %d faces" % (name, faces.shape[0], faces.shape[1], nc,
                                                                            nf, nfaces), flush=True)
    loc = torch.from_numpy(np.ascontiguousarray(local)).cuda()
    bt = torch.from_numpy(np.ascontiguousarray(BT)).cuda()
    M, b = G.condense_batched(ctx, faces, loc, nc, nf, nfaces, B_T=bt)
    A = csr(d, "A")
    S = M.scipy()
    scale = np.abs(A.data).max()
    assert np.abs((S - A).toarray()).max() <= 1e-12 * scale
    assert np.abs(b.cpu().numpy() - d["b"]).max() <= 1e-12 * np.abs(d["b"]).max()
    # the block pattern: every pair of faces of one cell, ascending columns
    rp, ci, _ = M.blocks()
    pairs = set()
    for row in faces:
        for f1 in row[row >= 0]:
            for f2 in row[row >= 0]:
                pairs.add((int(f1), int(f2)))
    got = {(r, int(c)) for r in range(nfaces) for c in ci[rp[r]:rp[r + 1]]}
    assert got == pairs
    assert all(np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0) for r in range(nfaces))
    # B_F and the matrix are also used by a second call (no state is kept)
    bf = torch.linspace(-1.0, 1.0, nfaces * nf, dtype=torch.float64, device="cuda")
    M2, b2 = G.condense_batched(ctx, faces, loc, nc, nf, nfaces, B_T=bt, B_F=bf)
    assert np.abs(b2.cpu().numpy() - bf.cpu().numpy() - d["b"]).max() <= 1e-12 * np.abs(d["b"]).max() + 1e-15
    # cell recovery from the face solution
    xf = spla.spsolve(A.tocsc(), d["b"])
    xt = G.cell_recover_batched(ctx, faces, loc, nc, nf, nfaces, torch.from_numpy(xf).cuda(), B_T=bt).cpu().numpy()
    worst = 0.0
    for e in range(faces.shape[0]):
        xe = np.concatenate([xf[f * nf:(f + 1) * nf] if f >= 0 else np.zeros(nf) for f in faces[e]])
        res = local[e, :nc, :nc] @ xt[e] + local[e, :nc, nc:] @ xe - BT[e]
        worst = max(worst, np.abs(res).max() / (np.abs(BT[e]).max() + np.abs(local[e]).max() * np.abs(xt[e]).max()))
    assert worst <= 1e-13, worst


def test_condense_rejects_out_of_range_faces(ctx):
    d = load("elem_quad4_k2")
    faces, local, BT, nc, nf, nfaces = elements(d)
    loc = torch.from_numpy(np.ascontiguousarray(local)).cuda()
    bad = faces.copy()
    bad[0, 0] = nfaces
    with pytest.raises(G.FhError, match="out of range"):
        G.condense_batched(ctx, bad, loc, nc, nf, nfaces)
    with pytest.raises(G.FhError, match="local"):
        G.condense_batched(ctx, faces, loc[:-1], nc, nf, nfaces)
