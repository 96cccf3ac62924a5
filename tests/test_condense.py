"""Batched static condensation: the element-matrix fixtures (the reference's
own Diff_HHOElement::A, face numbers and cell source terms, dumped by the
reference driver) and the C-ABI's argument checks, without a GPU.

The condensation of the fixtures' element matrices, cell by cell with dense
numpy solves, reproduces the reference's condensed system A and b
(Diffusion_HHO.h:624-637, :1026-1032): this pins the input convention the
GPU entry point fh_condense_batched is checked with (cell unknowns first,
faces in element order, -1 for Dirichlet faces) on a fixture-independent
computation."""
import numpy as np
import pytest

from conftest import csr, load

F = pytest.importorskip("fhhos4_amd")
G = F.fhgpu

CASES = ["elem_cart4_k1", "elem_stetra2_k2", "elem_quad4_k2", "elem_cart3_k3"]


def elements(d):
    """(cell_faces, local, B_T, nc, nf, nfaces) of an element fixture."""
    nf = int(d["info_nf"])
    A = csr(d, "A")
    nfaces = A.shape[0] // nf
    BT = d["elem_BT"]
    faces = d["elem_faces"]
    # nl^2 * ncells = len(local), nc * ncells = len(B_T), nfc * ncells = len(faces)
    for nc in range(1, 64):
        if len(BT) % nc:
            continue
        ncells = len(BT) // nc
        if len(faces) % ncells:
            continue
        nfc = len(faces) // ncells
        nl = nc + nfc * nf
        if nl * nl * ncells == len(d["elem_A"]):
            return (faces.reshape(ncells, nfc), d["elem_A"].reshape(ncells, nl, nl), BT.reshape(ncells, nc), nc, nf,
                    nfaces)
    raise AssertionError("inconsistent element fixture")


def condense_numpy(faces, local, BT, nc, nf, nfaces):
    N = nfaces * nf
    S = np.zeros((N, N))
    b = np.zeros(N)
    for e in range(faces.shape[0]):
        Att, Atf, Aff = local[e, :nc, :nc], local[e, :nc, nc:], local[e, nc:, nc:]
        X = np.linalg.solve(Att, np.column_stack([Atf, BT[e]]))
        Sc = Aff - Atf.T @ X[:, :-1]
        c = Atf.T @ X[:, -1]
        for j1, f1 in enumerate(faces[e]):
            if f1 < 0:
                continue
            b[f1 * nf:(f1 + 1) * nf] -= c[j1 * nf:(j1 + 1) * nf]
            for j2, f2 in enumerate(faces[e]):
                if f2 >= 0:
                    S[f1 * nf:(f1 + 1) * nf, f2 * nf:(f2 + 1) * nf] += Sc[j1 * nf:(j1 + 1) * nf, j2 * nf:(j2 + 1) * nf]
    return S, b


@pytest.mark.parametrize("name", CASES)
def test_element_fixtures_condense_to_the_reference_system(name):
    d = load(name)
    faces, local, BT, nc, nf, nfaces = elements(d)
    # every local matrix is symmetric and its A_TT positive definite
    assert np.abs(local - local.transpose(0, 2, 1)).max() <= 1e-12 * np.abs(local).max()
    assert all(np.linalg.eigvalsh(L[:nc, :nc]).min() > 0 for L in local)
    S, b = condense_numpy(faces, local, BT, nc, nf, nfaces)
    A = csr(d, "A").toarray()
    assert np.abs(S - A).max() <= 1e-12 * np.abs(A).max()
    assert np.abs(b - d["b"]).max() <= 1e-12 * np.abs(d["b"]).max()
    # the reference's pattern (exact zeros pruned) lies inside the pattern
    # of face pairs of one cell
    assert np.all((A != 0) <= (np.abs(S) > 0) | (np.abs(A) <= 1e-14 * np.abs(A).max()))


def test_condense_needs_a_device_context():
    import ctypes
    h = G.Context(-1)
    faces = np.zeros(2, dtype=np.int32)
    out = ctypes.c_void_p()
    rc = G.lib().fh_condense_batched(h.h, 1, 2, 1, 1, 2, faces, None, None, None, None, ctypes.byref(out))
    assert rc == -9 and b"device context" in G.lib().fh_last_error()
    rc = G.lib().fh_cell_recover_batched(h.h, 1, 2, 1, 1, 2, faces, None, None, None, None)
    assert rc == -9
