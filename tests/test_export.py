"""Exported linear-system format (Matrix Market text) against the reference's
own export: A.dat, b.dat and solutionFaces.dat written by ExportModule
(src/Utils/ExportModule.h:68-87) through Eigen's saveMarket, kept byte for
byte in the mm_* fixtures."""
import os

import numpy as np
import pytest

from conftest import csr, load
from fhhos4_amd import fhgpu as G

CASES = ["mm_c1_n4", "mm_tet2_k1"]


@pytest.fixture(scope="module")
def hctx():
    return G.Context(-1)


def _bytes(d, key):
    return d[key].tobytes()


@pytest.mark.parametrize("name", CASES)
def test_matrix_export_is_byte_identical(hctx, tmp_path, name):
    d = load(name)
    A = csr(d, "A")
    nf = int(d["info_nf"])
    M = G.Matrix.from_csr(hctx, A, nf, keep_host=True)
    out = tmp_path / "A.dat"
    M.write_market(str(out))
    assert out.read_bytes() == _bytes(d, "dat_A")


@pytest.mark.parametrize("name", CASES)
def test_matrix_import_round_trip(hctx, tmp_path, name):
    d = load(name)
    src = tmp_path / "ref_A.dat"
    src.write_bytes(_bytes(d, "dat_A"))
    nf = int(d["info_nf"])
    M = G.Matrix.read_market(hctx, str(src), nf)
    A = csr(d, "A")
    B = M.scipy()
    assert B.shape == A.shape
    assert abs(B - A).max() == 0.0
    again = tmp_path / "A.dat"
    M.write_market(str(again))
    assert again.read_bytes() == src.read_bytes()


@pytest.mark.parametrize("name", CASES)
def test_vector_export_and_import(tmp_path, name):
    d = load(name)
    for key, vec in (("dat_b", d["b"]), ("dat_solutionFaces", d["x"])):
        out = tmp_path / (key + ".dat")
        G.write_market_vector(vec, str(out))
        assert out.read_bytes() == _bytes(d, key)
        back = G.read_market_vector(str(out))
        assert np.array_equal(back, vec)


def test_import_sums_duplicates_and_sorts(hctx, tmp_path):
    p = tmp_path / "dup.dat"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% comment\n2 2 4\n2 2 1.5\n1 2 2\n2 2 0.5\n1 1 3\n")
    M = G.Matrix.read_market(hctx, str(p), 1)
    assert np.array_equal(M.scipy().toarray(), np.array([[3.0, 2.0], [0.0, 2.0]]))


@pytest.mark.parametrize("text,err", [
    ("%%MatrixMarket matrix array real general\n2 1\n1\n2\n", "coordinate real general"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n2 2 1\n", "entry count"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", "out of range"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1 1\n", "multiples of the block size"),
])
def test_import_rejects_malformed_files(hctx, tmp_path, text, err):
    p = tmp_path / "bad.dat"
    p.write_text(text)
    with pytest.raises(G.FhError, match=err):
        G.Matrix.read_market(hctx, str(p), 2 if "3 3" in text else 1)


def test_missing_file_is_an_error(hctx, tmp_path):
    with pytest.raises(G.FhError, match="cannot open"):
        G.Matrix.read_market(hctx, str(tmp_path / "none.dat"), 1)
    with pytest.raises(G.FhError, match="cannot open"):
        G.read_market_vector(str(tmp_path / "none.dat"))
