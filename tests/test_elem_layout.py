"""The element layout's storage convention and product scheme, on the CPU.

The GPU assembly stores every cell's condensed matrix by cyclic diagonals
(diagonal t = 0 .. m/2 holds S[l][(l + t) mod m] at position l; diagonal 0
first, then diagonals 2p + 1 and 2p + 2 interleaved), and the product kernel
reads diagonal t at lane l, adds S[l][l+t] x[l+t] and sends S[l][l+t] x[l] to
lane l + t for 0 < t < m/2, over two cell colours.  This restates that scheme
with numpy on a small hexahedral grid with random symmetric cell matrices and
checks it against the assembled operator sum_c P_c^T S_c P_c (a fact fixed by
the assembly, independent of the scheme), for the face block sizes of k = 0..3.
"""
import numpy as np
import pytest


def elem_stride(m):
    return m * (1 + 2 * ((m // 2 + 1) // 2))


def elem_pos(m, t, l):
    return l if t == 0 else m + 2 * m * ((t - 1) // 2) + 2 * l + (t - 1) % 2


def grid(n):
    """Cells of an n^3 grid with their six faces (x-, x+, y-, y+, z-, z+) as
    block rows of the interior faces (-1 on the boundary), and the parity
    colour of every cell."""
    def fid(d, a, b, c):  # face with coordinate a along direction d
        return d * (n + 1) * n * n + a + (n + 1) * (b + n * c)
    unk, nb = {}, 0
    for d in range(3):
        for c in range(n):
            for b in range(n):
                for a in range(1, n):
                    unk[fid(d, a, b, c)] = nb
                    nb += 1
    cells, colour = [], []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                f = [fid(0, i, j, k), fid(0, i + 1, j, k), fid(1, j, i, k), fid(1, j + 1, i, k),
                     fid(2, k, i, j), fid(2, k + 1, i, j)]
                cells.append([unk.get(g, -1) for g in f])
                colour.append((i + j + k) % 2)
    return np.array(cells), np.array(colour), nb


@pytest.mark.parametrize("nf", [1, 3, 6, 10])
def test_cyclic_diagonal_scheme_reproduces_the_assembled_product(nf):
    rng = np.random.default_rng(nf)
    cells, colour, nb = grid(3)
    m = 6 * nf
    N = nb * nf
    S = []
    A = np.zeros((N, N))
    for c in cells:
        R = rng.standard_normal((m, m))
        Sc = R + R.T
        S.append(Sc)
        idx = [r * nf + i if r >= 0 else -1 for r in c for i in range(nf)]
        for a, ia in enumerate(idx):
            for b2, ib in enumerate(idx):
                if ia >= 0 and ib >= 0:
                    A[ia, ib] += Sc[a, b2]
    # packing as the assembly does: entry (r, s) with t = (s - r) mod m <= m/2
    st = elem_stride(m)
    packed = np.zeros((len(cells), st))
    written = np.zeros((len(cells), st), dtype=int)
    for ci, Sc in enumerate(S):
        for r in range(m):
            for s in range(m):
                t = (s - r) % m
                if t <= m // 2:
                    packed[ci, elem_pos(m, t, r)] = Sc[r, s]
                    written[ci, elem_pos(m, t, r)] += 1
    # every stored slot is written once (one unused slot per pair when m/2 is odd)
    assert written.max() == 1
    assert written.sum() == len(cells) * m * (m // 2 + 1)
    x = rng.standard_normal(N)
    y = np.full(N, np.nan)
    lanes = np.arange(m)
    for q in (0, 1):  # colour passes: the first writes, the second adds
        for ci in np.nonzero(colour == q)[0]:
            rows = np.repeat(cells[ci], nf)
            ent = np.tile(np.arange(nf), 6)
            e = rows * nf + ent
            xl = np.where(rows >= 0, x[np.maximum(e, 0)], 0.0)
            acc = packed[ci, [elem_pos(m, 0, l) for l in lanes]] * xl
            for t in range(1, m // 2 + 1):
                u = packed[ci, [elem_pos(m, t, l) for l in lanes]]
                acc += u * xl[(lanes + t) % m]          # S[l][l+t] x[l+t]
                if t < m // 2:
                    acc += (u * xl)[(lanes - t) % m]    # S[l-t][l] x[l-t], sent from lane l - t
            ok = rows >= 0
            if q == 0:
                y[e[ok]] = acc[ok]
            else:
                y[e[ok]] += acc[ok]
    assert not np.isnan(y).any()  # every unknown face has a cell of each colour
    np.testing.assert_allclose(y, A @ x, rtol=1e-12, atol=1e-12 * np.abs(A).sum(1).max() * np.abs(x).max())
