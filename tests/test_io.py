"""Host input and ordering beside the solver path (no GPU needed):
MatrixMarket files read by the library against the reference's own reader
(HOST::ReadMtxTools, golden CSR arrays in tests/golden/mtx_*.npz written by
tests/golden/make_golden_io.py), and the library's reverse Cuthill-McKee
ordering against the reference's genrcm (tests/golden/rcm_*.npz), both bit
for bit; plus properties the golden files do not cover."""
import os

import numpy as np
import pytest

from oracle import ppo

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MTX = os.path.join(GOLD, "mtx")


@pytest.fixture(scope="module")
def pp():
    import prispack_amd
    return prispack_amd


@pytest.mark.parametrize("name", ["general", "symmetric", "skew", "zeros_dups", "rect"])
def test_read_mtx_matches_the_reference_reader(pp, name):
    g = np.load(os.path.join(GOLD, f"mtx_{name}.npz"))
    rp, ci, v = pp.read_mtx(os.path.join(MTX, name + ".mtx"))
    np.testing.assert_array_equal(rp, g["rp"])
    np.testing.assert_array_equal(ci, g["ci"])
    np.testing.assert_array_equal(v, g["v"])


def test_read_mtx_round_trip_and_errors(pp, tmp_path):
    # a generated operator written out and read back is the same CSR
    A = ppo.upwind(5, 4, 3, 20.0, 10.0, 5.0)
    n = len(A[0]) - 1
    rows = np.repeat(np.arange(n), np.diff(A[0]))
    p = tmp_path / "upwind.mtx"
    with open(p, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{n} {n} {len(A[1])}\n")
        for k in np.random.default_rng(0).permutation(len(A[1])):  # any entry order
            f.write(f"{rows[k] + 1} {A[1][k] + 1} {float(A[2][k])!r}\n")
    rp, ci, v = pp.read_mtx(str(p))
    np.testing.assert_array_equal(rp, A[0])
    np.testing.assert_array_equal(ci, A[1])
    np.testing.assert_array_equal(v, A[2])
    # a symmetric file of the lower triangle gives the full Laplacian
    L = ppo.laplacian(4, 3, 2)
    m = len(L[0]) - 1
    lr = np.repeat(np.arange(m), np.diff(L[0]))
    low = [(i, j, x) for i, j, x in zip(lr, L[1], L[2]) if j <= i]
    q = tmp_path / "lap.mtx"
    with open(q, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real symmetric\n{m} {m} {len(low)}\n")
        f.writelines(f"{i + 1} {j + 1} {float(x)!r}\n" for i, j, x in low)
    for got, want in zip(pp.read_mtx(str(q)), L):
        np.testing.assert_array_equal(got, want)
    # the dense right-hand side, column-major
    vals, shape = pp.read_mtx_dense(os.path.join(MTX, "rhs.mtx"))
    assert shape == (4, 2)
    np.testing.assert_array_equal(vals, [1.0, -2.0, 3.5, 0.25, -1.0, 0.0, 2.0, 1e-5])
    # refusals: complex data, a dense file as a matrix, an index beyond the size line, a missing file
    bad = {"complex.mtx": "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1.0 2.0\n",
           "range.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",
           "nobanner.mtx": "2 2 1\n1 1 1.0\n"}
    for name, text in bad.items():
        (tmp_path / name).write_text(text)
        with pytest.raises(pp.PrisPackError):
            pp.read_mtx(str(tmp_path / name))
    with pytest.raises(pp.PrisPackError):
        pp.read_mtx(os.path.join(MTX, "rhs.mtx"))
    with pytest.raises(pp.PrisPackError):
        pp.read_mtx(str(tmp_path / "missing.mtx"))


@pytest.mark.parametrize("name", ["lap2d_9x7", "lap3d_5x4x3", "upwind_6x5x4", "random_3comp"])
def test_rcm_matches_the_reference(pp, name):
    g = np.load(os.path.join(GOLD, f"rcm_{name}.npz"))
    perm = pp.rcm((g["rp"], g["ci"], np.ones(len(g["ci"]))))
    np.testing.assert_array_equal(perm, g["perm"])


def bandwidth(A, perm=None):
    rp, ci = np.asarray(A[0], np.int64), np.asarray(A[1], np.int64)
    n = len(rp) - 1
    inv = np.arange(n) if perm is None else np.argsort(perm)
    rows = np.repeat(np.arange(n), np.diff(rp))
    return int(np.max(np.abs(inv[rows] - inv[ci])))


def test_rcm_properties(pp):
    # a permutation; on a grid numbered against its short side it narrows the band
    # to the short side; every component keeps its nodes together
    A = ppo.laplacian(3, 40, 1)
    n = len(A[0]) - 1
    rng = np.random.default_rng(3)
    shuffle = rng.permutation(n).astype(np.uint32)
    from mc_order import permute_fast
    As = permute_fast(A, shuffle)
    perm = pp.rcm(As)
    assert sorted(perm.tolist()) == list(range(n))
    assert bandwidth(As) > 20 and bandwidth(As, perm) <= 3
    # two disjoint paths: positions of the first component come first
    rows = [[1], [0, 2], [1, 3], [2], [5], [4, 6], [5]]
    rp = np.zeros(8, np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([j for r in rows for j in r], np.uint32)
    perm = pp.rcm((rp, ci, np.ones(len(ci))))
    assert set(perm[:4].tolist()) == {0, 1, 2, 3} and set(perm[4:].tolist()) == {4, 5, 6}
    assert pp.rcm((np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0))).size == 0
