"""Writes the MatrixMarket and reverse Cuthill-McKee golden vectors under
tests/golden/ from the reference's own code (oracle/_ref/libppref.so):
HOST::ReadMtxTools<double>::loadMatrix (source/utils/TestTools/ReadMtxTools.cpp)
on the small files of tests/golden/mtx/, which this script also writes, and
the reference's genrcm (external_packages/RCM/lib/rcm.cpp) called as
HOST::rcmReorderCSR calls it (include/utils/ExternalTools/MatrixReorderTools.h:59-80).

Run in the container that holds /root/reference:
    make -C oracle ref && python tests/golden/make_golden_io.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MTX = os.path.join(HERE, "mtx")
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import ppo  # noqa: E402

# Small MatrixMarket files covering the reader's rules: storage schemes,
# comments and blank lines, entries out of row order, duplicates, explicit
# near-zero entries (dropped), upper-case banner words, a rectangular matrix
# and a dense right-hand side.
FILES = {
    "general.mtx": """%%MatrixMarket matrix coordinate real general
% a 4 x 4 general matrix, entries in no particular order
4 4 9
3 1 -2.5
1 1 4.0
2 3 1e-3
4 4 7.25
1 4 0.5

2 2 3.0
3 3 5.0
4 2 -1.0
1 2 2.0
""",
    "symmetric.mtx": """%%MatrixMarket matrix coordinate real symmetric
%-------------------------------------------
% lower triangle of a 5 x 5 SPD matrix
%-------------------------------------------
5 5 9
1 1 4.0
2 1 -1.0
2 2 4.0
3 2 -1.0
3 3 4.0
4 3 -1.0
4 4 4.0
5 1 -0.5
5 5 4.0
""",
    "skew.mtx": """%%MatrixMarket MATRIX Coordinate REAL Skew-Symmetric
3 3 3
2 1 1.5
3 1 -2.0
3 2 0.25
""",
    "zeros_dups.mtx": """%%MatrixMarket matrix coordinate real general
3 3 8
1 1 1.0
1 3 0.0
2 2 2.0
2 1 1e-300
3 3 3.0
1 3 4.0
1 3 -1.0
3 1 2.220446049250313e-15
""",
    "rect.mtx": """%%MatrixMarket matrix coordinate real general
2 5 4
2 5 1.0
1 4 -2.0
2 1 3.5
1 2 0.125
""",
    "rhs.mtx": """%%MatrixMarket matrix array real general
% column-major values of a 4 x 2 block of right-hand sides
4 2
1.0
-2.0
3.5
0.25
-1.0
0.0
2.0
1e-5
""",
}


def save(name, **arrays):
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print("wrote", name)


def rcm_cases():
    """Patterns with several components, missing diagonals and ties in degree."""
    cases = {
        "lap2d_9x7": ppo.laplacian(9, 7, 1),
        "lap3d_5x4x3": ppo.laplacian(5, 4, 3),
        "upwind_6x5x4": ppo.upwind(6, 5, 4, 20.0, 10.0, 5.0),
    }
    # a random structurally symmetric pattern of 3 components, some rows
    # without their diagonal and one isolated node
    rng = np.random.default_rng(17)
    n = 60
    edges = set()
    for lo, hi in ((0, 25), (25, 44), (45, 60)):
        for _ in range(3 * (hi - lo)):
            i, j = rng.integers(lo, hi, 2)
            if i != j:
                edges.add((int(i), int(j)))
                edges.add((int(j), int(i)))
    for i in range(n):
        if i % 7 != 3:
            edges.add((i, i))
    rows = [sorted(j for (a, j) in edges if a == i) for i in range(n)]
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([j for r in rows for j in r], np.uint32)
    cases["random_3comp"] = (rp, ci, np.ones(len(ci)))
    return cases


def main():
    R = ppo.Ref("strict")
    os.makedirs(MTX, exist_ok=True)
    for name, text in FILES.items():
        with open(os.path.join(MTX, name), "w") as f:
            f.write(text)
    cwd = os.getcwd()
    try:
        for name in FILES:
            if name == "rhs.mtx":
                continue
            path = os.path.join(MTX, name)
            h = R.L.refh_read_mtx(path.encode(), path.encode(), 1)
            rp, ci, v = R._csr_out(h)
            R.L.refh_csr_free(h)
            save("mtx_" + name[:-4], rp=rp, ci=ci, v=v)
    finally:
        os.chdir(cwd)  # the reader changes the working directory
    for name, A in rcm_cases().items():
        h = R.matrix(A)
        perm = np.zeros(len(A[0]) - 1, np.uint32)
        R.L.refh_rcm(h, perm)
        R.free_matrix(h)
        save("rcm_" + name, rp=np.asarray(A[0], np.uint32), ci=np.asarray(A[1], np.uint32), perm=perm)


if __name__ == "__main__":
    main()
