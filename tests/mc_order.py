"""An independent greedy distance-1 colouring and symmetric permutation, used
by the tests to build the oracle's view of a multicolour-ordered system."""
import numpy as np


def greedy_order(A):
    rp, ci, _ = A
    n = len(rp) - 1
    nb = [set() for _ in range(n)]
    for i in range(n):
        for j in ci[rp[i]:rp[i + 1]]:
            if j != i:
                nb[i].add(int(j))
                nb[int(j)].add(i)
    color = [-1] * n
    for i in range(n):
        used = {color[j] for j in nb[i] if color[j] >= 0}
        c = 0
        while c in used:
            c += 1
        color[i] = c
    return np.array(sorted(range(n), key=lambda i: (color[i], i)), np.uint32), max(color) + 1


def permute(A, order):
    rp, ci, v = A
    n = len(rp) - 1
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    rows, cols, vals = [], [], []
    for k, i in enumerate(order):
        c = inv[ci[rp[i]:rp[i + 1]]]
        s = np.argsort(c)
        rows.append(np.full(len(c), k))
        cols.append(c[s])
        vals.append(v[rp[i]:rp[i + 1]][s])
    rp2 = np.zeros(n + 1, np.uint32)
    rp2[1:] = np.cumsum([len(c) for c in cols])
    return rp2, np.concatenate(cols).astype(np.uint32), np.concatenate(vals)


def parity_order(nx, ny, nz=1):
    """Red-black order of an nx x ny x nz grid (x fastest): the nodes with
    (i + j + k) even first, then the odd ones, each in index order.  On 5- and
    7-point stencils this is the order greedy_order produces (the greedy
    colouring of a grid graph in index order gives node (i, j, k) the colour
    (i + j + k) mod 2); a vectorised form for grids too large for the loop."""
    idx = np.arange(nx * ny * nz)
    col = ((idx % nx) + (idx // nx) % ny + idx // (nx * ny)) & 1
    return np.argsort(col, kind="stable").astype(np.uint32)


def permute_fast(A, order):
    """permute() with numpy sorts: row k of the result is row order[k] of A,
    columns renumbered and sorted."""
    rp, ci, v = A
    n = len(rp) - 1
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(np.asarray(rp, np.int64)))
    newr, newc = inv[rows], inv[np.asarray(ci, np.int64)]
    s = np.lexsort((newc, newr))
    rp2 = np.zeros(n + 1, np.uint32)
    rp2[1:] = np.cumsum(np.bincount(newr, minlength=n))
    return rp2, newc[s].astype(np.uint32), np.asarray(v, np.float64)[s]
