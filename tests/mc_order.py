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
