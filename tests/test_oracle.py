"""Pins the oracle (oracle/ppo.c) before anything is checked against it.

1. Golden vectors written by the reference's own compiled C++ code
   (tests/golden/make_golden.py): bit-exact equality, because the oracle
   restates the reference's serial code paths in the same operation order.
2. Mathematics for what the reference does not cover: closed-form Laplacian
   eigenpairs, the adjoint identity, complete Cholesky at tau = 0, the
   defining pattern invariants of IC(0)/ILU(0), one-step convergence with an
   exact preconditioner.
CPU only.
"""
import glob
import os

import numpy as np
import pytest

from oracle import ppo

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz")))


def dense(A):
    rp, ci, v = A
    n = len(rp) - 1
    M = np.zeros((n, n))
    for i in range(n):
        M[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    return M


def dims_of(g):
    return tuple(int(t) for t in g["dims"])


# --------------------------------------------------------------------------- generators / SpMV
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "spmv_*.npz"))))
def test_generator_and_spmv_golden(path):
    g = dict(np.load(path))
    A = ppo.laplacian(*dims_of(g), *[float(c) for c in g["conv"]])
    assert np.array_equal(A[0], g["rp"]) and np.array_equal(A[1], g["ci"]) and np.array_equal(A[2], g["v"])
    assert np.array_equal(ppo.spmv(A, g["x"]), g["y"])
    assert np.array_equal(ppo.spmv_t(A, g["x"]), g["yt"])


@pytest.mark.parametrize("dims", [(9, 7, 1), (5, 6, 4)])
def test_spmv_closed_form_eigenvector(dims):
    nx, ny, nz = dims
    A = ppo.laplacian(*dims)
    p, q, r = 2, 3, 1
    i = np.arange(nx)[None, None, :] + 1
    j = np.arange(ny)[None, :, None] + 1
    k = np.arange(nz)[:, None, None] + 1
    v = np.sin(i * p * np.pi / (nx + 1)) * np.sin(j * q * np.pi / (ny + 1))
    lam = 4 - 2 * np.cos(p * np.pi / (nx + 1)) - 2 * np.cos(q * np.pi / (ny + 1))
    if nz > 1:
        v = v * np.sin(k * r * np.pi / (nz + 1))
        lam += 2 - 2 * np.cos(r * np.pi / (nz + 1))
    v = np.broadcast_to(v, (nz, ny, nx)).ravel()
    np.testing.assert_allclose(ppo.spmv(A, v), lam * v, atol=1e-13)


def test_spmv_t_adjoint_identity():
    A = ppo.upwind(7, 6, 5, 20.0, 10.0, 5.0)
    n = len(A[0]) - 1
    x, y = ppo.uniform(n, 1), ppo.uniform(n, 2)
    assert abs(ppo.spmv_t(A, x) @ y - x @ ppo.spmv(A, y)) < 1e-12 * n
    At = ppo.transpose(A)
    np.testing.assert_array_equal(dense(At), dense(A).T)


def test_seeded_generators():
    u = ppo.uniform(100000, 0)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01
    z = ppo.normal(100001, 3)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    # splitmix64 first output for seed 0 (published test value 0xE220A8397B1DCDAF)
    first = (0xE220A8397B1DCDAF >> 11) / 2.0**53 * 2 - 1
    assert u[0] == first


def test_upwind_and_vardiff_structure():
    A = ppo.upwind(4, 4, 4, 20.0, 10.0, 5.0)
    D = dense(A)
    h = 1 / 5
    assert D[0, 0] == 6 + h * 35 and D[1, 0] == -1 - h * 20 and D[0, 1] == -1.0
    assert D[4, 0] == -1 - h * 10 and D[16, 0] == -1 - h * 5
    ones = ppo.vardiff2d(6, np.ones(36))
    np.testing.assert_array_equal(dense(ones), dense(ppo.laplacian(6, 6, 1)))
    kap = np.exp(ppo.normal(36, 5))
    V = dense(ppo.vardiff2d(6, kap))
    np.testing.assert_allclose(V, V.T, rtol=0, atol=0)
    assert np.all(np.linalg.eigvalsh(V) > 0)


def test_vardiff_face_coefficients_by_hand():
    # 3x3 grid, node p = 3*j + i holds kappa = p + 1.  Face coefficients worked
    # out by hand: an interior face between kappa a and b is the harmonic mean
    # 2ab/(a+b); a boundary face takes the node's own kappa.  The arithmetic
    # (a+b)/2 or geometric sqrt(ab) means give other numbers at every entry below.
    V = dense(ppo.vardiff2d(3, np.arange(1.0, 10.0)))
    # node 0 (kappa 1): east to node 1 (kappa 2) 4/3, north to node 3 (kappa 4) 8/5,
    # south and west are boundary faces of weight 1 each
    assert V[0, 1] == -4 / 3 and V[0, 3] == -8 / 5
    assert abs(V[0, 0] - 74 / 15) < 1e-15
    # centre node 4 (kappa 5): neighbours kappa 2, 4, 6, 8
    assert V[4, 1] == -20 / 7 and V[4, 3] == -40 / 9 and V[4, 5] == -60 / 11 and V[4, 7] == -80 / 13
    assert abs(V[4, 4] - (20 / 7 + 40 / 9 + 60 / 11 + 80 / 13)) < 1e-14
    # corner node 8 (kappa 9): west to node 7 (kappa 8) 144/17, south to node 5 (kappa 6) 108/15,
    # east and north boundary faces of weight 9 each
    assert V[8, 7] == -144 / 17 and V[8, 5] == -108 / 15
    assert abs(V[8, 8] - (144 / 17 + 108 / 15 + 18)) < 1e-14
    # no coupling beyond the 5-point stencil
    assert V[0, 4] == 0 and V[0, 2] == 0 and np.count_nonzero(V) == 9 + 2 * 12


# --------------------------------------------------------------------------- factorizations
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ict_*.npz"))))
def test_ict_golden(path):
    g = dict(np.load(path))
    A = ppo.laplacian(*dims_of(g))
    P = ppo.Prec(A, "ict", float(g["tau"]))
    lt = P.factors()["lt"]
    assert np.array_equal(lt[0], g["lt_rp"]) and np.array_equal(lt[1], g["lt_ci"])
    assert np.array_equal(lt[2], g["lt_v"])
    assert P.mnnz == int(g["mnnz"])
    for part in ("apply", "lower", "upper"):
        assert np.array_equal(getattr(P, part)(g["x"]), g[part]), part


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ildlt_*.npz"))))
def test_ildlt_golden(path):
    g = dict(np.load(path))
    A = ppo.laplacian(*dims_of(g))
    P = ppo.Prec(A, "ildlt", float(g["tau"]))
    assert P.mnnz == int(g["mnnz"])
    for part in ("apply", "lower", "diag", "upper"):
        assert np.array_equal(getattr(P, part)(g["x"]), g[part]), part


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ilut_*.npz"))))
def test_ilut_golden(path):
    g = dict(np.load(path))
    A = (g["rp"], g["ci"], g["v"])
    P = ppo.Prec(A, "ilut", float(g["tau"]), int(g["lfil"]))
    f = P.factors()
    for key in ("l", "u"):
        for k, part in enumerate(("rp", "ci", "v")):
            assert np.array_equal(f[key][k], g[f"{key}_{part}"]), (key, part)
    assert np.array_equal(f["d"], g["d"])
    assert P.mnnz == int(g["mnnz"])
    assert np.array_equal(P.apply(g["x"]), g["apply"])


def test_ict_tau0_is_complete_cholesky():
    A = ppo.laplacian(7, 6, 1)
    lt = ppo.Prec(A, "ict", 0.0).factors()["lt"]
    L = dense(lt).T
    np.testing.assert_allclose(L, np.linalg.cholesky(dense(A)), atol=1e-13)


def _pattern_products(A, L, U):
    M = L @ U
    rp, ci, v = A
    for i in range(len(rp) - 1):
        for k in range(rp[i], rp[i + 1]):
            yield M[i, ci[k]], v[k]


@pytest.mark.parametrize("A", [ppo.laplacian(6, 5, 4), ppo.laplacian(12, 11, 1),
                               ppo.vardiff2d(9, np.exp(ppo.normal(81, 4)))])
def test_ic0_pattern_invariant(A):
    P = ppo.Prec(A, "ic0")
    assert P.status == 0
    Lt = dense(P.factors()["lt"])
    L = Lt.T
    pattern = dense((A[0], A[1], np.ones_like(A[2]))) != 0
    assert np.all((np.tril(L) != 0) <= np.tril(pattern))   # no fill outside pattern
    for got, want in _pattern_products(A, L, Lt):
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want))
    # a mistake in the invariant check itself would pass any factor: a factor
    # with one perturbed entry must fail it
    bad = Lt.copy()
    r, c = np.nonzero(np.triu(bad, 1))
    bad[r[0], c[0]] *= 1.001
    assert any(abs(g - w) > 1e-8 for g, w in _pattern_products(A, bad.T, bad))


@pytest.mark.parametrize("A", [ppo.upwind(6, 5, 4, 20.0, 10.0, 5.0), ppo.laplacian(9, 8, 1, 0.3, 0.1, 0, 0)])
def test_ilu0_pattern_invariant(A):
    P = ppo.Prec(A, "ilu0")
    f = P.factors()
    L = np.eye(len(A[0]) - 1) + dense(f["l"])
    U = dense(f["u"]) + np.diag(1.0 / f["d"])
    pattern = dense((A[0], A[1], np.ones_like(A[2]))) != 0
    assert np.all((L != 0) <= (pattern | np.eye(len(pattern), dtype=bool)))
    assert np.all((U != 0) <= pattern)
    for got, want in _pattern_products(A, L, U):
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want))


def test_zero_fill_factors_agree_on_tridiagonal():
    # The exact factors of a tridiagonal matrix have no fill, so IC(0) must
    # equal the complete threshold factor and ILU(0) the complete ILUT.
    A = ppo.laplacian(15, 1, 1)
    a = ppo.Prec(A, "ic0").factors()["lt"]
    b = ppo.Prec(A, "ict", 0.0).factors()["lt"]
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=1e-15, atol=0)
    C = ppo.laplacian(15, 1, 1, 0.3, 0, 0, 0)
    a = ppo.Prec(C, "ilu0").factors()
    b = ppo.Prec(C, "ilut", 0.0, 15).factors()
    np.testing.assert_allclose(dense(a["l"]), dense(b["l"]), rtol=1e-15)
    np.testing.assert_allclose(dense(a["u"]), dense(b["u"]), rtol=1e-15)
    np.testing.assert_allclose(a["d"], b["d"], rtol=1e-15)


def test_factor_error_paths():
    A = ppo.laplacian(5, 5, 1, 0, 0, 0, 6.0)          # diagonal -2: indefinite
    assert ppo.Prec(A, "ict", 1e-3).status == 1
    assert ppo.Prec(A, "ic0").status == 1
    Z = (np.array([0, 1, 1, 2], np.uint32), np.array([0, 2], np.uint32), np.array([1.0, 1.0]))
    assert ppo.Prec(Z, "ilut", 0.0, 5).status == 2      # zero row 1


# --------------------------------------------------------------------------- solvers
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "cg_*.npz"))))
def test_cg_golden(path):
    g = dict(np.load(path))
    A = ppo.laplacian(*dims_of(g))
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.uniform(n, int(g["seed"])))
    P = ppo.Prec(A, "ict", float(g["tau"]))
    out = ppo.cg(A, b, P, tol=float(g["tol"]), maxit=int(g["maxit"]))
    assert out["iters"] == int(g["iters"]) and out["converged"] == bool(g["converged"])
    assert np.array_equal(out["x"], g["x"])
    assert out["error"] == float(g["error"])
    outn = ppo.cg(A, b, P, tol=float(g["tol"]), maxit=int(g["maxit"]), flexible=False)
    assert outn["iters"] == int(g["normal_iters"]) and outn["error"] == float(g["normal_error"])


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "gmres_*.npz"))))
def test_gmres_golden(path):
    g = dict(np.load(path))
    A = (g["rp"], g["ci"], g["v"])
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.uniform(n, int(g["seed"])))
    P = ppo.Prec(A, "ilut", float(g["tau"]), int(g["lfil"]))
    out = ppo.gmres(A, b, P, tol=float(g["tol"]), restart=int(g["restart"]))
    assert out["iters"] == int(g["iters"]) and out["converged"] == bool(g["converged"])
    assert np.array_equal(out["x"], g["x"])


def test_cg_with_exact_preconditioner_takes_one_step():
    A = ppo.laplacian(10, 9, 1)
    n = len(A[0]) - 1
    xt = ppo.uniform(n, 0)
    out = ppo.cg(A, ppo.spmv(A, xt), ppo.Prec(A, "ict", 0.0))
    assert out["iters"] == 2 and out["converged"]        # converged at i=1, reported m+1
    np.testing.assert_allclose(out["x"], xt, rtol=0, atol=1e-12)


def test_cg_distinct_eigenvalues():
    # unpreconditioned CG on a diagonal matrix with 3 distinct eigenvalues
    n = 30
    d = np.repeat([1.0, 2.0, 5.0], 10)
    A = (np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), d)
    out = ppo.cg(A, np.ones(n), None, tol=1e-12, maxit=50)
    assert out["converged"] and out["iters"] <= 4
    np.testing.assert_allclose(out["x"], 1 / d, rtol=1e-12)


def test_gmres_exact_preconditioner_quirk():
    # With M = A the first Arnoldi step already meets the tolerance; the
    # reference back-solves with dim = j = 0 (GMRES.cpp:194-201), so x is left
    # at x0 while convergence is reported.  The oracle reproduces that.
    A = ppo.upwind(5, 5, 5, 20.0, 10.0, 5.0)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.uniform(n, 0))
    out = ppo.gmres(A, b, ppo.Prec(A, "ilut", 0.0, n), tol=1e-8, restart=30)
    assert out["converged"] and out["iters"] == 0
    assert np.all(out["x"] == 0)
