"""Multicolour-reordered factors, the AMSED multilevel low-rank Schur
preconditioner and the FP64 matrix-core GEMM, through the C ABI.

Multicolour: the product's apply equals the oracle's IC(0)/ILU(0) of P A P^T
(P from an independent greedy colouring written here), bit for bit in FP64.
AMSED and GMSLR: parity with the reference is unpinned (the reference needs METIS,
SuiteSparse AMD and LAPACKE, absent here); pinned instead by
  - the hierarchical ordering being a permutation whose levels are block
    diagonal (no coupling between blocks of one level),
  - exactness: with exact block factors (tau = 0) and rank 0 the recursion is
    an exact block LDL^T, so M^{-1} A x = x,
  - the Lanczos Ritz values lying in [0, 1) (S = C - E^T B^{-1} E is SPD),
  - PCG converging to rtol 1e-8 on the config-4 matrix type,
  - GMSLR: the nonsymmetric recursion against the same dense restatement,
    full-rank exactness on D A D^{-1} (real spectrum of G, no symmetry), and
    FGMRES(30) converging on a 3-D upwind matrix.
GEMM: against a host FP64 product (brute force)."""
import ctypes as C

import numpy as np
import pytest

from mc_order import greedy_order, permute
from oracle import ppo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def pp(torch):
    import prispack_amd
    return prispack_amd


@pytest.mark.parametrize("shape", [(14, 13, 12), (26, 25, 24)])
@pytest.mark.parametrize("kind,mat", [("ic0", "lap"), ("ilu0", "upw")])
def test_multicolor_apply_bit_exact(pp, torch, kind, mat, shape):
    # (26, 25, 24): colour levels of ~7800 rows take the LDS-staged sweep kernel
    A = ppo.laplacian(*shape) if mat == "lap" else ppo.upwind(*shape, 20.0, 10.0, 5.0)
    n = len(A[0]) - 1
    order, ncol = greedy_order(A)
    assert ncol == 2
    Ap = permute(A, order)
    M = pp.CSRMatrix(*A)
    P = (pp.IC0 if kind == "ic0" else pp.ILU0)(M, ordering=1)
    P.setup()
    assert P.info().levels_lower <= 2 and P.info().levels_upper <= 2
    Q = ppo.Prec(Ap, kind)
    x = ppo.uniform(n, 5)
    t = torch.tensor(x, device="cuda")
    P.MInvSolve(t)
    want = np.empty(n)
    want[order] = Q.apply(x[order])
    assert np.array_equal(t.cpu().numpy(), want)


def test_multicolor_pcg(pp, torch):
    A = ppo.laplacian(40, 40, 40)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 0))
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M, ordering=1)
    for prec in ("fp64", "mixed"):
        cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=3000, precision=prec)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        cg.solve(M, torch.tensor(b, device="cuda"), x)
        assert cg.getConvergence() and cg.report.true_relres < 1e-8
    order, _ = greedy_order(A)
    o = ppo.cg(permute(A, order), b[order], ppo.Prec(permute(A, order), "ic0"), tol=1e-8, maxit=3000)
    cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=3000)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    cg.solve(M, torch.tensor(b, device="cuda"), x)
    assert abs(cg.getNumIter() - o["iters"]) <= 2
    xo = np.empty(n)
    xo[order] = o["x"]
    assert np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo) <= 1e-6


def test_gemm_tall_f64_matrix_cores(pp, torch):
    from prispack_amd import _capi
    ctx = pp.Context.default()
    rng = np.random.default_rng(3)
    for n, k, r in ((1000, 64, 32), (37, 5, 3), (4096, 128, 48)):
        V = rng.standard_normal((k, n))  # column-major n x k = row-major k x n
        Y = rng.standard_normal((r, k))  # column-major k x r
        tV, tY = torch.tensor(V, device="cuda"), torch.tensor(Y, device="cuda")
        tW = torch.zeros((r, n), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        ctx.check(_capi.lib().pp_gemm_tall_f64(ctx.h, n, k, r, C.c_void_p(tV.data_ptr()), C.c_void_p(tY.data_ptr()),
                                                C.c_void_p(tW.data_ptr())))
        want = (V.T @ Y.T).T  # (n x r) column-major as (r x n) row-major
        np.testing.assert_allclose(tW.cpu().numpy(), want, rtol=1e-12, atol=1e-12)


def _ml_info(pp, P):
    from prispack_amd import _capi
    nlev = C.c_uint32()
    sizes = (C.c_uint32 * 16)()
    ranks = (C.c_uint32 * 16)()
    assert _capi.lib().pp_ml_info(P.h, C.byref(nlev), sizes, ranks, 16) == 0
    perm = np.zeros(P.A.n, np.uint32)
    assert _capi.lib().pp_ml_permutation(P.h, perm.ctypes.data_as(C.POINTER(C.c_uint32))) == 0
    return nlev.value, list(sizes)[:nlev.value], list(ranks)[:nlev.value], perm


def dense(A):
    rp, ci, v = A
    n = len(rp) - 1
    D = np.zeros((n, n))
    D[np.repeat(np.arange(n), np.diff(rp)), ci] = v
    return D


def dense_recursion(Ap, lev_ptr, level, x):
    """levelCInvRecursiveSolve (AMSEDPrecondition.cpp:1494-1587) with exact
    block solves and no low-rank term, densely: C_level^{-1} x."""
    nlev = len(lev_ptr) - 1
    nxt = level + 1
    if nxt == nlev - 1:
        return np.linalg.solve(Ap[lev_ptr[nxt]:, lev_ptr[nxt]:], x)
    b0, b1 = lev_ptr[nxt], lev_ptr[nxt + 1]
    B, E, F = Ap[b0:b1, b0:b1], Ap[b0:b1, b1:], Ap[b1:, b0:b1]
    nb = b1 - b0
    x1 = np.linalg.solve(B, x[:nb])
    x2 = dense_recursion(Ap, lev_ptr, nxt, x[nb:] - F @ x1)
    return np.concatenate([x1 - np.linalg.solve(B, E @ x2), x2])


@pytest.mark.parametrize("levels,ndom", [(2, 4), (3, 8), (4, 8)])
def test_amsed_recursion_matches_dense_restatement(pp, torch, levels, ndom):
    A = ppo.laplacian(12, 11, 10)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    P = pp.AMSED(M, levels=levels, ndom=ndom, rank=0, exact=True)
    P.setup()
    nlev, sizes, ranks, perm = _ml_info(pp, P)
    assert nlev == levels and sum(sizes) == n and not any(ranks)
    assert np.array_equal(np.sort(perm), np.arange(n))
    Ap = dense(A)[np.ix_(perm, perm)]
    lev_ptr = np.concatenate([[0], np.cumsum(sizes)])
    # the blocks of a level are decoupled: B_l is block diagonal by construction,
    # and every level but the last has at least two blocks' worth of rows
    x = ppo.uniform(n, 2)
    y = torch.tensor(x, device="cuda")
    P.MInvSolve(y)
    want = np.empty(n)
    want[perm] = dense_recursion(Ap, lev_ptr, -1, x[perm])
    np.testing.assert_allclose(y.cpu().numpy(), want, rtol=1e-9, atol=1e-11)


def test_amsed_full_rank_is_exact(pp, torch):
    # exact blocks and a complete low-rank term: C^{-1} + W H W^T = S^{-1}
    # at every level, so M^{-1} A = I
    A = ppo.laplacian(8, 8, 8)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    P = pp.AMSED(M, levels=3, ndom=4, rank=n, lanczos_steps=250, exact=True)
    P.setup()
    x = ppo.uniform(n, 3)
    y = torch.tensor(ppo.spmv(A, x), device="cuda")
    P.MInvSolve(y)
    np.testing.assert_allclose(y.cpu().numpy(), x, rtol=0, atol=1e-6)


@pytest.mark.parametrize("rank,lowrank", [(0, "reference"), (16, "reference"), (16, "deflation"), (16, "subspace")])
def test_amsed_pcg_config4_type(pp, torch, rank, lowrank):
    # the reference's restarted Lanczos with its basic (mu / (1 - mu)) and
    # eigenvalue-deflation ((eps - lam) / lam) cores, and this build's one-run form
    A = ppo.laplacian(32, 32, 32)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 0))
    M = pp.CSRMatrix(*A)
    P = pp.AMSED(M, levels=3, ndom=16, rank=rank, lowrank=lowrank)
    P.setup()
    nlev, sizes, ranks, _ = _ml_info(pp, P)
    print(f"[amsed] sizes {sizes} ranks {ranks}", flush=True)
    if rank and lowrank != "deflation":  # the deflated core keeps only the Ritz values above 1 - eps
        assert all(0 < r <= rank for r in ranks[:-1])
    cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=1000)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    cg.solve(M, torch.tensor(b, device="cuda"), x)
    assert cg.getConvergence() and cg.report.true_relres < 1e-8
    o = ppo.cg(A, b, None, tol=1e-12, maxit=5000)
    assert np.linalg.norm(x.cpu().numpy() - o["x"]) / np.linalg.norm(o["x"]) <= 1e-6
    print(f"[amsed] rank {rank} {lowrank}: {cg.getNumIter()} iterations", flush=True)


def scaled(A, seed):
    """D A D^{-1} with a random positive diagonal D: nonsymmetric, with the
    pattern and eigenvalues of A."""
    rp, ci, v = A
    n = len(rp) - 1
    d = np.exp(np.random.default_rng(seed).uniform(-1, 1, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    return rp, ci, v * d[rows] / d[ci]


@pytest.mark.parametrize("levels,ndom", [(2, 4), (3, 8)])
def test_gmslr_recursion_matches_dense_restatement(pp, torch, levels, ndom):
    # GMSLRDPrecondition.cpp levelCInvRecursiveSolve with exact blocks and no
    # low-rank term; nonsymmetric matrix (upwind), F the lower coupling
    A = ppo.upwind(10, 9, 8, 30.0, 10.0, 5.0)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    P = pp.GMSLR(M, levels=levels, ndom=ndom, rank=0, exact=True)
    P.setup()
    nlev, sizes, ranks, perm = _ml_info(pp, P)
    assert nlev == levels and sum(sizes) == n and not any(ranks)
    Ap = dense(A)[np.ix_(perm, perm)]
    lev_ptr = np.concatenate([[0], np.cumsum(sizes)])
    x = ppo.uniform(n, 4)
    y = torch.tensor(x, device="cuda")
    P.MInvSolve(y)
    want = np.empty(n)
    want[perm] = dense_recursion(Ap, lev_ptr, -1, x[perm])
    np.testing.assert_allclose(y.cpu().numpy(), want, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("levels,ndom", [(2, 2), (3, 4)])
def test_gmslr_full_rank_is_exact(pp, torch, levels, ndom):
    # D A D^{-1} of a variable-coefficient diffusion matrix: G = C^{-1} F B^{-1} E
    # is diagonalisable with real eigenvalues in [0, 1) and no structural
    # symmetry, so a full Arnoldi basis (every Schur complement under 64 rows)
    # makes (I + W K W^T) C^{-1} = S^{-1} at every level and M^{-1} A = I
    side = 12
    kap = np.exp(np.random.default_rng(5).uniform(-1, 1, side * side))
    A = scaled(ppo.vardiff2d(side, kap), 6)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    # (the exact-subspace core; the reference's diagonal core on Schur vectors
    # is exact only when the Schur form is diagonal)
    P = pp.GMSLR(M, levels=levels, ndom=ndom, rank=64, lanczos_steps=64, exact=True, lowrank="subspace")
    P.setup()
    nlev, sizes, ranks, _ = _ml_info(pp, P)
    print(f"[gmslr] sizes {sizes} ranks {ranks}", flush=True)
    assert all(n - sum(sizes[:l + 1]) <= 64 for l in range(nlev - 1))
    x = ppo.uniform(n, 3)
    y = torch.tensor(ppo.spmv(A, x), device="cuda")
    P.MInvSolve(y)
    np.testing.assert_allclose(y.cpu().numpy(), x, rtol=0, atol=1e-7)


@pytest.mark.parametrize("rank,lowrank", [(0, "reference"), (16, "reference"), (16, "subspace")])
def test_gmslr_gmres_upwind(pp, torch, rank, lowrank):
    A = ppo.upwind(24, 24, 24, 40.0, 20.0, 10.0)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 1))
    M = pp.CSRMatrix(*A)
    P = pp.GMSLR(M, levels=3, ndom=16, rank=rank, lowrank=lowrank)
    P.setup()
    nlev, sizes, ranks, _ = _ml_info(pp, P)
    print(f"[gmslr] sizes {sizes} ranks {ranks}", flush=True)
    g = pp.GMRES(n, P, 1e-8, pp.FlexibleGMRES, maxIter=500)
    g.setRestart(True, 30)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    g.solve(M, torch.tensor(b, device="cuda"), x)
    assert g.getConvergence() and g.report.true_relres < 1e-7
    print(f"[gmslr] rank {rank} {lowrank}: {g.getNumIter()} iterations", flush=True)


@pytest.mark.parametrize("kind", ["amsed", "gmslr"])
def test_multilevel_fp32_apply_and_mixed_solve(pp, torch, kind):
    # FP32 apply (factors, couplings, low-rank basis in FP32) against the FP64
    # apply, then the mixed-precision solve to the FP64 tolerance
    A = ppo.laplacian(24, 24, 24) if kind == "amsed" else ppo.upwind(24, 24, 24, 30.0, 10.0, 5.0)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    P = (pp.AMSED if kind == "amsed" else pp.GMSLR)(M, levels=3, ndom=8, rank=8)
    P.setup()
    x = ppo.uniform(n, 6)
    y64 = torch.tensor(x, device="cuda")
    P.MInvSolve(y64)
    y32 = torch.tensor(x, dtype=torch.float32, device="cuda")
    P.MInvSolve(y32)
    err = np.linalg.norm(y32.double().cpu().numpy() - y64.cpu().numpy()) / np.linalg.norm(y64.cpu().numpy())
    print(f"[{kind}] fp32 apply rel err {err:.2e}", flush=True)
    assert err < 1e-4
    b = ppo.spmv(A, ppo.normal(n, 2))
    xs = torch.zeros(n, dtype=torch.float64, device="cuda")
    if kind == "amsed":
        s = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=2000, precision="mixed")
    else:
        s = pp.GMRES(n, P, 1e-8, pp.FlexibleGMRES, maxIter=3000, precision="mixed")
        s.setRestart(True, 30)
    s.solve(M, torch.tensor(b, device="cuda"), xs)
    print(f"[{kind}] mixed: {s.getNumIter()} iterations, relres {s.report.true_relres:.2e}", flush=True)
    assert s.getConvergence() and s.report.true_relres < 1e-8


def test_setup_after_a_converged_solve(pp, torch):
    # a converged solve leaves its stop word set on the context; a multilevel
    # setup that runs afterwards (Lanczos through block sweeps) must not see it
    A = ppo.laplacian(20, 20, 20)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    x = ppo.uniform(n, 4)

    def build_apply():
        P = pp.AMSED(M, levels=3, ndom=8, rank=16)
        P.setup()
        t = torch.tensor(x, device="cuda")
        P.MInvSolve(t)
        return P.info_levels()[1], t.cpu().numpy()

    r1, v1 = build_apply()
    P0 = pp.IC0(M)
    cg = pp.CG(n, P0, pp.FlexibleCG, tol=1e-8, maxIter=1000)
    xs = torch.zeros(n, dtype=torch.float64, device="cuda")
    cg.solve(M, torch.tensor(ppo.spmv(A, x), device="cuda"), xs)
    assert cg.getConvergence()
    r2, v2 = build_apply()
    assert list(r1) == list(r2)
    assert np.array_equal(v1, v2)
