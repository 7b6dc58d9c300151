"""Solves on matrices read from MatrixMarket files and with the reverse
Cuthill-McKee ordering, through the C ABI, against the oracle.

Bars: the matrix read back from a written .mtx file is the generator's CSR bit
for bit; the FP64 solves on it meet the solver bar against the oracle's solve
of the same system (iterations within +-2 of the oracle's FlexibleCG /
FGMRES(30) with the same factor in the same order, true relative residual
<= tol, x within 1e-6)."""
import numpy as np
import pytest

from mc_order import permute_fast
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


def write_mtx(path, A, symmetric=False):
    rp, ci, v = A
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci <= rows if symmetric else np.ones(len(ci), bool)
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real {'symmetric' if symmetric else 'general'}\n")
        f.write("% written by the test\n")
        f.write(f"{n} {n} {int(keep.sum())}\n")
        f.writelines(f"{i + 1} {j + 1} {float(x)!r}\n" for i, j, x in zip(rows[keep], ci[keep], v[keep]))


def test_solve_on_mtx_files_with_rcm(pp, torch, tmp_path):
    # SPD: a shuffled 3D Poisson matrix stored as its lower triangle
    A0 = ppo.laplacian(18, 17, 16)
    n = len(A0[0]) - 1
    shuffle = np.random.default_rng(2).permutation(n).astype(np.uint32)
    A = permute_fast(A0, shuffle)
    write_mtx(tmp_path / "spd.mtx", A, symmetric=True)
    R = pp.read_mtx(tmp_path / "spd.mtx")
    for got, want in zip(R, A):
        np.testing.assert_array_equal(got, want)
    b = ppo.spmv(A, ppo.uniform(n, 3))
    M = pp.CSRMatrix(*R)
    perm = pp.rcm(R)
    Ar = permute_fast(A, perm)
    for ordering in (0, 2):  # natural, reverse Cuthill-McKee
        P = pp.IncompleteCholesky(M, 1e-3, ordering=ordering)
        P.setup()
        cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=2000)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        cg.solve(M, torch.tensor(b, device="cuda"), x)
        xs = x.cpu().numpy()
        if ordering == 0:
            o = ppo.cg(A, b, ppo.Prec(A, "ict", 1e-3), tol=1e-8, maxit=2000)
            xo = o["x"]
        else:
            o = ppo.cg(Ar, b[perm], ppo.Prec(Ar, "ict", 1e-3), tol=1e-8, maxit=2000)
            xo = np.empty(n)
            xo[perm] = o["x"]
        info = P.info()
        print(f"[mtx] ICT ordering {ordering}: {cg.getNumIter()} iterations (oracle {o['iters']}), "
              f"factor nnz {info.mnnz}, levels {info.levels_lower}", flush=True)
        assert cg.getConvergence() and cg.report.true_relres <= 1e-8
        assert abs(cg.getNumIter() - o["iters"]) <= 2
        assert np.linalg.norm(xs - xo) / np.linalg.norm(xo) <= 1e-6
        if ordering == 2:
            assert info.mnnz < nnz_nat  # the band order keeps less fill
        else:
            nnz_nat = info.mnnz

    # nonsymmetric: an upwind matrix in a general file, ILUT + FGMRES(30) in RCM order
    U = ppo.upwind(14, 13, 12, 20.0, 10.0, 5.0)
    nu = len(U[0]) - 1
    write_mtx(tmp_path / "upwind.mtx", U)
    RU = pp.read_mtx(tmp_path / "upwind.mtx")
    for got, want in zip(RU, U):
        np.testing.assert_array_equal(got, want)
    bu = ppo.spmv(U, ppo.uniform(nu, 4))
    MU = pp.CSRMatrix(*RU)
    P = pp.IncompleteLU(MU, lfill=10, tolerance=1e-3, ordering=2)
    gm = pp.GMRES(nu, P, 1e-8, pp.FlexibleGMRES)
    gm.setRestart(True, 30)
    x = torch.zeros(nu, dtype=torch.float64, device="cuda")
    gm.solve(MU, torch.tensor(bu, device="cuda"), x)
    pu = pp.rcm(RU)
    Ur = permute_fast(U, pu)
    o = ppo.gmres(Ur, bu[pu], ppo.Prec(Ur, "ilut", 1e-3, 10), tol=1e-8, restart=30)
    xo = np.empty(nu)
    xo[pu] = o["x"]
    print(f"[mtx] ILUT + FGMRES(30) in RCM order: {gm.getNumIter()} iterations (oracle {o['iters']})", flush=True)
    assert gm.getConvergence() and gm.report.true_relres <= 1e-8
    assert abs(gm.getNumIter() - o["iters"]) <= 2
    assert np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo) <= 1e-6
