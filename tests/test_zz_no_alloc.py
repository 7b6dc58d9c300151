"""pp_solve_noalloc (the FP64 CG solve whose work vectors live in one fixed
device array) through the C ABI, against pp_solve on the same inputs and
options: the same kernels in the same order, so the same iteration count and
the same x bit for bit.  Two calls in a row, the second on a smaller system;
a system too large for the array and an unsupported solver kind are refused
before any work, and the array's earlier contents never leak into a result."""
import ctypes as C

import numpy as np
import pytest

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


def run(pp, torch, fn, M, P, b, variant):
    from prispack_amd import _capi
    L = _capi.lib()
    o = _capi.SolveOpts()
    L.pp_solve_opts_default(C.byref(o), _capi.PP_SOLVER_CG)
    o.tol, o.maxit, o.cg_variant = 1e-10, 3000, variant
    o.launch = _capi.PP_LAUNCH_STREAM  # what the zero-allocation solve runs
    o.record_history = 0
    rep = _capi.SolveReport()
    db = torch.tensor(b, device="cuda")
    x = torch.full((len(b),), 0.25, dtype=torch.float64, device="cuda")  # a non-zero x0
    torch.cuda.synchronize()
    st = getattr(L, fn)(M.ctx.h, M.h, P.h if P is not None else None, C.byref(o), C.c_void_p(db.data_ptr()),
                        C.c_void_p(x.data_ptr()), C.byref(rep))
    return st, rep, x.cpu().numpy()


CASES = [  # (grid, preconditioner, CG variant): the second call of each pair is smaller
    ((40, 40, 40), "ic0_mc", 1),
    ((21, 20, 19), "ict", 1),
    ((30, 30, 30), "ildlt", 2),
    ((50, 50, 1), None, 1),
]


def make(pp, A, kind):
    M = pp.CSRMatrix(*A)
    if kind == "ic0_mc":
        P = pp.IC0(M, ordering=1)
    elif kind == "ict":
        P = pp.IncompleteCholesky(M, 1e-3)
    elif kind == "ildlt":
        P = pp.IncompleteLDLT(M, 1e-3)
    else:
        P = None
    if P is not None:
        P.setup()
    return M, P


def test_noalloc_matches_pp_solve(pp, torch):
    for dims, kind, variant in CASES:
        A = ppo.laplacian(*dims)
        n = len(A[0]) - 1
        b = ppo.spmv(A, ppo.normal(n, sum(dims)))
        M, P = make(pp, A, kind)
        print(f"[noalloc] {dims} {kind} variant {variant}", flush=True)
        st0, r0, x0 = run(pp, torch, "pp_solve", M, P, b, variant)
        st1, r1, x1 = run(pp, torch, "pp_solve_noalloc", M, P, b, variant)
        assert st0 == 0 and st1 == 0, (st0, st1)
        assert r1.converged and r1.iters == r0.iters and r1.true_relres <= 1e-10
        np.testing.assert_array_equal(x1, x0)
        # a second call right after, on the same system: the array still holds the
        # first call's vectors, and the result does not depend on them
        st2, r2, x2 = run(pp, torch, "pp_solve_noalloc", M, P, b, variant)
        assert st2 == 0 and r2.iters == r0.iters
        np.testing.assert_array_equal(x2, x0)


def test_noalloc_refusals(pp, torch):
    from prispack_amd import _capi
    L = _capi.lib()
    # too large for the array: a colour-ordered IC(0) solve keeps seven vectors
    # of n doubles, each rounded up to 256 bytes; 2,396,746 rows need 134,219,008
    # bytes against the 134,217,728 of the array
    big = 2_396_746
    i = np.arange(big)
    rows = np.concatenate([i[1:], i, i[:-1]])
    cols = np.concatenate([i[:-1], i, i[1:]])
    vals = np.concatenate([np.full(big - 1, -1.0), np.full(big, 4.0), np.full(big - 1, -1.0)])
    o = np.lexsort((cols, rows))
    rp = np.zeros(big + 1, np.uint32)
    rp[1:] = np.cumsum(np.bincount(rows, minlength=big))
    Mb = pp.CSRMatrix(rp, cols[o].astype(np.uint32), vals[o])
    Pb = pp.IC0(Mb, ordering=1)
    Pb.setup()
    b = np.ones(big)
    st, _, _ = run(pp, torch, "pp_solve_noalloc", Mb, Pb, b, 1)
    assert st == _capi.PP_ERR_INVALID
    assert b"scratch" in L.pp_last_error(Mb.ctx.h)
    # without the preconditioner five vectors fit, and the solve runs
    st, rep, x = run(pp, torch, "pp_solve_noalloc", Mb, None, b, 1)
    assert st == 0 and rep.converged and rep.true_relres <= 1e-10
    st0, rep0, x0 = run(pp, torch, "pp_solve", Mb, None, b, 1)
    assert st0 == 0 and rep0.iters == rep.iters
    np.testing.assert_array_equal(x, x0)
    M = pp.CSRMatrix(*ppo.laplacian(3, 3, 1))
    # mixed precision and GMRES are not zero-allocation solves
    o = _capi.SolveOpts()
    L.pp_solve_opts_default(C.byref(o), _capi.PP_SOLVER_GMRES)
    rep = _capi.SolveReport()
    bb = torch.ones(9, dtype=torch.float64, device="cuda")
    xx = torch.zeros(9, dtype=torch.float64, device="cuda")
    st = L.pp_solve_noalloc(M.ctx.h, M.h, None, C.byref(o), C.c_void_p(bb.data_ptr()), C.c_void_p(xx.data_ptr()),
                            C.byref(rep))
    assert st == _capi.PP_ERR_UNSUPPORTED
