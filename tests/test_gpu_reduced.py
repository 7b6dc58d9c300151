"""Reduced (Schur-complement) form of two-colour IC(0)-PCG on the GPU.

For a matrix whose rows split into two colours with diagonal colour blocks,
L^-1 A L^-T is the identity on the first colour and the scaled Schur
complement on the second, so CG on the second colour runs the same Krylov
iteration as IC(0)-PCG.  Checked against the oracle's FlexibleCG + IC(0) on
the colour-permuted matrix (the same ordering the library builds), to the
solver bar: iterations within +-2, true relative residual <= tol,
||x - x_ref|| / ||x_ref|| <= 1e-6; and mixed precision (FP32 reduced inner
solves, FP64 refinement) to the FP64 tolerance.
"""
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


def solve(pp, torch, M, P, b, precision, variant, tol=1e-8, check_every=4, launch=None):
    n = len(b)
    cg = pp.CG(n, P, pp.FlexibleCG, tol=tol, maxIter=3000, precision=precision).setCGVariant(variant)
    cg.opts.check_every = check_every
    if launch is not None:
        cg.opts.launch = launch
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    cg.solve(M, torch.tensor(b, device="cuda"), x)
    return cg, x.cpu().numpy()


@pytest.mark.parametrize("shape", [(20, 19, 18), (64, 63, 1), (33, 34, 35)])
def test_reduced_fp64_matches_ic0_pcg(pp, torch, shape):
    A = ppo.laplacian(*shape)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 11))
    print(f"[reduced] fp64 shape {shape}", flush=True)
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M, ordering=1)
    cg, x = solve(pp, torch, M, P, b, "fp64", "reduced")
    assert cg.report.cg_variant == 3  # PP_CG_REDUCED ran
    assert cg.getConvergence() and cg.report.true_relres < 1e-8
    order, ncol = greedy_order(A)
    assert ncol == 2
    Ap = permute(A, order)
    o = ppo.cg(Ap, b[order], ppo.Prec(Ap, "ic0"), tol=1e-8, maxit=3000)
    xo = np.empty(n)
    xo[order] = o["x"]
    print(f"[reduced] iters {cg.getNumIter()} (outer {cg.report.outer_iters}), oracle IC(0)-PCG {o['iters']}",
          flush=True)
    assert abs(cg.getNumIter() - o["iters"]) <= 2
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-6


@pytest.mark.parametrize("check_every", [1, 4])
def test_reduced_mixed(pp, torch, check_every):
    A = ppo.laplacian(40, 40, 40)
    n = len(A[0]) - 1
    xt = ppo.normal(n, 0)
    b = ppo.spmv(A, xt)
    print(f"[reduced] mixed 40^3 check_every {check_every}", flush=True)
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M, ordering=1, make_f32=True)
    cg, x = solve(pp, torch, M, P, b, "mixed", "auto", check_every=check_every)
    assert cg.report.cg_variant == 3  # auto picks the reduced inner solve here
    assert cg.getConvergence() and cg.report.true_relres < 1e-8
    ref, xr = solve(pp, torch, M, P, b, "fp64", "reference")
    red, _ = solve(pp, torch, M, P, b, "fp64", "reduced", check_every=check_every)
    print(f"[reduced] mixed inner {cg.report.inner_iters} / outer {cg.report.outer_iters}, "
          f"fp64 FlexibleCG {ref.getNumIter()}, fp64 reduced {red.getNumIter()}", flush=True)
    # refinement steps continue one Krylov iteration: about as many
    # iterations in all as the FP64 solve, not a restart per step
    assert cg.report.inner_iters <= red.getNumIter() + 6
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-6
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) <= 1e-5
    # the single-reduction inner solve is still available and agrees
    sr, xs = solve(pp, torch, M, P, b, "mixed", "single_reduction", check_every=check_every)
    assert sr.report.cg_variant == 2
    assert sr.getConvergence() and np.linalg.norm(xs - xr) / np.linalg.norm(xr) <= 1e-6


@pytest.mark.parametrize("rhs", ["normal", "ones", "smooth"])
def test_reduced_mixed_any_rhs(pp, torch, rhs):
    # b not of the form A x_true: for right-hand sides rich in the slowly
    # converging modes the correction is large against its residual and FP32
    # rounding limits what one FP32 step reaches; the solve still converges
    # to the FP64 tolerance in about the FP64 iteration count
    A = ppo.laplacian(48, 48, 48)
    n = len(A[0]) - 1
    if rhs == "normal":
        b = ppo.normal(n, 5)
    elif rhs == "ones":
        b = np.ones(n)
    else:
        g = np.arange(48) / 47.0
        b = (np.sin(np.pi * g)[:, None, None] * np.sin(np.pi * g)[None, :, None] *
             np.sin(np.pi * g)[None, None, :]).ravel() + 1e-3
    print(f"[reduced] mixed 48^3 rhs {rhs}", flush=True)
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M, ordering=1, make_f32=True)
    cg, x = solve(pp, torch, M, P, b, "mixed", "reduced")
    red, xr = solve(pp, torch, M, P, b, "fp64", "reduced")
    print(f"[reduced] rhs {rhs}: mixed inner {cg.report.inner_iters} / outer {cg.report.outer_iters} "
          f"(fp64 steps {cg.report.inner_fp64_steps}), fp64 reduced {red.getNumIter()}", flush=True)
    assert cg.getConvergence() and cg.report.true_relres < 1e-8
    assert cg.report.inner_iters <= 2 * red.getNumIter() + 10
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-6


def test_reduced_repeated_and_graph_reuse(pp, torch):
    # the reduced form and its iteration graphs are built once and reused;
    # a second right-hand side gives its own solution
    from prispack_amd._capi import PP_LAUNCH_GRAPH
    A = ppo.laplacian(24, 24, 24)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M, ordering=1, make_f32=True)
    for seed in (1, 2, 3):
        xt = ppo.normal(n, seed)
        b = ppo.spmv(A, xt)
        for prec in ("mixed", "fp64"):
            cg, x = solve(pp, torch, M, P, b, prec, "reduced", launch=PP_LAUNCH_GRAPH)
            assert cg.getConvergence() and cg.report.true_relres < 1e-8, (seed, prec)
            assert np.linalg.norm(x - xt) / np.linalg.norm(xt) <= 1e-5, (seed, prec)


def test_reduced_unsupported_structures(pp, torch):
    # natural ordering (no two-colour split) and a factor that is not IC(0)
    A = ppo.laplacian(12, 11, 10)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 3))
    M = pp.CSRMatrix(*A)
    for P in (pp.IC0(M, ordering=0), pp.IncompleteCholesky(M, 1e-3)):
        with pytest.raises(pp.PrisPackError) as e:
            solve(pp, torch, M, P, b, "fp64", "reduced")
        assert "PP_ERR_UNSUPPORTED" in str(e.value)
        # auto falls back to the single-reduction inner solve
        cg, _ = solve(pp, torch, M, P, b, "mixed", "auto")
        assert cg.report.cg_variant == 2 and cg.getConvergence()
    # a 27-point-like coupling inside a colour: three colours, no reduced form
    A9 = ppo.laplacian(16, 16, 1)
    rp, ci, v = A9
    rows = [list(zip(ci[rp[i]:rp[i + 1]], v[rp[i]:rp[i + 1]])) for i in range(len(rp) - 1)]
    side = 16
    for i in range(side * side):
        x, y = i % side, i // side
        for dx, dy in ((1, 1), (-1, -1), (1, -1), (-1, 1)):
            if 0 <= x + dx < side and 0 <= y + dy < side:
                rows[i].append((np.uint32((y + dy) * side + x + dx), -0.25))
    for i in range(len(rows)):
        rows[i].sort()
        rows[i] = [(c, (w + 1.0 if c == i else w)) for c, w in rows[i]]
    rp2 = np.zeros(len(rows) + 1, np.uint32)
    rp2[1:] = np.cumsum([len(r) for r in rows])
    ci2 = np.array([c for r in rows for c, _ in r], np.uint32)
    v2 = np.array([w for r in rows for _, w in r])
    M9 = pp.CSRMatrix(rp2, ci2, v2)
    P9 = pp.IC0(M9, ordering=1)
    b9 = ppo.spmv((rp2, ci2, v2), ppo.normal(len(rows), 4))
    with pytest.raises(pp.PrisPackError):
        solve(pp, torch, M9, P9, b9, "fp64", "reduced")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reduced_edge_cases(pp, torch, precision):
    from prispack_amd import PrisPackError
    from prispack_amd._capi import PP_ERR_NOT_FINITE
    # smallest two-colour systems: one row per colour, and 2 + 1
    for dims in ((2, 1, 1), (3, 1, 1), (5, 4, 1)):
        A = ppo.laplacian(*dims)
        n = len(A[0]) - 1
        xt = ppo.normal(n, 8)
        M = pp.CSRMatrix(*A)
        P = pp.IC0(M, ordering=1, make_f32=True)
        cg, x = solve(pp, torch, M, P, ppo.spmv(A, xt), precision, "reduced")
        assert cg.report.cg_variant == 3 and cg.getConvergence(), dims
        assert np.linalg.norm(x - xt) / np.linalg.norm(xt) <= 1e-6, dims
    A = ppo.laplacian(20, 18, 16)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M, ordering=1, make_f32=True)
    # b = 0: converged at once with x = 0
    cg, x = solve(pp, torch, M, P, np.zeros(n), precision, "reduced")
    assert cg.getConvergence() and not np.any(x)
    # a starting guess: the same solution
    xt = ppo.normal(n, 9)
    b = ppo.spmv(A, xt)
    cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=3000, precision=precision).setCGVariant("reduced")
    x0 = torch.tensor(ppo.normal(n, 10), device="cuda")
    cg.solve(M, torch.tensor(b, device="cuda"), x0)
    assert cg.getConvergence() and cg.report.true_relres < 1e-8
    assert np.linalg.norm(x0.cpu().numpy() - xt) / np.linalg.norm(xt) <= 1e-6
    # an iteration limit far below what the tolerance needs: no convergence, no error
    cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-12, maxIter=6, precision=precision).setCGVariant("reduced")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    cg.solve(M, torch.tensor(b, device="cuda"), x)
    assert not cg.getConvergence() and cg.report.inner_iters <= 8
    # a non-finite right-hand side is an error
    bad = b.copy()
    bad[3] = np.nan
    with pytest.raises(PrisPackError) as e:
        cg.solve(M, torch.tensor(bad, device="cuda"), torch.zeros(n, dtype=torch.float64, device="cuda"))
    assert e.value.status == PP_ERR_NOT_FINITE


def test_reduced_stream_launches_match_graphs(pp, torch):
    # the iteration bodies launched on the stream (no graph capture) give
    # the same iterates as the captured, replayed bodies, in both precisions
    from prispack_amd._capi import PP_LAUNCH_GRAPH, PP_LAUNCH_STREAM
    A = ppo.laplacian(30, 29, 28)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 12))
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M, ordering=1, make_f32=True)
    for prec in ("fp64", "mixed"):
        out = []
        for stream in (False, True):
            cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=3000, precision=prec).setCGVariant("reduced")
            cg.opts.launch = PP_LAUNCH_STREAM if stream else PP_LAUNCH_GRAPH
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            cg.solve(M, torch.tensor(b, device="cuda"), x)
            assert cg.getConvergence() and cg.report.true_relres < 1e-8
            out.append((cg.report.inner_iters, x.cpu().numpy()))
        assert out[0][0] == out[1][0], prec
        assert np.array_equal(out[0][1], out[1][1]), prec


def test_reduced_form_follows_value_updates(pp, torch):
    # new values on the same matrix, the preconditioner set up again: the
    # mixed solve must use reduced operators of the new values, i.e. behave
    # as a solve on a freshly created matrix and preconditioner does
    A = ppo.laplacian(24, 23, 22)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 3))
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M, ordering=1, make_f32=True)
    cg0, _ = solve(pp, torch, M, P, b, "mixed", "reduced")
    assert cg0.getConvergence()
    v2 = np.asarray(A[2]) * 1.5
    v2[np.asarray(A[1]) == np.repeat(np.arange(n), np.diff(A[0]))] += 0.25  # a diagonal shift: other iterates
    M.set_values(v2)
    P.setup()
    cg1, x1 = solve(pp, torch, M, P, b, "mixed", "reduced")
    A2 = (A[0], A[1], v2)
    M2 = pp.CSRMatrix(*A2)
    P2 = pp.IC0(M2, ordering=1, make_f32=True)
    cg2, x2 = solve(pp, torch, M2, P2, b, "mixed", "reduced")
    print(f"[reduced] updated values: outer {cg1.report.outer_iters} inner {cg1.report.inner_iters}; "
          f"fresh outer {cg2.report.outer_iters} inner {cg2.report.inner_iters}", flush=True)
    assert cg1.getConvergence() and cg1.report.true_relres < 1e-8
    assert (cg1.report.outer_iters, cg1.report.inner_iters) == (cg2.report.outer_iters, cg2.report.inner_iters)
    np.testing.assert_array_equal(x1, x2)


def test_reduced_form_device_build_matches_host(pp, torch, monkeypatch):
    # the reduced form built on the device from the colour-ordered matrix's
    # device arrays (64^3 class: its sliced-column copy exists) gives the
    # same mixed and FP64 reduced solves, bit for bit, as the host
    # construction (PP_REDUCED_HOST_BUILD=1); variable values D A D + I/4
    A = ppo.laplacian(64, 63, 62)
    n = len(A[0]) - 1
    rows = np.repeat(np.arange(n), np.diff(A[0]))
    cols = np.asarray(A[1], np.int64)
    rng = np.random.default_rng(5)
    dsc = 1.0 + 0.4 * rng.uniform(-1, 1, n)
    v = np.asarray(A[2], np.float64) * dsc[rows] * dsc[cols]
    v[A[1] == rows] += 0.25
    B = (A[0], A[1], v)
    b = ppo.spmv(B, ppo.normal(n, 21))
    out = {}
    for where in ("host", "device"):
        if where == "host":
            monkeypatch.setenv("PP_REDUCED_HOST_BUILD", "1")
        else:
            monkeypatch.delenv("PP_REDUCED_HOST_BUILD", raising=False)
        M = pp.CSRMatrix(*B)
        P = pp.IC0(M, ordering=1, make_f32=True)
        res = []
        for prec, var in (("mixed", "auto"), ("fp64", "reduced")):
            cg, x = solve(pp, torch, M, P, b, prec, var)
            assert cg.getConvergence() and cg.report.cg_variant == 3, (where, prec)
            res.append((cg.report.outer_iters, cg.report.inner_iters, cg.getNumIter(), x))
        assert P.info().reduced_form == (1 if where == "host" else 2), where
        out[where] = res
    print(f"[reduced] device build: {out['device'][0][:3]} / {out['device'][1][:3]}", flush=True)
    for h, d in zip(out["host"], out["device"]):
        assert h[:3] == d[:3]
        np.testing.assert_array_equal(h[3], d[3])
