"""HIP path through the C ABI against the oracle (oracle/ppo.c) and the golden
vectors of the compiled reference.  Bars:
  - SpMV / SpMV^T / the triangular sweeps in FP64: bit-identical (same
    operation order as the reference's serial loops, no fused multiply-add);
  - FP32 kernels: |err| <= 1e-5 relative to the FP64 result;
  - solvers (FP64): iterations within +-2 of the reference, true relative
    residual <= tol, ||x - x_ref|| / ||x_ref|| <= 1e-6.
"""
import os

import numpy as np
import pytest

from oracle import ppo

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


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


def dev(torch, a, dtype=None):
    return torch.tensor(np.asarray(a), dtype=dtype or torch.float64, device="cuda")


def host(t):
    return t.detach().cpu().numpy()


def long_row_matrix():
    # rows of very different lengths, including empty rows and one row longer
    # than the LDS stage of the SpMV kernel
    rng = np.random.default_rng(5)
    n = 3000
    rows = []
    for i in range(n):
        if i % 97 == 5:
            cols = np.array([], dtype=np.int64)
        elif i == 17:
            cols = np.arange(0, n, 1)
        else:
            cols = np.unique(rng.integers(0, n, size=rng.integers(1, 40)))
        rows.append(cols)
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows).astype(np.uint32)
    v = rng.standard_normal(len(ci))
    return rp, ci, v


def wide_span_matrix():
    # a banded matrix (one 15-bit window per row block) with long-range
    # couplings: some row blocks need two windows (two-base 16-bit offsets),
    # some three (32-bit indices)
    rng = np.random.default_rng(9)
    n = 200_000
    rows = []
    for i in range(n):
        cols = {max(0, i - 1), i, min(n - 1, i + 1)}
        if (i // 256) % 5 == 2 and i % 7 == 0:
            cols.add((i + 90_000) % n)
        if (i // 256) % 5 == 4 and i % 7 == 0:
            cols.add((i + 60_000) % n)
            cols.add((i + 130_000) % n)
        rows.append(np.array(sorted(cols), dtype=np.int64))
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows).astype(np.uint32)
    v = rng.standard_normal(len(ci))
    return rp, ci, v


MATS = {
    "lap3d_20": lambda: ppo.laplacian(20, 20, 20),
    "wide_span": wide_span_matrix,
    "lap2d_128": lambda: ppo.laplacian(128, 128, 1),
    "upwind_17": lambda: ppo.upwind(17, 17, 17, 20.0, 10.0, 5.0),
    "ragged": long_row_matrix,
}


@pytest.mark.parametrize("name", sorted(MATS))
def test_spmv_fp64_bit_exact(pp, torch, name):
    A = MATS[name]()
    n = len(A[0]) - 1
    print(f"[spmv] {name} n={n}", flush=True)
    M = pp.CSRMatrix(*A)
    x = ppo.uniform(n, 3)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    M.MatPVec(dev(torch, x), y)
    want = ppo.spmv(A, x)
    if name == "ragged":  # the long row is summed by a tree: FP64 rounding only
        long = np.zeros(n, bool)
        long[17] = True
        assert np.array_equal(host(y)[~long], want[~long])
        np.testing.assert_allclose(host(y)[long], want[long], rtol=1e-13)
    else:
        assert np.array_equal(host(y), want)
    # alpha / beta epilogue: y = -A x + b
    b = ppo.uniform(n, 4)
    yb = dev(torch, b)
    M.MatPVec(dev(torch, x), yb, alpha=-1.0, beta=1.0)
    np.testing.assert_allclose(host(yb), b - want, rtol=0, atol=1e-13 * (1 + np.abs(want).max()))


@pytest.mark.parametrize("name", ["lap3d_20", "upwind_17", "ragged", "wide_span"])
def test_spmv_t_and_fp32(pp, torch, name):
    A = MATS[name]()
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    x = ppo.uniform(n, 8)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    M.transMatPVec(dev(torch, x), y, alpha=1.0, beta=0.0)
    assert np.array_equal(host(y), ppo.spmv_t(A, x))
    y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    M.MatPVec(dev(torch, x, torch.float32), y32)
    ref = ppo.spmv(A, x)
    scale = np.abs(A[2]).max() * np.abs(x).max() * 40
    np.testing.assert_allclose(host(y32), ref, rtol=0, atol=1e-6 * scale)


@pytest.mark.parametrize("name", ["lap3d_40", "upwind_33"])
def test_spmv_variants_and_solves_agree(pp, torch, name):
    # every SpMV variant (CSR-stream 0-7, sliced-column 8, which a matrix
    # this size holds) gives the oracle's y bit for bit in FP64 and one FP32
    # result; a solve differs only through the inner products' partial sums
    import ctypes as C
    from prispack_amd import _capi
    L = _capi.lib()
    A = ppo.laplacian(40, 40, 40) if name == "lap3d_40" else ppo.upwind(33, 33, 33, 20.0, 10.0, 5.0)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    x = ppo.uniform(n, 21)
    want = ppo.spmv(A, x)
    y32s, runs = [], []
    for var in range(9):
        st = L.pp_csr_set_spmv_variant(M.h, 0, var)
        if st != 0:
            assert (var & 2) == 2 or var == 8
            continue
        assert L.pp_csr_set_spmv_variant(M.h, 1, var) == 0
        y = torch.zeros(n, dtype=torch.float64, device="cuda")
        M.MatPVec(dev(torch, x), y)
        assert np.array_equal(host(y), want), var
        b = ppo.uniform(n, 22)
        yb = dev(torch, b)
        M.MatPVec(dev(torch, x), yb, alpha=-1.0, beta=1.0)
        np.testing.assert_allclose(host(yb), b - want, rtol=0, atol=1e-13 * (1 + np.abs(want).max()))
        y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
        M.MatPVec(dev(torch, x, torch.float32), y32)
        y32s.append(host(y32))
        rhs = dev(torch, want)
        if name == "lap3d_40":
            P = pp.IC0(M, make_f32=True)
            s = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=1000)
        else:
            P = pp.ILU0(M, make_f32=True)
            s = pp.GMRES(n, P, 1e-8, pp.FlexibleGMRES, maxIter=1000)
            s.setRestart(True, 30)
        xs = torch.zeros(n, dtype=torch.float64, device="cuda")
        s.solve(M, rhs, xs)
        assert s.getConvergence(), var
        runs.append((var, s.getNumIter(), host(xs)))
        print(f"[spmv variant] {name} {var}: {s.getNumIter()} iterations", flush=True)
    assert any(v == 8 for v, _, _ in runs)  # the sliced-column copy exists at this size
    for y32 in y32s[1:]:
        assert np.array_equal(y32, y32s[0])
    for var, its, xs in runs[1:]:
        assert abs(its - runs[0][1]) <= 1, var
        assert np.linalg.norm(xs - runs[0][2]) <= 1e-7 * np.linalg.norm(runs[0][2]), var


def random_band(n, per_row, span, seed):
    # per_row distinct sorted columns within +-span of each row (diagonal
    # included): offsets that almost never repeat from row to row.  Row i's
    # other columns come one from each of per_row - 1 equal segments of its
    # window (so they are distinct), moved off the diagonal within a segment.
    rng = np.random.default_rng(seed)
    k = per_row - 1
    i = np.arange(n, dtype=np.int64)
    lo, hi = np.maximum(0, i - span), np.minimum(n, i + span + 1)
    w = (hi - lo) // k
    seg = lo[:, None] + np.arange(k)[None, :] * w[:, None]
    c = seg + (rng.random((n, k)) * w[:, None]).astype(np.int64)
    hit = c == i[:, None]
    c[hit] = np.where(c[hit] + 1 < (seg + w[:, None])[hit], c[hit] + 1, c[hit] - 1)
    ci = np.sort(np.concatenate([c, i[:, None]], axis=1), axis=1).astype(np.uint32).ravel()
    rp = np.arange(n + 1, dtype=np.uint32) * per_row
    v = rng.uniform(-1.0, 1.0, n * per_row)
    return rp, ci, v


@pytest.mark.parametrize("name", ["lap3d_40", "lap3d_41", "random_7", "rb_lap3d_40", "rb_upwind_33"])
def test_spmv_sliced_column_offset_forms(pp, torch, name):
    # the sliced-column copy stores a stencil's offsets as one pattern byte
    # per row (at most 256 distinct offset rows) and anything else as 16-bit
    # offsets per entry; both give the oracle's y bit for bit, with two rows
    # per thread (even n) and one (odd n)
    import ctypes as C
    from prispack_amd import _capi
    L = _capi.lib()
    if name == "random_7":
        A = random_band(40000, 7, 20000, 5)
    elif name.startswith("rb_"):
        # red-black (colour) order: every column about n/2 from its row, the
        # two-colour form with per-colour bases and the diagonal apart
        from mc_order import parity_order, permute_fast
        g = int(name[-2:])
        B = ppo.laplacian(g, g, g) if "lap" in name else ppo.upwind(g, g, g, 20.0, 10.0, 5.0)
        A = permute_fast(B, parity_order(g, g, g))
    else:
        g = int(name[-2:])
        A = ppo.laplacian(g, g, g)
    n = len(A[0]) - 1
    print(f"[ell forms] {name} n={n}", flush=True)
    M = pp.CSRMatrix(*A)
    for dt in (0, 1):
        assert L.pp_csr_set_spmv_variant(M.h, dt, 8) == 0
    nb = C.c_uint64()
    assert L.pp_csr_spmv_bytes(M.h, 0, C.byref(nb)) == 0
    dense_offsets = 10 * 7 * n + 16 * n
    if name == "random_7":
        assert nb.value == dense_offsets
    elif name.startswith("rb_"):
        assert nb.value < 8 * 6 * n + 8 * n + 2 * n + 16 * n  # six slots, the diagonal, a byte per row
    else:
        assert nb.value < 8 * 7 * n + 2 * n + 16 * n  # a byte per row, a small table
    x = ppo.uniform(n, 41)
    want = ppo.spmv(A, x)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    M.MatPVec(dev(torch, x), y)
    assert np.array_equal(host(y), want)
    b = ppo.uniform(n, 42)
    yb = dev(torch, b)
    M.MatPVec(dev(torch, x), yb, alpha=-1.0, beta=1.0)
    np.testing.assert_allclose(host(yb), b - want, rtol=0, atol=1e-13 * (1 + np.abs(want).max()))
    y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    M.MatPVec(dev(torch, x, torch.float32), y32)
    np.testing.assert_allclose(host(y32), want, rtol=0, atol=1e-5 * (1 + np.abs(want).max()))
    # new values on the same pattern refill the copy (its diagonal included)
    v2 = np.asarray(A[2], np.float64) * (1.0 + 0.5 * ppo.uniform(len(A[2]), 43))
    M.set_values(v2)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    M.MatPVec(dev(torch, x), y)
    assert np.array_equal(host(y), ppo.spmv((A[0], A[1], v2), x))


def test_sliced_column_copy_follows_value_updates(pp, torch):
    # new values on the same pattern (on the host, then on the device) reach
    # the sliced-column copy in both precisions
    import ctypes as C
    from prispack_amd import _capi
    L = _capi.lib()
    A = ppo.laplacian(30, 30, 30)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    for dt in (0, 1):
        assert L.pp_csr_set_spmv_variant(M.h, dt, 8) == 0
    x = ppo.uniform(n, 31)
    y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    M.MatPVec(dev(torch, x, torch.float32), y32)  # the FP32 copies exist before the updates
    for k, on_device in ((2.0, False), (0.5, True)):
        v = A[2] * k
        vals = dev(torch, v) if on_device else v
        M.set_values(vals)
        y = torch.zeros(n, dtype=torch.float64, device="cuda")
        M.MatPVec(dev(torch, x), y)
        want = ppo.spmv((A[0], A[1], v), x)
        assert np.array_equal(host(y), want), k
        y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
        M.MatPVec(dev(torch, x, torch.float32), y32)
        np.testing.assert_allclose(host(y32), want, rtol=0, atol=1e-5 * (1 + np.abs(want).max()))


def test_sweep_variants_agree_on_wide_levels(pp, torch):
    # colour levels of 280k-560k rows: the staged kernel's variants and the
    # sliced-column form (8) give one apply bit for bit, in FP64 and FP32,
    # and one FlexibleCG solve
    import ctypes as C
    from prispack_amd import _capi
    L = _capi.lib()
    A = ppo.laplacian(104, 104, 104)
    n = len(A[0]) - 1
    M = pp.CSRMatrix(*A)
    x = ppo.uniform(n, 41)
    b = ppo.spmv(A, ppo.normal(n, 42))
    ref = None
    for var in (0, 6, 7, 8):
        P = pp.IC0(M, ordering=1, make_f32=True)
        P.setup()
        assert L.pp_precond_set_sweep_variant(P.h, var) == 0, var
        assert P.info().sweep_variant == var
        y = P.MInvSolve(dev(torch, x), torch.zeros(n, dtype=torch.float64, device="cuda"))
        y32 = P.MInvSolve(dev(torch, x, torch.float32), torch.zeros(n, dtype=torch.float32, device="cuda"))
        cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=1000)
        xs = torch.zeros(n, dtype=torch.float64, device="cuda")
        cg.solve(M, dev(torch, b), xs)
        got = (host(y), host(y32), cg.getNumIter(), host(xs))
        print(f"[sweep variant] {var}: {cg.getNumIter()} iterations", flush=True)
        if ref is None:
            ref = got
            continue
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), var
        assert got[2] == ref[2] and np.array_equal(got[3], ref[3]), var
    assert L.pp_precond_set_sweep_variant(P.h, 9) != 0


def test_blas1(pp, torch):
    from prispack_amd import _capi
    import ctypes as C
    ctx = pp.Context.default()
    L = _capi.lib()
    for n in (1, 1000, 1_000_003):
        x, y = ppo.uniform(n, 1), ppo.uniform(n, 2)
        tx, ty = dev(torch, x), dev(torch, y)
        torch.cuda.synchronize()
        out = C.c_double()
        ctx.check(L.pp_dot(ctx.h, 0, n, C.c_void_p(tx.data_ptr()), C.c_void_p(ty.data_ptr()), C.byref(out)))
        assert abs(out.value - x @ y) <= 1e-13 * n
        ctx.check(L.pp_nrm2(ctx.h, 0, n, C.c_void_p(tx.data_ptr()), C.byref(out)))
        assert abs(out.value - np.linalg.norm(x)) <= 1e-13 * np.sqrt(n)
        tz = torch.empty_like(tx)
        ctx.check(L.pp_axpby(ctx.h, 0, n, 2.0, C.c_void_p(tx.data_ptr()), -3.0, C.c_void_p(ty.data_ptr()),
                             C.c_void_p(tz.data_ptr())))
        ctx.sync()
        assert np.array_equal(host(tz), 2.0 * x + -3.0 * y)


# --------------------------------------------------------------------------- preconditioner sweeps
PRECS = [("ict", (1e-3,), "lap"), ("ict", (0.0,), "lap_small"), ("ict", (1e-2,), "lap"), ("ildlt", (1e-3,), "lap"),
         ("ic0", (), "lap"), ("ilut", (1e-2, 20), "upw"), ("ilut", (0.0, 3), "upw"), ("ilu0", (), "upw")]


def prec_matrix(which):
    return {"lap": ppo.laplacian(16, 15, 14), "lap_small": ppo.laplacian(9, 8, 7),
            "upw": ppo.upwind(16, 15, 14, 20.0, 10.0, 5.0)}[which]


def make_prec(pp, A, kind, args, trsv=0):
    cls = {"ict": pp.IncompleteCholesky, "ildlt": pp.IncompleteLDLT, "ic0": pp.IC0, "ilu0": pp.ILU0}
    if kind == "ilut":
        P = pp.IncompleteLU(pp.CSRMatrix(*A), lfill=args[1], tolerance=args[0], trsv=trsv)
    else:
        P = cls[kind](pp.CSRMatrix(*A), *args, trsv=trsv)
    P.setup()
    return P


TRSV = {"levels": 1, "syncfree": 2, "block": 3, "pipe": 4}


@pytest.mark.parametrize("trsv", ["levels", "syncfree", "block", "pipe"])
@pytest.mark.parametrize("kind,args,mat", PRECS)
def test_precond_apply_bit_exact(pp, torch, kind, args, mat, trsv):
    A = prec_matrix(mat)
    n = len(A[0]) - 1
    print(f"[precond] {kind}{args} {mat} n={n} {trsv}", flush=True)
    P = make_prec(pp, A, kind, args, TRSV[trsv])
    Q = ppo.Prec(A, kind, *args)
    assert P.getPreconditionNonZeroCount() == Q.mnnz
    x = ppo.uniform(n, 21)
    t = dev(torch, x)
    P.MInvSolve(t)
    assert np.array_equal(host(t), Q.apply(x))
    for part in ("lower", "diag", "upper"):
        t = dev(torch, x)
        getattr(P, {"lower": "MSolveLower", "diag": "MSolveDiagonal", "upper": "MSolveUpper"}[part])(t)
        assert np.array_equal(host(t), getattr(Q, part)(x)), part
    # out-of-place form and FP32
    out = torch.zeros(n, dtype=torch.float64, device="cuda")
    P.MInvSolve(dev(torch, x), out)
    assert np.array_equal(host(out), Q.apply(x))
    t32 = dev(torch, x, torch.float32)
    P.MInvSolve(t32)
    ref = Q.apply(x)
    np.testing.assert_allclose(host(t32), ref, rtol=0, atol=2e-5 * np.abs(ref).max())


def test_precond_errors(pp, torch):
    from prispack_amd import PrisPackError
    from prispack_amd._capi import PP_ERR_NOT_READY, PP_ERR_NOT_SPD
    A = ppo.laplacian(6, 6, 1, 0, 0, 0, 6.0)
    P = pp.IncompleteCholesky(pp.CSRMatrix(*A), 1e-3)
    with pytest.raises(PrisPackError) as e:
        P.MInvSolve(dev(torch, np.ones(36)))
    assert e.value.status == PP_ERR_NOT_READY
    with pytest.raises(PrisPackError) as e:
        P.setup()
    assert e.value.status == PP_ERR_NOT_SPD


# --------------------------------------------------------------------------- solvers
def _golden(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz")))


@pytest.mark.parametrize("launch", ["stream", "graph"])
@pytest.mark.parametrize("trsv", ["levels", "syncfree"])
@pytest.mark.parametrize("name", ["cg_128x128x1_tau0.001", "cg_128x128x1_tau0.01", "cg_12x12x12_tau0.001",
                                  "cg_20x20x20_tau0.01", "cg_16x16x1_tau0.001"])
def test_cg_parity_with_reference(pp, torch, name, trsv, launch):
    from prispack_amd._capi import PP_LAUNCH_GRAPH, PP_LAUNCH_STREAM
    g = _golden(name)
    dims = tuple(int(t) for t in g["dims"])
    A = ppo.laplacian(*dims)
    n = len(A[0]) - 1
    print(f"[cg] {name} n={n} {trsv}", flush=True)
    xt = ppo.uniform(n, int(g["seed"]))
    b = ppo.spmv(A, xt)
    M = pp.CSRMatrix(*A)
    P = pp.IncompleteCholesky(M, float(g["tau"]), trsv=TRSV[trsv])
    cg = pp.CG(n, P, pp.FlexibleCG, tol=float(g["tol"]), maxIter=int(g["maxit"]))
    cg.opts.launch = PP_LAUNCH_GRAPH if launch == "graph" else PP_LAUNCH_STREAM
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    cg.solve(M, dev(torch, b), x)
    xs = host(x)
    assert abs(cg.getNumIter() - int(g["iters"])) <= 2
    assert cg.getConvergence()
    assert cg.report.true_relres <= float(g["tol"])
    xr = g["x"]
    assert np.linalg.norm(xs - xr) / np.linalg.norm(xr) <= 1e-6
    # the residual history follows the oracle's to rounding
    o = ppo.cg(A, b, ppo.Prec(A, "ict", float(g["tau"])), tol=float(g["tol"]), maxit=int(g["maxit"]))
    h = cg.history()
    k = min(len(h), len(o["hist"]))
    np.testing.assert_allclose(h[:k], o["hist"][:k], rtol=1e-6)


def test_normal_cg_and_host_boundary(pp, torch):
    g = _golden("cg_16x16x1_tau0.001")
    A = ppo.laplacian(16, 16, 1)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.uniform(n, 0))
    M = pp.CSRMatrix(*A)
    P = pp.IncompleteCholesky(M, 1e-3)
    cg = pp.CG(n, P, pp.NormalCG, tol=1e-8, maxIter=1000)
    x = np.zeros(n)
    cg.solve(M, b, x)  # host in / host out
    assert abs(cg.getNumIter() - int(g["normal_iters"])) <= 2


def test_cg_limits_and_errors(pp, torch):
    from prispack_amd import PrisPackError
    from prispack_amd._capi import PP_ERR_NOT_FINITE
    A = ppo.laplacian(30, 30, 1)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.uniform(n, 0))
    M = pp.CSRMatrix(*A)
    P = pp.IncompleteCholesky(M, 1e-2)
    o = ppo.cg(A, b, ppo.Prec(A, "ict", 1e-2), tol=1e-12, maxit=5)
    cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-12, maxIter=5)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    cg.solve(M, dev(torch, b), x)
    assert not cg.getConvergence() and cg.getNumIter() == o["iters"] == 6
    assert abs(cg.getError() - o["error"]) <= 1e-10 * abs(o["error"])
    bad = b.copy()
    bad[7] = np.nan
    with pytest.raises(PrisPackError) as e:
        cg.solve(M, dev(torch, bad), torch.zeros(n, dtype=torch.float64, device="cuda"))
    assert e.value.status == PP_ERR_NOT_FINITE


@pytest.mark.parametrize("trsv", ["levels", "syncfree"])
@pytest.mark.parametrize("dims", [(64, 64, 1), (24, 24, 24)])
def test_mixed_precision_cg(pp, torch, dims, trsv):
    A = ppo.laplacian(*dims)
    n = len(A[0]) - 1
    xt = ppo.normal(n, 0)
    b = ppo.spmv(A, xt)
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M, trsv=TRSV[trsv])
    x64 = torch.zeros(n, dtype=torch.float64, device="cuda")
    c64 = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=2000)
    c64.solve(M, dev(torch, b), x64)
    xm = torch.zeros(n, dtype=torch.float64, device="cuda")
    cm = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=2000, precision="mixed")
    cm.solve(M, dev(torch, b), xm)
    assert cm.getConvergence() and cm.report.true_relres < 1e-8
    assert cm.report.outer_iters >= 2
    # against the FP64 oracle solve with the same IC(0) factor
    o = ppo.cg(A, b, ppo.Prec(A, "ic0"), tol=1e-8, maxit=2000)
    xo = o["x"]
    assert abs(c64.getNumIter() - o["iters"]) <= 2
    assert np.linalg.norm(host(x64) - xo) / np.linalg.norm(xo) <= 1e-6
    err = np.linalg.norm(host(xm) - xo) / np.linalg.norm(xo)
    print(f"mixed: outer {cm.report.outer_iters} inner {cm.report.inner_iters} fp64 {c64.getNumIter()} err {err:.2e}")
    assert err <= 1e-6


@pytest.mark.parametrize("launch", ["stream", "graph"])
@pytest.mark.parametrize("trsv", ["levels", "syncfree"])
@pytest.mark.parametrize("name", ["gmres_lapconv_12x12x12", "gmres_upwind_16x16x16", "gmres_upwind_10x10x10_l3"])
def test_gmres_parity_with_reference(pp, torch, name, trsv, launch):
    # graph launches capture the FGMRES cycle before its first apply: the
    # sync-free schedule's buffers must exist from setup on
    from prispack_amd._capi import PP_LAUNCH_GRAPH, PP_LAUNCH_STREAM
    g = _golden(name)
    A = (g["rp"], g["ci"], g["v"])
    n = len(A[0]) - 1
    print(f"[gmres] {name} n={n} {trsv}", flush=True)
    xt = ppo.uniform(n, int(g["seed"]))
    b = ppo.spmv(A, xt)
    M = pp.CSRMatrix(*A)
    P = pp.IncompleteLU(M, lfill=int(g["lfil"]), tolerance=float(g["tau"]), trsv=TRSV[trsv])
    for quirks in (True, False):
        gm = pp.GMRES(n, P, float(g["tol"]), pp.FlexibleGMRES, reference_quirks=quirks)
        gm.setRestart(True, int(g["restart"]))
        gm.opts.launch = PP_LAUNCH_GRAPH if launch == "graph" else PP_LAUNCH_STREAM
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        gm.solve(M, dev(torch, b), x)
        xs = host(x)
        assert abs(gm.getNumIter() - int(g["iters"])) <= 2, (quirks, gm.getNumIter(), int(g["iters"]))
        assert gm.getConvergence()
        xr = g["x"]
        assert np.linalg.norm(xs - xr) / np.linalg.norm(xr) <= 1e-6
        if not quirks:  # the fixed back-substitution meets the tolerance on the true residual
            assert gm.report.true_relres <= float(g["tol"])


def test_gmres_mixed(pp, torch):
    A = ppo.upwind(24, 24, 24, 20.0, 10.0, 5.0)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.uniform(n, 0))
    M = pp.CSRMatrix(*A)
    P = pp.ILU0(M)
    gm = pp.GMRES(n, P, 1e-8, pp.FlexibleGMRES, precision="mixed")
    gm.setRestart(True, 30)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    gm.solve(M, dev(torch, b), x)
    assert gm.getConvergence() and gm.report.true_relres <= 1e-8
    o = ppo.gmres(A, b, ppo.Prec(A, "ilu0"), tol=1e-10, restart=30)
    err = np.linalg.norm(host(x) - o["x"]) / np.linalg.norm(o["x"])
    assert err <= 1e-6


def test_single_reduction_fp64_syncfree(pp, torch):
    # the single-reduction body's first apply is inside the captured graph
    from prispack_amd._capi import PP_LAUNCH_GRAPH
    A = ppo.laplacian(16, 16, 12)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 2))
    M = pp.CSRMatrix(*A)
    out = {}
    for trsv in ("levels", "syncfree"):
        P = pp.IncompleteCholesky(M, 1e-3, trsv=TRSV[trsv])
        cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-10, maxIter=2000).setCGVariant("single_reduction")
        cg.opts.launch = PP_LAUNCH_GRAPH
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        cg.solve(M, dev(torch, b), x)
        assert cg.getConvergence() and cg.report.true_relres < 1e-10, trsv
        out[trsv] = (cg.getNumIter(), host(x))
    # the schedules apply the same factor
    assert abs(out["levels"][0] - out["syncfree"][0]) <= 1
    xl, xs = out["levels"][1], out["syncfree"][1]
    assert np.linalg.norm(xs - xl) / np.linalg.norm(xl) <= 1e-9


@pytest.mark.parametrize("kind", ["ic0", "ict"])
def test_single_reduction_pcg(pp, torch, kind):
    # Chronopoulos-Gear PCG: same Krylov iterates as CG in exact arithmetic, so
    # iteration counts within a few of FlexibleCG and the same solution
    A = ppo.laplacian(24, 22, 20)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 4))
    M = pp.CSRMatrix(*A)
    P = pp.IC0(M) if kind == "ic0" else pp.IncompleteCholesky(M, 1e-3)
    out = {}
    for variant in ("reference", "single_reduction"):
        cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-10, maxIter=2000).setCGVariant(variant)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        cg.solve(M, dev(torch, b), x)
        assert cg.getConvergence() and cg.report.true_relres < 1e-9, variant
        out[variant] = (cg.getNumIter(), host(x))
    print(f"[sr] {kind}: reference {out['reference'][0]}, single reduction {out['single_reduction'][0]}", flush=True)
    assert abs(out["reference"][0] - out["single_reduction"][0]) <= 3
    xr, xs = out["reference"][1], out["single_reduction"][1]
    assert np.linalg.norm(xs - xr) / np.linalg.norm(xr) <= 1e-7
    # mixed precision with each inner variant reaches the FP64 tolerance
    for variant in ("reference", "single_reduction"):
        cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=2000, precision="mixed").setCGVariant(variant)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        cg.solve(M, dev(torch, b), x)
        assert cg.getConvergence() and cg.report.true_relres < 1e-8, variant


@pytest.mark.parametrize("kind,args", [("ict", (1e-3,)), ("ilut", (1e-3, 10)), ("ildlt", (1e-3,))])
def test_block_schedule_solves(pp, torch, kind, args):
    # one-workgroup sweeps (PP_TRSV_BLOCK): the applies are bit-identical to
    # the level schedule, so the FP64 solves are too; AUTO reports what it chose
    A = ppo.laplacian(64, 64, 1) if kind != "ilut" else ppo.upwind(64, 64, 1, 20.0, 10.0, 5.0)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 6))
    print(f"[block] {kind} n={n}", flush=True)
    M = pp.CSRMatrix(*A)
    out = {}
    for trsv in (1, 3, 4, 0):
        P = make_prec(pp, A, kind, args, trsv)
        info = P.info()
        assert info.trsv == (trsv if trsv else info.trsv) and info.trsv in (1, 3, 4)
        if kind == "ilut":
            s = pp.GMRES(n, P, 1e-8, pp.FlexibleGMRES, maxIter=3000)
            s.setRestart(True, 30)
        else:
            s = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=2000)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        s.solve(M, dev(torch, b), x)
        assert s.getConvergence() and s.report.true_relres < 1e-8
        out[trsv] = (s.getNumIter(), host(x))
    print(f"[block] {kind}: levels {out[1][0]} it, block {out[3][0]} it, pipelined {out[4][0]} it", flush=True)
    assert out[1][0] == out[3][0] == out[4][0] == out[0][0]
    for t in (3, 4, 0):
        assert np.array_equal(out[1][1], out[t][1]), t
    # mixed precision runs the FP32 one-workgroup sweeps
    for trsv in (3, 4):
        P = make_prec(pp, A, kind, args, trsv)
        if kind == "ilut":
            s = pp.GMRES(n, P, 1e-8, pp.FlexibleGMRES, maxIter=3000, precision="mixed")
            s.setRestart(True, 30)
        else:
            s = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=2000, precision="mixed")
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        s.solve(M, dev(torch, b), x)
        assert s.getConvergence() and s.report.true_relres < 1e-8, trsv



def block_diag(*mats):
    rps, cis, vs, off, nnz = [np.zeros(1, np.uint32)], [], [], 0, 0
    for rp, ci, v in mats:
        rps.append(rp[1:].astype(np.uint32) + nnz)
        cis.append(ci.astype(np.uint32) + off)
        vs.append(v)
        off += len(rp) - 1
        nnz += int(rp[-1])
    return np.concatenate(rps), np.concatenate(cis), np.concatenate(vs)


@pytest.mark.parametrize("kind,args", [("ict", (1e-3,)), ("ildlt", (1e-3,)), ("ilut", (1e-3, 10))])
def test_block_schedule_chunks(pp, torch, kind, args):
    # independent diagonal blocks of 9000, 9000 and 5000 rows: the chunked
    # schedule runs them in separate workgroups (chunks of at most the LDS
    # size); applies bit-identical to the level schedule
    mk = ppo.laplacian if kind != "ilut" else (lambda a, b, c: ppo.upwind(a, b, c, 20.0, 10.0, 5.0))
    A = block_diag(mk(90, 100, 1), mk(100, 90, 1), mk(50, 100, 1))
    n = len(A[0]) - 1
    print(f"[chunks] {kind} n={n}", flush=True)
    P1 = make_prec(pp, A, kind, args, 1)
    x = ppo.uniform(n, 8)
    t1, f1 = dev(torch, x), dev(torch, x, torch.float32)
    P1.MInvSolve(t1)
    P1.MInvSolve(f1)
    for trsv in (3, 4):
        P = make_prec(pp, A, kind, args, trsv)
        assert P.info().trsv == trsv
        t, f = dev(torch, x), dev(torch, x, torch.float32)
        P.MInvSolve(t)
        P.MInvSolve(f)
        assert np.array_equal(host(t1), host(t)), trsv
        assert np.array_equal(host(f1), host(f)), trsv
    # one connected block larger than the LDS holds: PP_TRSV_BLOCK / PIPE fall back
    B = ppo.laplacian(160, 160, 1)
    for trsv in (3, 4):
        PB = make_prec(pp, B, kind if kind != "ilut" else "ict", (1e-3,), trsv)
        assert PB.info().trsv == 1


@pytest.mark.parametrize("case", ["wide_levels", "many_rows", "long_rows", "two_colour"])
def test_pipe_sweep_shapes(pp, torch, case):
    # the pipelined one-workgroup sweep (PP_TRSV_PIPE) on the shapes its
    # paths serve: levels of more than 64 rows staged in the ring (200
    # independent 1D blocks), levels too large for the ring read straight
    # from memory (400 blocks: 400-row levels), rows longer than the
    # registers hold (ICT(1e-5) fill), and the two levels of a two-colour
    # IC(0); every apply bit-identical to the level schedule, FP64 and FP32
    if case in ("wide_levels", "many_rows"):
        nb = 200 if case == "wide_levels" else 400
        A = block_diag(*[ppo.laplacian(20, 1, 1) for _ in range(nb)])
        kind, args = "ict", (1e-3,)
    elif case == "long_rows":
        A = ppo.laplacian(48, 48, 1)
        kind, args = "ict", (1e-5,)
    else:
        A = ppo.laplacian(40, 40, 1)
        kind, args = "ic0", ()
    n = len(A[0]) - 1
    print(f"[pipe] {case} n={n}", flush=True)
    if case == "two_colour":
        P1, P4 = (pp.IC0(pp.CSRMatrix(*A), ordering=1, trsv=t) for t in (1, 4))
        P1.setup()
        P4.setup()
    else:
        P1 = make_prec(pp, A, kind, args, 1)
        P4 = make_prec(pp, A, kind, args, 4)
    assert P4.info().trsv == 4
    for seed in (3, 4):
        x = ppo.uniform(n, seed)
        for dt in (torch.float64, torch.float32):
            a, b = dev(torch, x, dt), dev(torch, x, dt)
            P1.MInvSolve(a)
            P4.MInvSolve(b)
            assert np.array_equal(host(a), host(b)), (seed, dt)
