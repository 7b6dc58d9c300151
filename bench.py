"""Benchmark: time-to-rtol 1e-8 of IC(0)-preconditioned CG on the 3D 7-point
Poisson problem on a 128^3 grid (n = 2,097,152, nnz = 14,581,760, FP64 CSR),
b = A x_true with x_true ~ N(0, 1) from SplitMix64(0), x0 = 0.  Mixed
precision (FP32 factor, sweeps and inner CG; FP64 residual and refinement)
against the plain FP64 solve, plus the HBM throughput of the SpMV and of the
triangular sweeps measured with HIP events on the library's stream.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

One "step" = one complete solve from x0 = 0 to relative residual 1e-8 with
A, the factor and b resident in HBM.  Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GRID = 128


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=GRID)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-natural", action="store_true", help="skip the natural-ordering record")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary records (128^2 plain FP64, 100^3 upwind GMRES(30), 64^3 AMSED/GMSLR, "
                         "256-system batch)")
    ap.add_argument("--record", default="",
                    help="only run these secondary records (comma-separated: plain_fp64_2d_128, upwind_gmres_100, "
                         "amsed_64, gmslr_64, batch_256), print them as one JSON line and exit")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--trsv", type=int, default=0, help="SpTRSV schedule: 0 auto, 1 levels, 2 sync-free")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized FP64 buffer 5 times (elementwise copy kernel) and exit: "
                         "the known-size reference for scaling FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only build the system and time the SpMV and preconditioner-apply launches (no solves)")
    ap.add_argument("--launch", type=int, default=None,
                    help="iteration bodies: the library's default (stream launches) unless 0 hipGraph replay, "
                         "1 stream launches")
    ap.add_argument("--check-every", type=int, default=4, help="iterations per graph body (per host convergence poll)")
    ap.add_argument("--spmv-variant", type=int, default=-1,
                    help="pin the SpMV load variant of the timed launches (0-7 CSR-stream, 8 sliced-column; "
                         "default: the library's timed choice)")
    ap.add_argument("--ordering", type=int, default=1,
                    help="IC(0) row ordering: 1 multicolour (2 sweep levels, default), 0 natural (382 levels)")
    return ap.parse_args()


# HBM bytes per launch of the FP64 SpMV read from two rocprofv3 --pmc passes
# (FETCH_SIZE, WRITE_SIZE, KiB per dispatch) over this script, each scaled by
# what the same counter reads for the streaming copy of a known size
# (--copy-probe).  Keyed by (grid side, SpMV storage variant); the commands and
# the raw per-dispatch figures are in profiles/spmv_ell2_poisson128_pmc.md.
PMC_TRAFFIC = {(128, v): b for v, b in enumerate((224627870, 226711825, 194751552, 197276754, 219112328, 229056082,
                                                        189937730, 192875293, 154977738))}


ELL_VARIANT = 8  # the SpMV variant that runs the sliced-column form


def measured_traffic(grid, variant):
    return PMC_TRAFFIC.get((grid, variant))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_time_ms(torch, stream_handle, fn, reps):
    """Average duration of fn() (which launches on the library's stream),
    measured with HIP events recorded on that same stream."""
    ext = torch.cuda.current_stream() if stream_handle is None else torch.cuda.ExternalStream(stream_handle)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    e0.record(ext)
    for _ in range(reps):
        fn()
    e1.record(ext)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def ref_solve(variant, A, b, tau):
    """One reference FlexibleCG + IncompleteCholesky(tau) solve: (setup s, solve s, result)."""
    from oracle import ppo
    R = ppo.Ref(variant)
    hA = R.matrix(A)
    t0 = time.perf_counter()
    hp = R.prec(hA, "ict", tau)
    t1 = time.perf_counter()
    out = R.cg(hA, hp, b, tol=1e-8, maxit=1000)
    t2 = time.perf_counter()
    R.L.refh_prec_free(hp)
    R.free_matrix(hA)
    return t1 - t0, t2 - t1, out


def cpu_baseline(A, b, n):
    """The reference's own CPU path (FlexibleCG + IncompleteCholesky, compiled
    from its sources with its release flags and 4 OpenMP threads) on the same
    matrix and right-hand side: one complete solve; beside it the same build
    with 16 threads (the host cores a one-GPU share of the machine has)."""
    tau = 1e-2
    setup, solve, out = ref_solve("omp", A, b, tau)
    all_cores = None
    try:
        s16, t16, o16 = ref_solve("all", A, b, tau)
        all_cores = {"value": round(t16, 4), "cores": 16, "setup_s": round(s16, 4), "iters": o16["iters"],
                     "sample": "the same solve with the reference's OpenMP thread count raised from 4 to 16 at "
                               "compile time (oracle/threads_override.h); its triangular sweeps stay serial"}
    except FileNotFoundError:
        pass
    return {"value": round(solve, 4), "unit": "s", "cores": 4, "kind": "reference", "all_cores": all_cores,
            "sample": (f"one full solve of the same {n}-row system: reference FlexibleCG + IncompleteCholesky "
                       f"(tau={tau}, the reference has no IC(0)), its own 4 OpenMP threads (THREAD_NUM 4 in "
                       f"config/config.h:28 is in every num_threads clause; its triangular sweeps are serial); "
                       f"release flags; {out['iters']} iterations, setup {setup:.3f} s excluded"),
            "iters": out["iters"], "converged": out["converged"], "setup_s": round(setup, 4),
            "host_cpus": os.cpu_count()}


def oracle_errors(A, b, g, xs):
    """The FP64 oracle's solve of the same system (FlexibleCG + IC(0) on the
    red-black ordered matrix, the order the library's colouring gives a grid,
    to rtol 1e-12) and each GPU solution's relative error against it."""
    from oracle import ppo
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from mc_order import parity_order, permute_fast
    order = parity_order(g, g, g)
    Ap = permute_fast(A, order)
    t0 = time.perf_counter()
    o = ppo.cg(Ap, b[order], ppo.Prec(Ap, "ic0"), tol=1e-12, maxit=4000)
    xo = np.empty(len(b))
    xo[order] = o["x"]
    out = {"oracle": f"FlexibleCG + IC(0) on the red-black matrix to rtol 1e-12, {o['iters']} iterations, "
                     f"{time.perf_counter() - t0:.1f} s"}
    for name, x in xs.items():
        out[name] = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
    return out


def timed(torch, fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t, out


def extra_records(torch, pp, args):
    """Secondary workloads, one record each (timed after one untimed run);
    a failure is recorded as an error and never affects the headline."""
    out = {}

    def guard(name, fn):
        try:
            out[name] = fn()
        except Exception as e:  # recorded, not raised
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        log(f"[bench] {name}: {out[name]}")

    def plain_2d():
        A = pp.generateLaplacianCSR(128, 128, 1)
        n = len(A[0]) - 1
        M = pp.CSRMatrix(*A)
        xt = torch.tensor(pp.uniform(n, 0), device="cuda")
        b = torch.zeros(n, dtype=torch.float64, device="cuda")
        M.MatPVec(xt, b)
        P = pp.IncompleteCholesky(M, 1e-3)
        P.setup()
        def run():
            cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=1000)
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            cg.solve(M, b, x)
            return cg
        timed(torch, run)
        dt, cg = timed(torch, run)
        return {"workload": "ICT(1e-3)-FlexibleCG FP64, 2D Poisson 128^2, natural order, b=A x_true U(-1,1) seed 0",
                "s": round(dt, 5), "iters": cg.getNumIter(), "relres": cg.report.true_relres}

    def upwind_gmres():
        A = pp.generate_upwind(100, 100, 100, 20.0, 10.0, 5.0)
        n = len(A[0]) - 1
        M = pp.CSRMatrix(*A)
        xt = torch.tensor(pp.uniform(n, 0), device="cuda")
        b = torch.zeros(n, dtype=torch.float64, device="cuda")
        M.MatPVec(xt, b)
        P = pp.ILU0(M, make_f32=True, ordering=1)
        P.setup()
        rec = {"workload": "ILU(0) multicolour + FGMRES(30), 3D upwind 100^3, beta=(20,10,5)h, rtol 1e-8"}
        for prec in ("fp64", "mixed"):
            def run():
                g = pp.GMRES(n, P, 1e-8, pp.FlexibleGMRES, maxIter=3000, precision=prec)
                g.setRestart(True, 30)
                x = torch.zeros(n, dtype=torch.float64, device="cuda")
                g.solve(M, b, x)
                return g
            timed(torch, run)
            dt, g = timed(torch, run)
            rec[prec] = {"s": round(dt, 5), "iters": g.getNumIter(), "relres": g.report.true_relres,
                         "converged": g.getConvergence()}
        rec["speedup_mixed_vs_fp64"] = round(rec["fp64"]["s"] / rec["mixed"]["s"], 3)
        return rec

    def multilevel(kind):
        if kind == "amsed":
            A = pp.generateLaplacianCSR(64, 64, 64)
        else:
            A = pp.generate_upwind(64, 64, 64, 20.0, 10.0, 5.0)
        n = len(A[0]) - 1
        M = pp.CSRMatrix(*A)
        xt = torch.tensor(pp.uniform(n, 0), device="cuda")
        b = torch.zeros(n, dtype=torch.float64, device="cuda")
        M.MatPVec(xt, b)
        t0 = time.perf_counter()
        P = (pp.AMSED if kind == "amsed" else pp.GMSLR)(M, levels=3, ndom=16, rank=32)
        P.setup()
        setup = time.perf_counter() - t0
        sizes, ranks = P.info_levels()
        info = P.info()
        rec = {"workload": f"{kind.upper()} 3 levels, 16 subdomains, rank 32, 64^3 "
                           f"{'Poisson, FlexibleCG' if kind == 'amsed' else 'upwind (20,10,5)h, FGMRES(30)'}",
               "setup_s": round(setup, 3), "level_sizes": sizes, "ranks": ranks,
               "block_factor": {"levels": [info.levels_lower, info.levels_upper], "trsv": info.trsv}}
        for prec in ("fp64", "mixed"):
            def run():
                if kind == "amsed":
                    s = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=2000, precision=prec)
                else:
                    s = pp.GMRES(n, P, 1e-8, pp.FlexibleGMRES, maxIter=3000, precision=prec)
                    s.setRestart(True, 30)
                x = torch.zeros(n, dtype=torch.float64, device="cuda")
                s.solve(M, b, x)
                return s
            timed(torch, run)
            dt, s = timed(torch, run)
            rec[prec] = {"s": round(dt, 5), "iters": s.getNumIter(), "relres": s.report.true_relres,
                         "converged": s.getConvergence()}
        return rec

    def batch():
        nsys = 256
        sides = [32 + int(225 * u) for u in (pp.uniform(nsys, 7) + 1) / 2]
        systems = [pp.generate_vardiff2d(s, np.exp(pp.normal(s * s, 1000 + i))) for i, s in enumerate(sides)]
        t0 = time.perf_counter()
        B = pp.BatchedSystems(systems, ordering=1)
        setup = time.perf_counter() - t0
        blk = B.tune()
        b = torch.tensor(pp.normal(B.n, 5), device="cuda")
        def run():
            x = torch.zeros_like(b)
            its = B.solve(b, x, precision="mixed", tol=1e-8)
            return its
        timed(torch, run)
        dt, its = timed(torch, run)
        return {"workload": f"{nsys} 2D variable-coefficient diffusion systems, side U[32,256] (sum n={B.n}), "
                            "one persistent IC(0)-PCG launch (colour-ordered systems), mixed precision, rtol 1e-8",
                "s": round(dt, 5), "setup_s": round(setup, 3), "block": blk, "iters_max": int(its.max()),
                "iters_mean": float(its.mean()), "relres": B.report.true_relres,
                "converged": bool(B.report.converged)}

    records = {"plain_fp64_2d_128": plain_2d, "upwind_gmres_100": upwind_gmres,
               "amsed_64": lambda: multilevel("amsed"), "gmslr_64": lambda: multilevel("gmslr"),
               "batch_256": batch}
    only = args.record.split(",") if args.record else list(records)
    for name in only:
        if name not in records:
            raise SystemExit(f"unknown record {name!r}: one of {', '.join(records)}")
        guard(name, records[name])
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the solver path is single-GPU; "
              "sharding across GPUs is not implemented.", file=sys.stderr)
        sys.exit(2)
    import torch

    if args.copy_probe:
        import prispack_amd as pp
        from prispack_amd import _capi
        L = _capi.lib()
        ctx = pp.Context.default()
        nb = args.copy_probe << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):  # the streaming copy kernel: reads nb bytes, writes nb bytes
            ctx.check(L.pp_copy_stream(ctx.h, C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), nb, 0))
        ctx.sync()
        print(json.dumps({"copy_probe_bytes": nb, "launches": 5, "kernel": "k_copy_stream"}), flush=True)
        return

    import prispack_amd as pp
    from prispack_amd import _capi

    if args.record:
        print(json.dumps(extra_records(torch, pp, args)), flush=True)
        return

    g = args.grid
    A = pp.generateLaplacianCSR(g, g, g)
    n, nnz = len(A[0]) - 1, int(A[0][-1])
    log(f"[bench] 3D Poisson {g}^3 n={n} nnz={nnz}")
    xt = pp.normal(n, 0)
    M = pp.CSRMatrix(*A)
    ctx = M.ctx
    L = _capi.lib()
    d_xt = torch.tensor(xt, device="cuda")
    d_b = torch.zeros(n, dtype=torch.float64, device="cuda")
    M.MatPVec(d_xt, d_b)
    b = d_b.cpu().numpy()

    t0 = time.perf_counter()
    P = pp.IC0(M, make_f32=True, trsv=args.trsv, ordering=args.ordering)
    P.setup()
    setup_s = time.perf_counter() - t0
    info = P.info()
    log(f"[bench] IC(0) setup {setup_s:.2f} s (ordering {info.phase_ms[0]:.0f} ms, factorization "
        f"{info.phase_ms[1]:.0f} ms, analysis + upload {info.phase_ms[2]:.0f} ms, schedule {info.phase_ms[3]:.0f} ms), "
        f"levels {info.levels_lower}/{info.levels_upper}")

    def solve(precision, P=P, variant="auto"):
        cg = pp.CG(n, P, pp.FlexibleCG, tol=1e-8, maxIter=2000, precision=precision)
        cg.setCGVariant(variant)
        cg.opts.record_history = 0
        if args.launch is not None:
            cg.opts.launch = args.launch
        cg.opts.check_every = args.check_every
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        cg.solve(M, d_b, x)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        return dt, cg.report, x

    first_mixed_s = None
    reduced_rec = None
    if not args.kernels_only:
        for w in range(args.warmup):
            solve("fp64")
            dt_w, _, _ = solve("mixed")
            if w == 0:
                first_mixed_s = dt_w  # includes the colour-ordered matrix and reduced form built on first use
                ri = P.info()
                reduced_rec = {"built": {0: None, 1: "host", 2: "device", 3: "refreshed"}[ri.reduced_form],
                               "ms": round(float(ri.reduced_form_ms), 2)}
        t64, tmx, rep64, repmx = [], [], None, None
        torch.cuda.synchronize()
        for _ in range(args.steps):
            dt, rep64, x64 = solve("fp64")
            t64.append(dt)
        for _ in range(args.steps):
            dt, repmx, xmx = solve("mixed")
            tmx.append(dt)
        # the inner iteration of the mixed solve run in FP64 (the precision
        # effect alone), and the FP64 single-reduction PCG and mixed solve
        # with that inner iteration, for the record
        variant_names = {1: "reference", 2: "single_reduction", 3: "reduced"}
        inner_variant = variant_names.get(repmx.cg_variant, "reference")
        same = {}
        for name, prec, var in (("fp64_same", "fp64", inner_variant), ("fp64_sr", "fp64", "single_reduction"),
                                ("mixed_sr", "mixed", "single_reduction")):
            solve(prec, variant=var)
            ts, rp = [], None
            for _ in range(args.steps):
                dt, rp, _ = solve(prec, variant=var)
                ts.append(dt)
            same[name] = (float(np.median(ts)), rp)
        tsr, repsr = [same["fp64_sr"][0]], same["fp64_sr"][1]
        torch.cuda.synchronize()
        err_mixed = float(torch.linalg.norm(xmx - x64) / torch.linalg.norm(x64))
        err_true = float(torch.linalg.norm(xmx - d_xt) / torch.linalg.norm(d_xt))
        log(f"[bench] fp64 {np.median(t64):.4f} s ({rep64.iters} it), mixed {np.median(tmx):.4f} s "
            f"({repmx.outer_iters} outer / {repmx.inner_iters} inner, {inner_variant}); same iteration in fp64 "
            f"{same['fp64_same'][0]:.4f} s ({same['fp64_same'][1].iters} it); fp64 single-reduction "
            f"{np.median(tsr):.4f} s ({repsr.iters} it); mixed with single-reduction inner {same['mixed_sr'][0]:.4f} s")
        # new values on the same pattern (A scaled by 1.5, then 2): the
        # device re-factorization (the first call also builds its symbolic
        # maps on the host) and the next mixed solve, whose reduced form is
        # refreshed on the device; then the original values again
        v0 = np.asarray(A[2], np.float64)
        resetup = {}
        for scale in (1.5, 2.0):
            M.set_values(v0 * scale)
            torch.cuda.synchronize()
            t = time.perf_counter()
            P.refactor()
            torch.cuda.synchronize()
            t_ref = time.perf_counter() - t
            dt_s, rp_s, _ = solve("mixed")
            resetup[scale] = (t_ref, dt_s, rp_s)
        M.set_values(v0)
        P.refactor()
        resetup_rec = {"workload": "the same matrix with its values scaled (1.5, then 2): pp_precond_refactor, "
                                   "then the first mixed solve on the new values (its reduced form refreshed on "
                                   "the device)",
                       "first_refactor_s": round(resetup[1.5][0], 4),
                       "refactor_s": round(resetup[2.0][0], 4), "first_mixed_solve_s": round(resetup[2.0][1], 4),
                       "total_s": round(resetup[2.0][0] + resetup[2.0][1], 4),
                       "inner_iters": resetup[2.0][2].inner_iters,
                       "converged": bool(resetup[2.0][2].converged)}
        log(f"[bench] new values: refactor {resetup_rec['refactor_s'] * 1e3:.1f} ms (first "
            f"{resetup_rec['first_refactor_s'] * 1e3:.0f} ms, maps built) + first mixed solve "
            f"{resetup_rec['first_mixed_solve_s'] * 1e3:.1f} ms")

    # ---- kernel throughput on the library's stream
    stream = L.pp_ctx_stream(ctx.h)
    spmv_var = C.c_int()
    if args.spmv_variant >= 0:
        ctx.check(L.pp_csr_set_spmv_variant(M.h, 0, args.spmv_variant))
    ctx.check(L.pp_csr_spmv_variant(M.h, 0, C.byref(spmv_var)))
    # the FP64 SpMV alternates between two identical copies of the matrix and
    # of x, y (2 x 188 MB, more than the 256 MB last-level cache): each launch
    # streams its operands from HBM instead of finding them cached by the
    # launch before
    M2 = pp.CSRMatrix(*A, ctx=ctx)
    ctx.check(L.pp_csr_set_spmv_variant(M2.h, 0, spmv_var.value))
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    x2, y2 = d_xt.clone(), torch.zeros_like(y)
    px, py = C.c_void_p(d_xt.data_ptr()), C.c_void_p(y.data_ptr())
    px2, py2 = C.c_void_p(x2.data_ptr()), C.c_void_p(y2.data_ptr())
    flip = [0]
    def spmv_alt():
        flip[0] ^= 1
        if flip[0]:
            L.pp_spmv(ctx.h, M.h, 0, 1.0, px, 0.0, py)
        else:
            L.pp_spmv(ctx.h, M2.h, 0, 1.0, px2, 0.0, py2)
    spmv_ms = kernel_time_ms(torch, stream, spmv_alt, args.kernel_reps)
    del M2, x2, y2
    def spmv_bytes_of(dtype):
        nb = C.c_uint64()
        ctx.check(L.pp_csr_spmv_bytes(M.h, dtype, C.byref(nb)))
        return int(nb.value)
    spmv_bytes = spmv_bytes_of(0)
    x32 = d_xt.float()
    y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
    p32x, p32y = C.c_void_p(x32.data_ptr()), C.c_void_p(y32.data_ptr())
    spmv32_ms = kernel_time_ms(torch, stream, lambda: L.pp_spmv(ctx.h, M.h, 1, 1.0, p32x, 0.0, p32y),
                               args.kernel_reps)
    spmv32_bytes = spmv_bytes_of(1)
    spmv32_var = C.c_int()
    ctx.check(L.pp_csr_spmv_variant(M.h, 1, C.byref(spmv32_var)))
    # one preconditioner apply = forward sweep on L + backward sweep on L^T,
    # priced on the bytes the library's sweep storage streams (direct apply:
    # with the colour-order gather and scatter)
    def apply_bytes_of(dtype, in_solve=0):
        nb = C.c_uint64()
        ctx.check(L.pp_precond_apply_bytes(P.h, dtype, in_solve, C.byref(nb)))
        return int(nb.value)
    trsv_bytes = apply_bytes_of(0)
    trsv32_bytes = apply_bytes_of(1)
    tv = torch.tensor(xt, device="cuda")
    ptv = C.c_void_p(tv.data_ptr())
    trsv_ms = kernel_time_ms(torch, stream, lambda: L.pp_precond_apply(ctx.h, P.h, 0, ptv, ptv),
                             max(3, args.kernel_reps // 10))
    tv32 = tv.float()
    ptv32 = C.c_void_p(tv32.data_ptr())
    trsv32_ms = kernel_time_ms(torch, stream, lambda: L.pp_precond_apply(ctx.h, P.h, 1, ptv32, ptv32),
                               max(3, args.kernel_reps // 10))
    gbs = lambda by, ms: by / (ms * 1e-3) / 1e9
    # what a kernel that only moves bytes reaches on this device: the
    # library's streaming copy (16-byte non-temporal loads and stores) of
    # 1 GiB (read 1 GiB, write 1 GiB) at four launch widths, the best kept,
    # timed with HIP events on the library stream; torch's elementwise copy
    # beside it for reference
    src = torch.ones((1 << 30) // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    ps, pd = C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr())
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    copy_by_groups = {}
    for per_cu in (8, 16, 32, 64):
        ms = kernel_time_ms(torch, stream, lambda: L.pp_copy_stream(ctx.h, ps, pd, 1 << 30, cus * per_cu), 10)
        copy_by_groups[cus * per_cu] = round(gbs(2 << 30, ms), 1)
    copy_gbs = max(copy_by_groups.values())
    torch_copy_gbs = gbs(2 << 30, kernel_time_ms(torch, None, lambda: torch.mul(src, 1.0, out=dst), 10))
    del src, dst
    log(f"[bench] SpMV fp64 {spmv_ms * 1e3:.1f} us ({gbs(12 * nnz + 20 * n + 4, spmv_ms):.0f} GB/s algorithmic, "
        f"{gbs(spmv_bytes, spmv_ms):.0f} GB/s streamed); fp32 {spmv32_ms * 1e3:.1f} us; "
        f"M^-1 apply fp64 {trsv_ms * 1e3:.1f} us ({gbs(trsv_bytes, trsv_ms):.1f} GB/s streamed); "
        f"fp32 {trsv32_ms * 1e3:.1f} us ({gbs(trsv32_bytes, trsv32_ms):.1f} GB/s streamed)")

    sched = {0: "levels", 1: "levels", 2: "syncfree"}[args.trsv]
    # algorithmic bytes (32-bit indices, x read once, y written once):
    # SpMV 12 nnz + 20 n + 4 (FP64), 8 nnz + 12 n + 4 (FP32); one sweep over
    # the IC(0) factor 12 nnz_L + 4 (n + 1) + 16 n (FP64), 8 nnz_L + 4 (n + 1)
    # + 8 n (FP32), nnz_L = (nnz + n) / 2, and an apply is two sweeps
    alg_spmv = 12 * nnz + 20 * n + 4
    alg_spmv32 = 8 * nnz + 12 * n + 4
    nnz_l = (nnz + n) // 2
    alg_apply = 2 * (12 * nnz_l + 4 * (n + 1) + 16 * n)
    alg_apply32 = 2 * (8 * nnz_l + 4 * (n + 1) + 8 * n)
    traffic = measured_traffic(g, spmv_var.value)
    spmv_kernel = ("k_ell2<double, 7, true> (y = A x, FP64 sliced-column form, offsets from one pattern byte per "
                   "row, two rows per thread"
                   if spmv_var.value == ELL_VARIANT else "k_spmv<double> (y = A x, FP64 CSR")
    dominant = {"kernel": spmv_kernel + ", launches alternating between two copies of A, x, y on the library stream)",
                "bound": "hbm", "achieved": round(gbs(alg_spmv, spmv_ms), 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(gbs(alg_spmv, spmv_ms) / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes": alg_spmv, "avg_us": round(spmv_ms * 1e3, 2),
                "launches_timed": args.kernel_reps, "bytes_streamed": spmv_bytes,
                "copy_kernel_GBs": round(copy_gbs, 1), "copy_kernel": "k_copy_stream (pp_copy_stream), best of "
                f"{sorted(copy_by_groups)} workgroups: {copy_by_groups}",
                "torch_copy_GBs": round(torch_copy_gbs, 1),
                "frac_of_copy": round(gbs(alg_spmv, spmv_ms) / copy_gbs, 4),
                # the same launch priced on the HBM bytes the counters measured
                # (the sliced-column form stores one pattern byte per row, so it
                # moves fewer bytes than the 32-bit-index algorithmic figure)
                "traffic_GBs": round(gbs(traffic, spmv_ms), 1) if traffic else None,
                "traffic_frac": round(gbs(traffic, spmv_ms) / HBM_PEAK_GBS, 4) if traffic else None,
                "traffic_frac_of_copy": round(gbs(traffic, spmv_ms) / copy_gbs, 4) if traffic else None,
                "streamed_GBs": round(gbs(spmv_bytes, spmv_ms), 1), "variant": spmv_var.value}
    apply_rec = {"kernel": f"IC(0) apply = forward + backward sweep ({sched} schedule, {info.levels_lower} levels)",
                 "bound": "hbm", "achieved": round(gbs(alg_apply, trsv_ms), 1), "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": round(gbs(alg_apply, trsv_ms) / HBM_PEAK_GBS, 4),
                 "algorithmic_bytes": alg_apply, "avg_us": round(trsv_ms * 1e3, 2),
                 "bytes_streamed": trsv_bytes, "sweep_variant": info.sweep_variant,
                 "fused_levels": info.fused_levels,
                 "fp32": {"achieved": round(gbs(alg_apply32, trsv32_ms), 1), "algorithmic_bytes": alg_apply32,
                          "avg_us": round(trsv32_ms * 1e3, 2), "bytes_streamed": trsv32_bytes}}
    if args.kernels_only:
        print(json.dumps({"roofline": dominant, "roofline_apply": apply_rec}), flush=True)
        return

    # the other ordering, one solve of each precision, for the record
    other = None
    if args.ordering == 1 and not args.no_natural:
        Pn = pp.IC0(M, make_f32=True, trsv=args.trsv, ordering=0)
        Pn.setup()
        solve("fp64", Pn)
        dn64, rn64, _ = solve("fp64", Pn)
        dnmx, rnmx, _ = solve("mixed", Pn)
        other = {"ordering": "natural (382 sweep levels)", "fp64_s": round(dn64, 4), "mixed_s": round(dnmx, 4),
                 "iters_fp64": rn64.iters, "iters_mixed_inner": rnmx.inner_iters,
                 "speedup_mixed_vs_fp64": round(dn64 / dnmx, 3)}
        log(f"[bench] natural ordering: fp64 {dn64:.4f} s ({rn64.iters} it), mixed {dnmx:.4f} s")
        del Pn

    extra = None if args.no_extra else extra_records(torch, pp, args)

    cpu = None
    errs = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(A, b, n)
        except FileNotFoundError as e:
            cpu = {"value": None, "unit": "s", "cores": 0, "kind": "reference", "sample": f"unavailable: {e}"}
        if args.ordering == 1:
            errs = oracle_errors(A, b, g, {"mixed": xmx.cpu().numpy(), "fp64": x64.cpu().numpy()})
            log(f"[bench] error against the FP64 oracle: {errs}")

    med64, medmx = float(np.median(t64)), float(np.median(tmx))
    out = {
        "metric": "time-to-rtol 1e-8 (s) mixed-prec vs FP64 on GPU; SpMV/SpTRSV HBM GB/s",
        "value": round(medmx, 5), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(medmx * 1e3, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-inner/f64-outer", "data": "synthetic",
        "config": {"workload": f"IC(0)-PCG, 3D 7-point Poisson {g}^3 (n={n}, nnz={nnz}), b=A x_true, "
                               "x_true~N(0,1) seed 0, rtol 1e-8", "precision": "mixed vs fp64",
                   "ordering": "multicolour" if args.ordering == 1 else "natural"},
        "other_ordering": other,
        "fp64_s": round(med64, 5), "speedup_mixed_vs_fp64": round(med64 / medmx, 3),
        "mixed_inner_iteration": inner_variant,
        "fp64_same_algorithm_s": round(same["fp64_same"][0], 5), "iters_fp64_same_algorithm": same["fp64_same"][1].iters,
        "speedup_mixed_vs_fp64_same_algorithm": round(same["fp64_same"][0] / medmx, 3),
        "fp64_single_reduction_s": round(float(np.median(tsr)), 5), "iters_fp64_single_reduction": repsr.iters,
        "mixed_single_reduction_inner_s": round(same["mixed_sr"][0], 5),
        "iters_mixed_single_reduction_inner": same["mixed_sr"][1].inner_iters,
        "iters_fp64": rep64.iters, "iters_mixed_inner": repmx.inner_iters, "outer_mixed": repmx.outer_iters,
        "relres_fp64": rep64.true_relres, "relres_mixed": repmx.true_relres,
        "err_mixed_vs_fp64": err_mixed, "err_vs_xtrue": err_true,
        "err_vs_fp64_oracle": errs, "setup_s": round(setup_s, 3),
        "levels": [info.levels_lower, info.levels_upper],
        "setup_phases_ms": [round(float(v), 1) for v in info.phase_ms],
        "first_mixed_solve_s": None if first_mixed_s is None else round(first_mixed_s, 4),
        "reduced_form": reduced_rec,
        "new_values": resetup_rec,
        "roofline": dominant,
        "roofline_apply": apply_rec,
        "roofline_spmv_f32": {"kernel": ("k_ell2<float> (sliced-column, row patterns)" if spmv32_var.value == ELL_VARIANT
                                         else "k_spmv<float>"),
                              "variant": spmv32_var.value, "achieved": round(gbs(alg_spmv32, spmv32_ms), 1),
                              "unit": "GB/s", "frac": round(gbs(alg_spmv32, spmv32_ms) / HBM_PEAK_GBS, 4),
                              "algorithmic_bytes": alg_spmv32, "avg_us": round(spmv32_ms * 1e3, 2),
                              "bytes_streamed": spmv32_bytes,
                              "streamed_GBs": round(gbs(spmv32_bytes, spmv32_ms), 1),
                              "streamed_frac": round(gbs(spmv32_bytes, spmv32_ms) / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "secondary": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
