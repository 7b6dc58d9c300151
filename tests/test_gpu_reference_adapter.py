"""The drop-in at the reference's own preconditioner base: the reference's
compiled HOST::CG<T> / HOST::GMRES<T> (oracle/_ref/libppref_gpu.so, built by
oracle/Makefile from the reference's unmodified sources) driving the GPU
preconditioners of include/prispack_amd_reference.hpp, which derive from the
reference's TriangularPrecondition<T>.

Bars: with the FP64 GPU apply bit-identical to the reference's serial sweeps,
the reference's solver running on the GPU preconditioner returns the same
iteration count and the same x, bit for bit, as with its own CPU
preconditioner; on the 128^2 Poisson workload with ICT(1e-3) that is the
golden vector's 19 iterations and x within 1e-6.  Every entry of the
interface (four MInvSolve overloads, MSolveLower / Diagonal / Upper, isReady,
the nonzero count, the type) agrees bit for bit.  T = float (the FP32 factor)
converges to the FP32 tolerance.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import ppo

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libppref_gpu.so")
GOLD = os.path.join(ROOT, "tests", "golden")
KIND = {"ict": 1, "ilut": 2, "ildlt": 3}


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import prispack_amd  # noqa: F401  (the HIP library, loaded once for both)
    if not os.path.exists(LIB):
        pytest.skip("oracle/_ref/libppref_gpu.so is built only where the reference sources exist")
    L = C.CDLL(LIB)
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
    args = [C.c_int, C.c_int, C.c_uint32, u32p, u32p, f64p, f64p, f64p, C.c_double, C.c_uint32, C.c_int,
            C.c_double, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    for nm in ("refgpu_cg", "refgpu_cg_f32", "refgpu_gmres"):
        getattr(L, nm).restype = C.c_int
        getattr(L, nm).argtypes = args
    L.refgpu_interface.restype = C.c_int
    L.refgpu_interface.argtypes = [C.c_int, C.c_int, C.c_uint32, u32p, u32p, f64p, C.c_double, C.c_uint32, C.c_int,
                                   f64p, f64p, u32p]
    L.refgpu_error.restype = C.c_char_p
    return L


def arrays(A):
    return (np.ascontiguousarray(A[0], np.uint32), np.ascontiguousarray(A[1], np.uint32),
            np.ascontiguousarray(A[2], np.float64))


def solve(lib, fn, gpu, kind, A, b, tau, lfil=0, trsv=0, tol=1e-8, maxit=1000):
    rp, ci, v = arrays(A)
    n = len(rp) - 1
    x = np.zeros(n)
    conv, err = C.c_int(), C.c_double()
    it = getattr(lib, fn)(gpu, KIND[kind], n, rp, ci, v, np.ascontiguousarray(b, np.float64), x, tau, lfil, trsv,
                          tol, maxit, C.byref(conv), C.byref(err))
    assert it >= 0, lib.refgpu_error().decode()
    return it, x, bool(conv.value)


@pytest.mark.parametrize("trsv", [0, 1, 3])
def test_reference_cg_drives_gpu_ict_on_c1(lib, trsv):
    g = np.load(os.path.join(GOLD, "cg_128x128x1_tau0.001.npz"))
    dims = tuple(int(t) for t in g["dims"])
    A = ppo.laplacian(*dims)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.uniform(n, int(g["seed"])))
    it_g, x_g, conv_g = solve(lib, "refgpu_cg", 1, "ict", A, b, float(g["tau"]), trsv=trsv, tol=float(g["tol"]),
                              maxit=int(g["maxit"]))
    it_c, x_c, conv_c = solve(lib, "refgpu_cg", 0, "ict", A, b, float(g["tau"]), tol=float(g["tol"]),
                              maxit=int(g["maxit"]))
    print(f"[adapter] HOST::CG + GPU ICT (trsv {trsv}): {it_g} iterations; with the CPU ICT {it_c}; "
          f"golden {int(g['iters'])}", flush=True)
    assert conv_g and it_g == it_c == int(g["iters"])
    np.testing.assert_array_equal(x_g, x_c)
    assert np.linalg.norm(x_g - g["x"]) / np.linalg.norm(g["x"]) <= 1e-6


@pytest.mark.parametrize("kind,tau,lfil", [("ict", 1e-2, 0), ("ildlt", 1e-3, 0), ("ilut", 1e-3, 10)])
def test_reference_solvers_bit_identical_on_gpu_factors(lib, kind, tau, lfil):
    A = ppo.laplacian(14, 13, 12)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.normal(n, 3))
    it_g, x_g, conv = solve(lib, "refgpu_cg", 1, kind, A, b, tau, lfil)
    it_c, x_c, _ = solve(lib, "refgpu_cg", 0, kind, A, b, tau, lfil)
    assert conv and it_g == it_c
    np.testing.assert_array_equal(x_g, x_c)
    if kind == "ilut":  # the nonsymmetric path: HOST::GMRES(30) on an upwind matrix
        U = ppo.upwind(12, 11, 10, 20.0, 10.0, 5.0)
        nu = len(U[0]) - 1
        bu = ppo.spmv(U, ppo.uniform(nu, 4))
        it_g, x_g, conv = solve(lib, "refgpu_gmres", 1, kind, U, bu, tau, lfil, maxit=30)
        it_c, x_c, _ = solve(lib, "refgpu_gmres", 0, kind, U, bu, tau, lfil, maxit=30)
        print(f"[adapter] HOST::GMRES(30) + GPU ILUT: {it_g} iterations (CPU ILUT {it_c})", flush=True)
        assert conv and it_g == it_c
        np.testing.assert_array_equal(x_g, x_c)


@pytest.mark.parametrize("kind,tau,lfil", [("ict", 1e-3, 0), ("ildlt", 1e-3, 0), ("ilut", 1e-3, 10)])
def test_every_interface_entry_matches_the_reference(lib, kind, tau, lfil):
    A = ppo.laplacian(9, 8, 7) if kind != "ilut" else ppo.upwind(9, 8, 7, 20.0, 10.0, 5.0)
    rp, ci, v = arrays(A)
    n = len(rp) - 1
    x = ppo.uniform(n, 8)
    outs, infos = [], []
    for gpu in (1, 0):
        out = np.zeros(7 * n)
        info = np.zeros(3, np.uint32)
        st = lib.refgpu_interface(gpu, KIND[kind], n, rp, ci, v, tau, lfil, 0, x, out, info)
        assert st == 0, lib.refgpu_error().decode()
        outs.append(out.reshape(7, n))
        infos.append(info)
    np.testing.assert_array_equal(infos[0], infos[1])
    assert infos[0][0] == 1
    for k in range(7):
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=f"entry {k}")


def test_reference_cg_float_on_gpu_fp32_factor(lib):
    # HOST::CG<float> with the GPU factor applied in FP32 against the
    # reference's IncompleteCholesky<float, float>: both reach the FP32 tolerance
    A = ppo.laplacian(64, 64, 1)
    n = len(A[0]) - 1
    xt = ppo.uniform(n, 0)
    b = ppo.spmv(A, xt)
    it_g, x_g, conv_g = solve(lib, "refgpu_cg_f32", 1, "ict", A, b, 1e-3, tol=1e-5)
    it_c, x_c, conv_c = solve(lib, "refgpu_cg_f32", 0, "ict", A, b, 1e-3, tol=1e-5)
    print(f"[adapter] HOST::CG<float>: GPU FP32 ICT {it_g} iterations, CPU <float,float> ICT {it_c}", flush=True)
    # both stop at an FP32 relative residual of 1e-5; with the 64^2 Poisson
    # condition number (about 1.7e3) the solutions agree to about 1e-2
    assert conv_g and conv_c and abs(it_g - it_c) <= 2
    assert np.linalg.norm(x_g - xt) / np.linalg.norm(xt) <= 1e-2
    assert np.linalg.norm(x_g - x_c) / np.linalg.norm(x_c) <= 1e-2
