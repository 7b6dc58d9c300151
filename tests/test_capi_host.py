"""The C ABI without a GPU: the library loads, exports every entry point that
include/prispack_amd.h declares, and its host-side parts (model generators,
seeded inputs, the setup() factorizations) agree bit for bit with the oracle.
CPU only."""
import os
import numpy as np
import pytest

from oracle import ppo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def pp():
    import prispack_amd
    return prispack_amd


def test_library_exports_every_declared_symbol(pp):
    from prispack_amd import _capi
    syms = _capi.header_symbols()
    assert len(syms) > 40
    L = _capi.lib()
    assert [s for s in syms if not hasattr(L, s)] == []
    # every declared symbol has a ctypes signature in the binding
    assert set(syms) <= set(_capi.SIGS)


def eq(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("args", [(8, 8, 1, 0, 0, 0, 0), (6, 5, 4, 0.2, 0.1, 0.05, 0.3), (1, 9, 1, 0, 0, 0, 0),
                                  (128, 128, 1, 0, 0, 0, 0)])
def test_laplacian_generator(pp, args):
    assert eq(pp.generateLaplacianCSR(*args), ppo.laplacian(*args))


def test_other_generators_and_inputs(pp):
    assert eq(pp.generate_upwind(9, 8, 7, 20.0, 10.0, 5.0), ppo.upwind(9, 8, 7, 20.0, 10.0, 5.0))
    k = np.exp(ppo.normal(49, 1))
    assert eq(pp.generate_vardiff2d(7, k), ppo.vardiff2d(7, k))
    assert np.array_equal(pp.uniform(5000, 42), ppo.uniform(5000, 42))
    # Box-Muller goes through libm cos/sin/log, which hipcc and gcc may
    # evaluate with a one-ulp difference; inputs are generated once and handed
    # to both sides, so one ulp is the bar here.
    a, b = pp.normal(5001, 42), ppo.normal(5001, 42)
    assert np.all(np.abs(a - b) <= 2.3e-16 * np.maximum(1.0, np.abs(b)))


@pytest.mark.parametrize("dims", [(16, 16, 1), (12, 12, 12), (6, 5, 4)])
@pytest.mark.parametrize("kind,tau", [("ict", 0.0), ("ict", 1e-3), ("ict", 1e-2), ("ildlt", 1e-3), ("ic0", None)])
def test_symmetric_factorizations(pp, dims, kind, tau):
    A = ppo.laplacian(*dims)
    args = () if tau is None else (tau,)
    f = pp.host_factor(A, kind, *args)
    P = ppo.Prec(A, kind, *args)
    g = P.factors()
    assert f["status"] == 0
    assert eq(f["lt"], g["lt"])
    assert f["mnnz"] == P.mnnz
    if kind == "ildlt":
        assert np.array_equal(f["d"], g["d"])


@pytest.mark.parametrize("A", [ppo.upwind(10, 9, 8, 20.0, 10.0, 5.0), ppo.laplacian(16, 16, 1, 0.3, 0.2, 0, 0)])
@pytest.mark.parametrize("kind,tau,lfil", [("ilut", 1e-2, 20), ("ilut", 0.0, 3), ("ilut", 1e-3, 5), ("ilu0", 0, 0)])
def test_ilu_factorizations(pp, A, kind, tau, lfil):
    f = pp.host_factor(A, kind, tau, lfil) if kind == "ilut" else pp.host_factor(A, kind)
    P = ppo.Prec(A, kind, tau, lfil) if kind == "ilut" else ppo.Prec(A, kind)
    g = P.factors()
    assert eq(f["l"], g["l"]) and eq(f["u"], g["u"])
    assert np.array_equal(f["d"], g["d"]) and f["mnnz"] == P.mnnz


def test_factorization_failures(pp):
    from prispack_amd._capi import PP_ERR_NOT_SPD, PP_ERR_ZERO_ROW
    A = ppo.laplacian(5, 5, 1, 0, 0, 0, 6.0)
    f = pp.host_factor(A, "ict", 1e-3)
    assert f["status"] == PP_ERR_NOT_SPD and f["fail_row"] == ppo.Prec(A, "ict", 1e-3).fail_row
    assert pp.host_factor(A, "ic0")["status"] == PP_ERR_NOT_SPD
    Z = (np.array([0, 1, 1, 2], np.uint32), np.array([0, 2], np.uint32), np.array([1.0, 1.0]))
    assert pp.host_factor(Z, "ilut", 0.0, 5)["status"] == PP_ERR_ZERO_ROW


def shuffled_rows(A, seed):
    # the same matrix with the entries of every row in a random order
    rp, ci, v = A
    rng = np.random.default_rng(seed)
    ci2, v2 = ci.copy(), v.copy()
    for i in range(len(rp) - 1):
        s = rng.permutation(int(rp[i + 1] - rp[i])) + rp[i]
        ci2[rp[i]:rp[i + 1]] = ci[s]
        v2[rp[i]:rp[i + 1]] = v[s]
    return rp, ci2, v2


def upper_extra(A, shift):
    # an entry (i, i + shift) that has no mirror (i + shift, i): a pattern
    # that is not symmetric; IC(0) reads only the lower part
    rp, ci, v = A
    n = len(rp) - 1
    rows = []
    for i in range(n):
        r = list(zip(ci[rp[i]:rp[i + 1]].tolist(), v[rp[i]:rp[i + 1]].tolist()))
        if i + shift < n and all(c != i + shift for c, _ in r):
            r.append((i + shift, -0.01))
        rows.append(sorted(r))
    rp2 = np.zeros(n + 1, np.uint32)
    rp2[1:] = np.cumsum([len(r) for r in rows])
    return (rp2, np.array([c for r in rows for c, _ in r], np.uint32), np.array([x for r in rows for _, x in r]))


@pytest.mark.parametrize("case", ["sorted", "shuffled", "nonsymmetric"])
@pytest.mark.parametrize("dims", [(9, 8, 7), (82, 81, 80)])
def test_ic0_row_layouts(pp, case, dims):
    # sorted rows with a symmetric pattern take the level-parallel merge and
    # the scatter-free L^T; shuffled rows the serial slot table; a pattern
    # that is not symmetric the transpose (parallel beyond 2^20 entries):
    # all bit for bit the oracle's
    A = ppo.laplacian(*dims)
    if case == "shuffled":
        A = shuffled_rows(A, 3)
    elif case == "nonsymmetric":
        A = upper_extra(A, 2)
    f = pp.host_factor(A, "ic0")
    g = ppo.Prec(A, "ic0").factors()
    assert f["status"] == 0 and eq(f["lt"], g["lt"])


@pytest.mark.parametrize("clique", [0, 70])
def test_multicolour_order_matches_greedy(pp, clique):
    # a 70-row clique needs more colours than one 64-bit mask holds: the
    # adjacency-list colouring; otherwise the one-pass masked colouring
    from mc_order import greedy_order, permute
    A = ppo.laplacian(11, 10, 9)
    if clique:
        rp, ci, v = A
        n = len(rp) - 1
        rows = []
        for i in range(n):
            r = dict(zip(ci[rp[i]:rp[i + 1]].tolist(), v[rp[i]:rp[i + 1]].tolist()))
            if i < clique:
                for j in range(clique):
                    if j != i:
                        r.setdefault(j, -1e-3)
                r[i] += clique * 1e-3
            rows.append(sorted(r.items()))
        A = (np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint32),
             np.array([c for r in rows for c, _ in r], np.uint32), np.array([x for r in rows for _, x in r]))
    order, ncol = greedy_order(A)
    assert (ncol > 64) == (clique > 64)
    f = pp.host_factor(A, "ic0", ordering=1)
    g = ppo.Prec(permute(A, order), "ic0").factors()
    assert f["status"] == 0 and eq(f["lt"], g["lt"])
    assert f["levels_lower"] == ncol


def test_level_counts_of_natural_order_ic0(pp):
    # hyperplanes i+j+k = const of a natural-order 7-point IC(0): 3(n-1)+1 levels
    f = pp.host_factor(ppo.laplacian(9, 8, 7), "ic0")
    assert f["levels_lower"] == f["levels_upper"] == 8 + 7 + 6 + 1


# ---------------------------------------------------------------- C++ drop-in layer
def test_cpp_header_compiles_strict():
    import subprocess
    src = "#include \"prispack_amd.hpp\"\nint main() { return 0; }\n"
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Wpedantic", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), "-x", "c++", "-"], input=src, text=True,
                       capture_output=True)
    assert p.returncode == 0, p.stderr


def test_cpp_dropin_driver_fails_loudly_without_gpu(tmp_path):
    # no GPU in this process's view: the first device call throws prispack::Error
    # and the driver exits non-zero; there is no CPU fallback
    import dropin_io
    if not os.path.exists(dropin_io.DRIVER):
        pytest.skip("driver not built (run the build first)")
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    A = ppo.laplacian(6, 6, 1)
    r = dropin_io.run(str(tmp_path), A, np.ones(36), "cg", "ict", 1e-2, 1e-8)
    assert r["returncode"] == 1 and r.get("status", 0) not in (0, None) and "exception" in r
