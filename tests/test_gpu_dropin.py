"""The C++ drop-in classes (include/prispack_amd.hpp) driven the way
PrisPack's own drivers drive its host classes: build A, make a
preconditioner, setup(), BaseSolver::solve(A, b, x) on host vectors, read
getNumIter/getError/getConvergence.  The driver is a separate host program
(tests/cpp/dropin_driver.cpp) linked against the library.

Checks: the host drop-in solve and the device-resident solve agree bit for
bit; MInvSolve on host and device vectors, its offset and in->out forms and
the MSolveLower/Diagonal/Upper composition agree; an invalid call throws;
iteration counts match the oracle (CG exactly, GMRES within 2) and x matches
the oracle's to 1e-6 relative."""
import numpy as np
import pytest

import dropin_io
from oracle import ppo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import os
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(dropin_io.DRIVER):
        pytest.fail("tests/cpp/dropin_driver is not built")


CASES = [
    ("cg", "ict", 1e-2, "lap"),
    ("cg", "ildlt", 1e-3, "lap"),
    ("cg", "ic0", 0.0, "lap"),
    ("cg", "none", 0.0, "lap"),
    ("gmres", "ilut", 1e-3, "upw"),
    ("gmres", "ilu0", 0.0, "upw"),
]


@pytest.mark.parametrize("solver,pc,tau,mat", CASES)
def test_cpp_dropin(tmp_path, solver, pc, tau, mat):
    A = ppo.laplacian(24, 20, 16) if mat == "lap" else ppo.upwind(20, 18, 16, 30.0, 10.0, 5.0)
    n = len(A[0]) - 1
    b = ppo.spmv(A, ppo.uniform(n, 11))
    print(f"[dropin] {solver} {pc} tau={tau} n={n}", flush=True)
    r = dropin_io.run(str(tmp_path), A, b, solver, pc, tau, 1e-8)
    assert r["returncode"] == 0, (r.get("exception"), r["stderr"])
    print(f"[dropin] iters {r['iters']} error {r['error']:.3e}", flush=True)
    assert r["converged"] == 1 and r["converged_device"] == 1
    assert r["iters"] == r["iters_device"]
    assert np.array_equal(r["x_host"], r["x_dev"])
    assert r["invalid_threw"] == 1
    if pc != "none":
        assert r["ready_before"] == 0 and r["ready_after"] == 1
        assert r["d_host_dev"] == 0.0 and r["d_offset"] == 0.0 and r["d_inout"] == 0.0
        assert 0.0 <= r["d_parts"] <= 1e-13
        assert r["fill"] > 0.0
    prec = None if pc == "none" else ppo.Prec(A, pc, tau=tau, lfil=20)
    if solver == "cg":
        o = ppo.cg(A, b, prec, tol=1e-8, maxit=1000)
        assert r["iters"] == o["iters"]
    else:
        o = ppo.gmres(A, b, prec, tol=1e-8, restart=30)
        assert abs(r["iters"] - o["iters"]) <= 2
    assert np.linalg.norm(r["x_host"] - o["x"]) / np.linalg.norm(o["x"]) <= 1e-6
