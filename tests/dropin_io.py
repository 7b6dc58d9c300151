"""Input/output files of the C++ drop-in driver (tests/cpp/dropin_driver.cpp)."""
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "dropin_driver")


def write_input(path, A, b):
    rp, ci, v = A
    n = len(rp) - 1
    with open(path, "wb") as f:
        np.array([n, rp[-1]], np.uint32).tofile(f)
        np.asarray(rp, np.uint32).tofile(f)
        np.asarray(ci, np.uint32).tofile(f)
        np.asarray(v, np.float64).tofile(f)
        np.asarray(b, np.float64).tofile(f)


def run(tmp_path, A, b, solver, pc, tau, tol, timeout=120):
    inp, out = os.path.join(tmp_path, "in.bin"), os.path.join(tmp_path, "out.bin")
    write_input(inp, A, b)
    p = subprocess.run([DRIVER, inp, out, solver, pc, repr(float(tau)), repr(float(tol))], capture_output=True,
                       text=True, timeout=timeout)
    line = p.stdout.strip().splitlines()[-1] if p.stdout.strip() else "{}"
    res = json.loads(line)
    res["returncode"] = p.returncode
    res["stderr"] = p.stderr
    if p.returncode == 0:
        n = len(A[0]) - 1
        x = np.fromfile(out, np.float64)
        res["x_host"], res["x_dev"] = x[:n], x[n:]
    return res
