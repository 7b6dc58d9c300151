"""Builds prispack_amd/libprispack_amd.so (HIP for gfx950) in-tree.

    python prispack_amd/build.py        (or __graft_entry__.build())

Every source under csrc/ is compiled by hipcc with --offload-arch=gfx950; the
objects go to _build/ at the repository root and the shared library next to
this file, so it travels with the tree.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libprispack_amd.so")
BUILD = os.path.join(ROOT, "_build")

ARCH = os.environ.get("PP_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
          "-I" + os.path.join(ROOT, "include")]
# packed FP32 instructions are disabled for all device code on this hardware
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc()] + COMMON + DEVICE + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = min(8, len(srcs), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lamdhip64"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources", flush=True)
    build_dropin_driver(verbose)
    return OUT


DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "cpp", "dropin_driver")


def build_dropin_driver(verbose=True):
    """The C++ drop-in test driver: host code only (g++), over
    include/prispack_amd.hpp, linked against the library built above."""
    if not os.path.exists(DRIVER_SRC):
        return None
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
           DRIVER_SRC, "-L" + HERE, "-lprispack_amd", "-Wl,-rpath,$ORIGIN/../../prispack_amd",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", DRIVER + ".tmp"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print(f"built {DRIVER}", flush=True)
    return DRIVER


if __name__ == "__main__":
    build()
    sys.exit(0)
