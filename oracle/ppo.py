"""ctypes bindings for the oracle (libppo.so) and the reference shim (_ref/).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The prispack_amd package never imports it.

A CSR matrix is passed around as a tuple (rp, ci, v) of numpy arrays
(uint32, uint32, float64), 0-based, rows sorted by column.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class _Csr(C.Structure):
    _fields_ = [("nrow", C.c_uint32), ("ncol", C.c_uint32), ("nnz", C.c_uint32),
                ("rp", C.POINTER(C.c_uint32)), ("ci", C.POINTER(C.c_uint32)),
                ("v", C.POINTER(C.c_double))]


class _Prec(C.Structure):
    _fields_ = [("kind", C.c_int), ("status", C.c_int), ("fail_row", C.c_uint32),
                ("lt", C.POINTER(_Csr)), ("l", C.POINTER(_Csr)), ("u", C.POINTER(_Csr)),
                ("d", C.POINTER(C.c_double)), ("mnnz", C.c_uint32)]


class _Report(C.Structure):
    _fields_ = [("iters", C.c_int32), ("converged", C.c_int32),
                ("final_error", C.c_double), ("nhist", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libppo.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = C.CDLL(path)
        P = C.POINTER(_Csr)
        Q = C.POINTER(_Prec)
        L.ppo_csr_from.restype = P
        L.ppo_csr_from.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p, _f64p]
        L.ppo_csr_free.argtypes = [P]
        L.ppo_csr_transpose.restype = P
        L.ppo_csr_transpose.argtypes = [P]
        L.ppo_fill_uniform.argtypes = [_f64p, C.c_uint64, C.c_uint64]
        L.ppo_fill_normal.argtypes = [_f64p, C.c_uint64, C.c_uint64]
        L.ppo_laplacian.restype = P
        L.ppo_laplacian.argtypes = [C.c_int] * 3 + [C.c_double] * 4
        L.ppo_upwind.restype = P
        L.ppo_upwind.argtypes = [C.c_int] * 3 + [C.c_double] * 3
        L.ppo_vardiff2d.restype = P
        L.ppo_vardiff2d.argtypes = [C.c_int, _f64p]
        L.ppo_spmv.argtypes = [P, _f64p, _f64p]
        L.ppo_spmv_t_acc.argtypes = [P, _f64p, _f64p]
        for name in ("ppo_ic0", "ppo_ilu0"):
            getattr(L, name).restype = Q
            getattr(L, name).argtypes = [P]
        L.ppo_ict.restype = Q
        L.ppo_ict.argtypes = [P, C.c_double]
        L.ppo_ildlt.restype = Q
        L.ppo_ildlt.argtypes = [P, C.c_double]
        L.ppo_ilut.restype = Q
        L.ppo_ilut.argtypes = [P, C.c_uint32, C.c_double]
        L.ppo_prec_free.argtypes = [Q]
        for name in ("ppo_prec_apply", "ppo_prec_lower", "ppo_prec_diag", "ppo_prec_upper"):
            getattr(L, name).argtypes = [Q, _f64p]
        L.ppo_cg.argtypes = [P, Q, C.c_int, _f64p, _f64p, C.c_double, C.c_uint32,
                             _f64p, C.c_uint32, C.POINTER(_Report)]
        L.ppo_gmres.argtypes = [P, Q, C.c_int, _f64p, _f64p, C.c_double, C.c_uint32,
                                C.c_uint32, _f64p, C.c_uint32, C.POINTER(_Report)]
        _lib = L
    return _lib


def _to_np(p):
    a = p.contents
    n, nnz = a.nrow, a.nnz
    rp = np.ctypeslib.as_array(a.rp, shape=(n + 1,)).copy()
    ci = np.ctypeslib.as_array(a.ci, shape=(max(nnz, 1),))[:nnz].copy()
    v = np.ctypeslib.as_array(a.v, shape=(max(nnz, 1),))[:nnz].copy()
    return rp, ci, v


def _own(p):
    out = _to_np(p)
    lib().ppo_csr_free(p)
    return out


class Csr:
    """Owns a ppo_csr built from numpy arrays for the duration of a call."""

    def __init__(self, A):
        rp, ci, v = (np.ascontiguousarray(A[0], np.uint32), np.ascontiguousarray(A[1], np.uint32),
                     np.ascontiguousarray(A[2], np.float64))
        n = len(rp) - 1
        if len(ci) == 0:
            ci = np.zeros(1, np.uint32)
            v = np.zeros(1, np.float64)
        self.n = n
        self.p = lib().ppo_csr_from(n, n, rp, ci, v)

    def __del__(self):
        if getattr(self, "p", None) is not None:
            lib().ppo_csr_free(self.p)
            self.p = None


# ---------------- inputs ----------------
def uniform(n, seed):
    x = np.empty(n, np.float64)
    lib().ppo_fill_uniform(x, n, seed)
    return x


def normal(n, seed):
    x = np.empty(max(n, 1), np.float64)
    lib().ppo_fill_normal(x, n, seed)
    return x[:n]


def laplacian(nx, ny, nz, ax=0.0, ay=0.0, az=0.0, shift=0.0):
    return _own(lib().ppo_laplacian(nx, ny, nz, ax, ay, az, shift))


def upwind(nx, ny, nz, bx, by, bz):
    return _own(lib().ppo_upwind(nx, ny, nz, bx, by, bz))


def vardiff2d(side, kappa):
    return _own(lib().ppo_vardiff2d(side, np.ascontiguousarray(kappa, np.float64)))


def transpose(A):
    c = Csr(A)
    return _own(lib().ppo_csr_transpose(c.p))


# ---------------- kernels ----------------
def spmv(A, x):
    c = Csr(A)
    y = np.empty(c.n, np.float64)
    lib().ppo_spmv(c.p, np.ascontiguousarray(x, np.float64), y)
    return y


def spmv_t(A, x, y0=None):
    c = Csr(A)
    y = np.zeros(c.n, np.float64) if y0 is None else np.array(y0, np.float64)
    lib().ppo_spmv_t_acc(c.p, np.ascontiguousarray(x, np.float64), y)
    return y


# ---------------- preconditioners ----------------
class Prec:
    """Oracle preconditioner; kind in {'ict','ilut','ildlt','ic0','ilu0'}."""

    def __init__(self, A, kind, tau=1e-3, lfil=20):
        self.A = Csr(A)
        L = lib()
        if kind == "ict":
            self.p = L.ppo_ict(self.A.p, tau)
        elif kind == "ildlt":
            self.p = L.ppo_ildlt(self.A.p, tau)
        elif kind == "ilut":
            self.p = L.ppo_ilut(self.A.p, lfil, tau)
        elif kind == "ic0":
            self.p = L.ppo_ic0(self.A.p)
        elif kind == "ilu0":
            self.p = L.ppo_ilu0(self.A.p)
        else:
            raise ValueError(kind)
        self.kind = kind
        pc = self.p.contents
        self.status = pc.status
        self.fail_row = pc.fail_row
        self.mnnz = pc.mnnz
        self.n = self.A.n

    def factors(self):
        pc = self.p.contents
        out = {}
        if pc.lt:
            out["lt"] = _to_np(pc.lt)
        if pc.l:
            out["l"] = _to_np(pc.l)
        if pc.u:
            out["u"] = _to_np(pc.u)
        if pc.d:
            out["d"] = np.ctypeslib.as_array(pc.d, shape=(self.n,)).copy()
        return out

    def _run(self, fn, x):
        y = np.array(x, np.float64, copy=True)
        fn(self.p, y)
        return y

    def apply(self, x):
        return self._run(lib().ppo_prec_apply, x)

    def lower(self, x):
        return self._run(lib().ppo_prec_lower, x)

    def diag(self, x):
        return self._run(lib().ppo_prec_diag, x)

    def upper(self, x):
        return self._run(lib().ppo_prec_upper, x)

    def __del__(self):
        if getattr(self, "p", None) is not None:
            lib().ppo_prec_free(self.p)
            self.p = None


# ---------------- solvers ----------------
def cg(A, b, prec=None, x0=None, tol=1e-8, maxit=1000, flexible=True):
    c = Csr(A)
    x = np.zeros(c.n) if x0 is None else np.array(x0, np.float64)
    hist = np.zeros(maxit + 1)
    rep = _Report()
    lib().ppo_cg(c.p, prec.p if prec else None, int(flexible), np.ascontiguousarray(b, np.float64),
                 x, tol, maxit, hist, len(hist), C.byref(rep))
    return dict(x=x, iters=rep.iters, converged=bool(rep.converged),
                error=rep.final_error, hist=hist[:min(rep.nhist, len(hist))])


def gmres(A, b, prec=None, x0=None, tol=1e-8, restart=30, maxit_total=500, flexible=True):
    c = Csr(A)
    x = np.zeros(c.n) if x0 is None else np.array(x0, np.float64)
    hist = np.zeros((maxit_total // restart + 2) * (restart + 1))
    rep = _Report()
    lib().ppo_gmres(c.p, prec.p if prec else None, int(flexible), np.ascontiguousarray(b, np.float64),
                    x, tol, restart, maxit_total, hist, len(hist), C.byref(rep))
    return dict(x=x, iters=rep.iters, converged=bool(rep.converged),
                error=rep.final_error, hist=hist[:min(rep.nhist, len(hist))])


# ---------------- the compiled reference (oracle/_ref) ----------------
class Ref:
    """The reference's own C++ hot path, compiled from its sources.

    variant 'strict' = serial, strict IEEE (golden vectors);
    variant 'omp'    = the reference release flags + 4 OpenMP threads (timing);
    variant 'all'    = the same with 16 OpenMP threads (threads_override.h).
    Only present where oracle/_ref was built.
    """

    def __init__(self, variant="strict"):
        name = {"strict": "libppref.so", "omp": "libppref_omp.so", "all": "libppref_all.so"}[variant]
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        vp = C.c_void_p
        L.refh_csr.restype = vp
        L.refh_csr.argtypes = [C.c_uint32, _u32p, _u32p, _f64p]
        L.refh_laplacian.restype = vp
        L.refh_laplacian.argtypes = [C.c_int] * 3 + [C.c_double] * 4
        L.refh_csr_free.argtypes = [vp]
        L.refh_csr_n.restype = C.c_uint32
        L.refh_csr_n.argtypes = [vp]
        L.refh_csr_nnz.restype = C.c_uint32
        L.refh_csr_nnz.argtypes = [vp]
        L.refh_csr_get.argtypes = [vp, _u32p, _u32p, _f64p]
        L.refh_spmv.argtypes = [vp, _f64p, _f64p]
        L.refh_spmv_t_acc.argtypes = [vp, _f64p, _f64p]
        for nm, extra in (("refh_ict", [C.c_double]), ("refh_ildlt", [C.c_double]),
                          ("refh_ilut", [C.c_uint32, C.c_double])):
            getattr(L, nm).restype = vp
            getattr(L, nm).argtypes = [vp] + extra
        L.refh_prec_free.argtypes = [vp]
        L.refh_prec_ready.restype = C.c_int
        L.refh_prec_ready.argtypes = [vp]
        L.refh_prec_mnnz.restype = C.c_uint32
        L.refh_prec_mnnz.argtypes = [vp]
        L.refh_factor_nnz.restype = C.c_uint32
        L.refh_factor_nnz.argtypes = [vp, C.c_int]
        L.refh_factor_get.argtypes = [vp, C.c_int, _u32p, _u32p, _f64p]
        L.refh_ilu_diag.argtypes = [vp, _f64p, C.c_uint32]
        L.refh_prec_apply.argtypes = [vp, _f64p, C.c_uint32]
        L.refh_prec_part.argtypes = [vp, C.c_int, _f64p, C.c_uint32]
        L.refh_cg.restype = C.c_int
        L.refh_cg.argtypes = [vp, vp, C.c_int, _f64p, _f64p, C.c_double, C.c_uint32,
                              C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.refh_read_mtx.restype = vp
        L.refh_read_mtx.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.refh_rcm.argtypes = [vp, _u32p]
        L.refh_gmres.restype = C.c_int
        L.refh_gmres.argtypes = [vp, vp, C.c_int, _f64p, _f64p, C.c_double, C.c_uint32,
                                 C.POINTER(C.c_int), C.POINTER(C.c_double)]
        self.L = L

    def _csr_out(self, h):
        n, nnz = self.L.refh_csr_n(h), self.L.refh_csr_nnz(h)
        rp = np.zeros(n + 1, np.uint32)
        ci = np.zeros(max(nnz, 1), np.uint32)
        v = np.zeros(max(nnz, 1), np.float64)
        self.L.refh_csr_get(h, rp, ci, v)
        return rp, ci[:nnz], v[:nnz]

    def laplacian(self, nx, ny, nz, ax=0.0, ay=0.0, az=0.0, shift=0.0):
        h = self.L.refh_laplacian(nx, ny, nz, ax, ay, az, shift)
        out = self._csr_out(h)
        self.L.refh_csr_free(h)
        return out

    def matrix(self, A):
        rp, ci, v = A
        ci = np.ascontiguousarray(ci, np.uint32) if len(ci) else np.zeros(1, np.uint32)
        v = np.ascontiguousarray(v, np.float64) if len(v) else np.zeros(1)
        return self.L.refh_csr(len(rp) - 1, np.ascontiguousarray(rp, np.uint32), ci, v)

    def free_matrix(self, h):
        self.L.refh_csr_free(h)

    def spmv(self, A, x):
        h = self.matrix(A)
        y = np.zeros(len(A[0]) - 1)
        self.L.refh_spmv(h, np.ascontiguousarray(x, np.float64), y)
        self.L.refh_csr_free(h)
        return y

    def spmv_t(self, A, x):
        h = self.matrix(A)
        y = np.zeros(len(A[0]) - 1)
        self.L.refh_spmv_t_acc(h, np.ascontiguousarray(x, np.float64), y)
        self.L.refh_csr_free(h)
        return y

    def prec(self, hA, kind, tau=1e-3, lfil=20):
        if kind == "ict":
            return self.L.refh_ict(hA, tau)
        if kind == "ildlt":
            return self.L.refh_ildlt(hA, tau)
        if kind == "ilut":
            return self.L.refh_ilut(hA, lfil, tau)
        raise ValueError(kind)

    def factor(self, hp, which, n):
        nnz = self.L.refh_factor_nnz(hp, which)
        rp = np.zeros(n + 1, np.uint32)
        ci = np.zeros(max(nnz, 1), np.uint32)
        v = np.zeros(max(nnz, 1))
        self.L.refh_factor_get(hp, which, rp, ci, v)
        return rp, ci[:nnz], v[:nnz]

    def ilu_diag(self, hp, n):
        d = np.zeros(n)
        self.L.refh_ilu_diag(hp, d, n)
        return d

    def apply(self, hp, x, part=None):
        y = np.array(x, np.float64, copy=True)
        if part is None:
            self.L.refh_prec_apply(hp, y, len(y))
        else:
            self.L.refh_prec_part(hp, part, y, len(y))
        return y

    def cg(self, hA, hp, b, x0=None, tol=1e-8, maxit=1000, flexible=True):
        n = len(b)
        x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
        conv, err = C.c_int(), C.c_double()
        it = self.L.refh_cg(hA, hp, int(flexible), np.ascontiguousarray(b, np.float64), x, tol, maxit,
                            C.byref(conv), C.byref(err))
        return dict(x=x, iters=it, converged=bool(conv.value), error=err.value)

    def gmres(self, hA, hp, b, x0=None, tol=1e-8, restart=30, flexible=True):
        n = len(b)
        x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
        conv, err = C.c_int(), C.c_double()
        it = self.L.refh_gmres(hA, hp, int(flexible), np.ascontiguousarray(b, np.float64), x, tol,
                               restart, C.byref(conv), C.byref(err))
        return dict(x=x, iters=it, converged=bool(conv.value), error=err.value)
