"""prispack_amd -- the PrisPack preconditioned-Krylov hot path on AMD Instinct
MI355X (gfx950), behind the reference's operator interface.

Host-side mirror of the reference classes (PrisPack, include/...):
  CSRMatrix            HOST::CSRMatrix / DEVICE::DeviceCSRMatrix
                       (include/MatrixClass/CSRMatrix.h, DeviceCSRMatrix.cuh)
  IncompleteCholesky   HOST::IncompleteCholesky<Low,High>(A, threshold)
  IncompleteLU         HOST::IncompleteLU<Low,High>(A, lfill, tolerance)
  IncompleteLDLT       HOST::IncompleteLDLT<Low,High>(A, tolerance)
  IC0, ILU0            zero-fill variants (no reference counterpart)
  CG                   HOST::CG<T>(n, precond, CGType, tol, maxIter)
  GMRES                HOST::GMRES<T>(n, precond, tol, GMRESType) + setRestart
  generateLaplacianCSR HOST::generateLaplacianCSR (ModelProblem.h)

Every numerical operation runs in libprispack_amd.so on the GPU through the C
ABI of include/prispack_amd.h; importing the package without the built library
raises ImportError (no CPU fallback).  Vectors are torch CUDA tensors
(device-resident solves) or numpy arrays (host in / host out).
"""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import (PP_F32, PP_F64, PP_FP64, PP_MIXED, PP_PC_IC0, PP_PC_ICT, PP_PC_ILDLT, PP_PC_ILU0,
                    PP_PC_ILUT, PP_PC_NONE, PP_SOLVER_CG, PP_SOLVER_GMRES, PrisPackError, check)

_capi.lib()  # fail loudly at import when the HIP library is missing

__all__ = ["Context", "CSRMatrix", "IncompleteCholesky", "IncompleteLU", "IncompleteLDLT", "IC0", "ILU0",
           "CG", "GMRES", "NormalCG", "FlexibleCG", "NormalGMRES", "FlexibleGMRES", "generateLaplacianCSR",
           "generate_upwind", "generate_vardiff2d", "uniform", "normal", "PrisPackError", "host_factor",
           "read_mtx", "read_mtx_dense", "rcm"]

# CGType / GMRESType (include/Solver/CG.h:14-17, include/Solver/GMRES.h:15-18)
NormalCG, FlexibleCG = 0, 1
NormalGMRES, FlexibleGMRES = 0, 1

# PreconditionType (include/Preconditioner/BasePreconditon.h:12-19)
PreconditionNone, PreconditionIC, PreconditionILUT, PreconditionILDLT, PreconditionGMSLR, PreconditionAMSED = range(6)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


# ----------------------------------------------------------------------------- model problems
def _gen(fn, *args):
    L = _capi.lib()
    n, nnz = C.c_uint32(), C.c_uint32()
    check(fn(*args, C.byref(n), C.byref(nnz), None, None, None))
    rp = np.zeros(n.value + 1, np.uint32)
    ci = np.zeros(max(nnz.value, 1), np.uint32)
    v = np.zeros(max(nnz.value, 1), np.float64)
    check(fn(*args, C.byref(n), C.byref(nnz), _p(rp, C.c_uint32), _p(ci, C.c_uint32), _p(v, C.c_double)))
    del L
    return rp, ci[:nnz.value], v[:nnz.value]


def generateLaplacianCSR(nx, ny, nz, alphax=0.0, alphay=0.0, alphaz=0.0, shift=0.0):
    """HOST::generateLaplacianCSR (include/MatrixClass/ModelProblem.h:45-52)."""
    return _gen(_capi.lib().pp_generate_laplacian, nx, ny, nz, alphax, alphay, alphaz, shift)


def generate_upwind(nx, ny, nz, bx, by, bz):
    """First-order upwind convection-diffusion (scaled by h^2, h = 1/(nx+1))."""
    return _gen(_capi.lib().pp_generate_upwind, nx, ny, nz, bx, by, bz)


def generate_vardiff2d(side, kappa):
    """-div(kappa grad u) on a side x side grid, harmonic-mean face coefficients."""
    k = _f64(kappa)
    return _gen(_capi.lib().pp_generate_vardiff2d, side, _p(k, C.c_double))


def _fill(kind, n, seed):
    out = np.zeros(max(n, 1), np.float64)
    check(_capi.lib().pp_fill_random(kind, seed, n, _p(out, C.c_double)))
    return out[:n]


def uniform(n, seed):
    """x_i ~ U(-1, 1) from SplitMix64(seed)."""
    return _fill(0, n, seed)


def normal(n, seed):
    """x_i ~ N(0, 1) from SplitMix64(seed) + Box-Muller."""
    return _fill(1, n, seed)


# ----------------------------------------------------------------------------- input and ordering
def read_mtx(path, base=1):
    """HOST::ReadMtxTools<double>(.., path, base).loadMatrix(csr)
    (source/utils/TestTools/ReadMtxTools.cpp): a coordinate real MatrixMarket
    file as CSR (rp, ci, v) with the reference reader's rules."""
    L = _capi.lib()
    h = C.c_void_p()
    check(L.pp_mtx_read(str(path).encode(), int(base), C.byref(h)))
    try:
        nr, nc, nz = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(L.pp_mtx_dims(h, C.byref(nr), C.byref(nc), C.byref(nz)))
        rp = np.zeros(nr.value + 1, np.uint32)
        ci = np.zeros(max(nz.value, 1), np.uint32)
        v = np.zeros(max(nz.value, 1), np.float64)
        check(L.pp_mtx_csr(h, _p(rp, C.c_uint32), _p(ci, C.c_uint32), _p(v, C.c_double)))
    finally:
        L.pp_mtx_destroy(h)
    return rp, ci[:nz.value], v[:nz.value]


def read_mtx_dense(path):
    """A dense (array) MatrixMarket file, column-major values and (nrow, ncol)
    (ReadMtxTools::loadRightHand)."""
    L = _capi.lib()
    nr, nc = C.c_uint32(), C.c_uint32()
    check(L.pp_mtx_read_dense(str(path).encode(), C.byref(nr), C.byref(nc), None))
    vals = np.zeros(max(nr.value * nc.value, 1), np.float64)
    check(L.pp_mtx_read_dense(str(path).encode(), C.byref(nr), C.byref(nc), _p(vals, C.c_double)))
    return vals[:nr.value * nc.value], (nr.value, nc.value)


def rcm(A):
    """HOST::rcmReorderCSR (include/utils/ExternalTools/MatrixReorderTools.h:59-80):
    reverse Cuthill-McKee order of a structurally symmetric CSR pattern;
    perm[k] = the row placed at position k."""
    rp, ci = _u32(A[0]), _u32(A[1])
    n = len(rp) - 1
    perm = np.zeros(max(n, 1), np.uint32)
    cip = ci if len(ci) else np.zeros(1, np.uint32)
    check(_capi.lib().pp_reorder_rcm(n, _p(rp, C.c_uint32), _p(cip, C.c_uint32), _p(perm, C.c_uint32)))
    return perm[:n]


# ----------------------------------------------------------------------------- host factor
def _popts(kind, tau=1e-3, lfil=20, make_f32=False, **extra):
    o = _capi.PrecondOpts()
    _capi.lib().pp_precond_opts_default(C.byref(o))
    o.kind = kind
    o.tau = tau
    o.lfil = lfil
    o.make_f32 = int(make_f32)
    for k, v in extra.items():
        setattr(o, k, v)
    return o


def host_factor(A, kind, tau=1e-3, lfil=20, ordering=0):
    """setup() on the host only (no GPU): returns dict of factor arrays,
    mnnz, fail_row, status and the SpTRSV level counts.  ordering=1 factors
    the multicolour-reordered matrix."""
    kinds = {"ict": PP_PC_ICT, "ilut": PP_PC_ILUT, "ildlt": PP_PC_ILDLT, "ic0": PP_PC_IC0, "ilu0": PP_PC_ILU0}
    rp, ci, v = _u32(A[0]), _u32(A[1]), _f64(A[2])
    if len(ci) == 0:
        ci, v = np.zeros(1, np.uint32), np.zeros(1)
    L = _capi.lib()
    o = _popts(kinds[kind], tau, lfil, ordering=ordering)
    h = C.c_void_p()
    st = L.pp_factor_host(C.byref(o), len(rp) - 1, _p(rp, C.c_uint32), _p(ci, C.c_uint32), _p(v, C.c_double),
                          C.byref(h))
    out = {"status": st}
    n = len(rp) - 1
    for which, key in ((0, "lt"), (1, "l"), (2, "u")):
        nnz = C.c_uint32()
        if L.pp_hfactor_get(h, which, C.byref(nnz), None, None, None, None) != 0:
            continue
        frp = np.zeros(n + 1, np.uint32)
        fci = np.zeros(max(nnz.value, 1), np.uint32)
        fv = np.zeros(max(nnz.value, 1))
        d = np.zeros(max(n, 1))
        L.pp_hfactor_get(h, which, C.byref(nnz), _p(frp, C.c_uint32), _p(fci, C.c_uint32), _p(fv, C.c_double),
                         _p(d, C.c_double))
        out[key] = (frp, fci[:nnz.value], fv[:nnz.value])
        if kind in ("ilut", "ilu0", "ildlt"):
            out["d"] = d[:n]
    mnnz, fail, lo, up = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    L.pp_hfactor_info(h, C.byref(mnnz), C.byref(fail), C.byref(lo), C.byref(up))
    out.update(mnnz=mnnz.value, fail_row=fail.value, levels_lower=lo.value, levels_upper=up.value)
    L.pp_hfactor_destroy(h)
    return out


# ----------------------------------------------------------------------------- device context
class Context:
    """One device + one HIP stream (replaces the reference's per-call
    cudaSetDevice / StreamController)."""

    _default = {}

    def __init__(self, device=0):
        self.device = device
        h = C.c_void_p()
        check(_capi.lib().pp_ctx_create(device, C.byref(h)))
        self.h = h

    @classmethod
    def default(cls, device=0):
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def check(self, status):
        check(status, self.h)

    def sync(self):
        self.check(_capi.lib().pp_ctx_sync(self.h))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _capi.lib().pp_ctx_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _torch():
    import torch
    return torch


def _dev_ptr(t, dtype=None):
    """(pointer, pp dtype) of a contiguous CUDA tensor."""
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a CUDA tensor")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    dt = {torch.float64: PP_F64, torch.float32: PP_F32}.get(t.dtype)
    if dt is None or (dtype is not None and dt != dtype):
        raise TypeError(f"unsupported dtype {t.dtype}")
    return C.c_void_p(t.data_ptr()), dt


def _before(ctx):
    # producer work queued by torch must be finished before the context's stream reads it
    _torch().cuda.current_stream(ctx.device).synchronize()


class CSRMatrix:
    """Device CSR (DEVICE::DeviceCSRMatrix, include/MatrixClass/DeviceCSRMatrix.cuh:19-93)
    built from host arrays (HOST::CSRMatrix layout: u32 row offsets / columns,
    FP64 values)."""

    def __init__(self, rp, ci, v, ncol=None, ctx=None):
        self.ctx = ctx or Context.default()
        self.rp, self.ci, self.v = _u32(rp), _u32(ci), _f64(v)
        n = len(self.rp) - 1
        self.n = n
        self.ncol = n if ncol is None else ncol
        ci_, v_ = (self.ci, self.v) if len(self.ci) else (np.zeros(1, np.uint32), np.zeros(1))
        h = C.c_void_p()
        self.ctx.check(_capi.lib().pp_csr_create(self.ctx.h, n, self.ncol, _p(self.rp, C.c_uint32),
                                                 _p(ci_, C.c_uint32), _p(v_, C.c_double), C.byref(h)))
        self.h = h

    @classmethod
    def from_scipy(cls, M, ctx=None):
        M = M.tocsr()
        M.sort_indices()
        return cls(M.indptr, M.indices, M.data, ncol=M.shape[1], ctx=ctx)

    def getRowNum(self):
        return self.n

    def getColNum(self):
        return self.ncol

    def getNNZnum(self):
        return int(self.rp[-1])

    def host(self):
        return self.rp, self.ci, self.v

    def set_values(self, values):
        """Replace the values (same pattern); numpy (host) or CUDA tensor."""
        L = _capi.lib()
        if isinstance(values, np.ndarray):
            v = _f64(values)
            self.v = v.copy()
            self.ctx.check(L.pp_csr_set_values(self.ctx.h, self.h, v.ctypes.data_as(C.c_void_p), 0))
        else:
            _before(self.ctx)
            p, _ = _dev_ptr(values, PP_F64)
            self.ctx.check(L.pp_csr_set_values(self.ctx.h, self.h, p, 1))
            self.v = values.detach().cpu().numpy().copy()
        self.values_version = getattr(self, "values_version", 0) + 1

    def MatPVec(self, x, y, alpha=1.0, beta=0.0):
        """y = alpha A x + beta y (CSRMatrix::MatPVec, CSRMatrix.h:379-447)."""
        _before(self.ctx)
        px, dt = _dev_ptr(x)
        py, _ = _dev_ptr(y, dt)
        self.ctx.check(_capi.lib().pp_spmv(self.ctx.h, self.h, dt, alpha, px, beta, py))
        self.ctx.sync()
        return y

    def transMatPVec(self, x, y, alpha=1.0, beta=1.0):
        """y = alpha A^T x + beta y (CSRMatrix::transMatPVec accumulates: beta = 1)."""
        _before(self.ctx)
        px, dt = _dev_ptr(x)
        py, _ = _dev_ptr(y, dt)
        self.ctx.check(_capi.lib().pp_spmv_t(self.ctx.h, self.h, dt, alpha, px, beta, py))
        self.ctx.sync()
        return y

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _capi.lib().pp_csr_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ----------------------------------------------------------------------------- preconditioners
class BasePrecondition:
    """BasePrecondition<T> (include/Preconditioner/BasePreconditon.h:34-145)."""

    _kind = PP_PC_NONE
    _type = PreconditionNone

    def __init__(self, A, ctx=None, make_f32=False, **opts):
        self.A = A if isinstance(A, CSRMatrix) else CSRMatrix(*A, ctx=ctx)
        self.ctx = self.A.ctx
        self._opts = _popts(self._kind, make_f32=make_f32, **opts)
        steps = getattr(self, "_lanczos_steps", 0)
        if steps:
            self._opts.lanczos_steps = steps
        h = C.c_void_p()
        self.ctx.check(_capi.lib().pp_precond_create(self.ctx.h, self.A.h, C.byref(self._opts), C.byref(h)))
        self.h = h

    def setup(self):
        self.ctx.check(_capi.lib().pp_precond_setup(self.ctx.h, self.h))

    def refactor(self):
        """Numeric re-factorization from the matrix's current values (same
        pattern): on the device for IC(0) / ILU(0), the full setup otherwise
        (no reference counterpart; the reference's setup() is a no-op once ready)."""
        self.ctx.check(_capi.lib().pp_precond_refactor(self.ctx.h, self.h))

    def isReady(self):
        return self.info().ready

    def getPreconditionType(self):
        return self._type

    def info(self):
        i = _capi.PrecondInfo()
        check(_capi.lib().pp_precond_info_get(self.h, C.byref(i)))
        return i

    def getPreconditionNonZeroCount(self):
        return self.info().mnnz

    def getOriginalNonZeroCount(self):
        return self.info().annz

    def getPreconditionFillinRatio(self):
        return self.info().fill_ratio

    def MInvSolve(self, vec, out=None):
        """vec = M^{-1} vec in place, or out = M^{-1} vec (BasePreconditon.h:102-128)."""
        _before(self.ctx)
        px, dt = _dev_ptr(vec)
        po = px if out is None else _dev_ptr(out, dt)[0]
        self.ctx.check(_capi.lib().pp_precond_apply(self.ctx.h, self.h, dt, px, po))
        self.ctx.sync()
        return vec if out is None else out

    def _part(self, part, vec):
        _before(self.ctx)
        px, dt = _dev_ptr(vec)
        self.ctx.check(_capi.lib().pp_precond_apply_part(self.ctx.h, self.h, part, dt, px))
        self.ctx.sync()
        return vec

    def MSolveLower(self, vec):
        return self._part(0, vec)

    def MSolveDiagonal(self, vec):
        return self._part(1, vec)

    def MSolveUpper(self, vec):
        return self._part(2, vec)

    def factor(self, which):
        """Host copy of the factor: 0 L^T, 1 L, 2 U -> (rp, ci, v, d)."""
        L = _capi.lib()
        nnz = C.c_uint32()
        check(L.pp_precond_factor(self.h, which, C.byref(nnz), None, None, None, None))
        n = self.A.n
        rp = np.zeros(n + 1, np.uint32)
        ci = np.zeros(max(nnz.value, 1), np.uint32)
        v = np.zeros(max(nnz.value, 1))
        d = np.zeros(max(n, 1))
        check(L.pp_precond_factor(self.h, which, C.byref(nnz), _p(rp, C.c_uint32), _p(ci, C.c_uint32),
                                  _p(v, C.c_double), _p(d, C.c_double)))
        return rp, ci[:nnz.value], v[:nnz.value], d[:n]

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _capi.lib().pp_precond_destroy(self.h)
                self.h = None
        except Exception:
            pass


class IncompleteCholesky(BasePrecondition):
    """IncompleteCholesky<Low,High>(A, threshold) (IncompleteCholesky.cpp:12-33)."""
    _kind, _type = PP_PC_ICT, PreconditionIC

    def __init__(self, A, threshold=1e-3, **kw):
        super().__init__(A, tau=threshold, **kw)


class IncompleteLU(BasePrecondition):
    """IncompleteLU<Low,High>(A, lfill, tolerance) (IncompleteLU.cpp:12-35)."""
    _kind, _type = PP_PC_ILUT, PreconditionILUT

    def __init__(self, A, lfill=20, tolerance=1e-2, **kw):
        super().__init__(A, tau=tolerance, lfil=lfill, **kw)


class IncompleteLDLT(BasePrecondition):
    """IncompleteLDLT<Low,High>(A, tolerance) (IncompleteLDLT.cpp)."""
    _kind, _type = PP_PC_ILDLT, PreconditionILDLT

    def __init__(self, A, tolerance=1e-3, **kw):
        super().__init__(A, tau=tolerance, **kw)


class IC0(BasePrecondition):
    """Zero-fill incomplete Cholesky (no reference counterpart)."""
    _kind, _type = PP_PC_IC0, PreconditionIC


class ILU0(BasePrecondition):
    """Zero-fill incomplete LU (no reference counterpart)."""
    _kind, _type = PP_PC_ILU0, PreconditionILUT


# ----------------------------------------------------------------------------- solvers
class _Solver:
    _solver = PP_SOLVER_CG

    def __init__(self, n, precond=None, tol=1e-8, maxIter=1000, flexible=True, precision="fp64"):
        self.n = n
        self.precond = precond
        self.opts = _capi.SolveOpts()
        _capi.lib().pp_solve_opts_default(C.byref(self.opts), self._solver)
        self.opts.tol = tol
        self.opts.maxit = maxIter
        self.opts.flexible = int(flexible)
        self.opts.precision = PP_MIXED if precision == "mixed" else PP_FP64
        self.report = _capi.SolveReport()

    def setCGVariant(self, variant):
        """"auto" (reference FlexibleCG in FP64; for the FP32 inner solves of
        mixed precision the reduced CG where it applies, else the
        single-reduction PCG), "reference", "single_reduction" or "reduced"
        (IC(0) of a two-colour ordering applied in Schur-complement form: CG
        on the second colour's unknowns, no triangular sweeps)."""
        self.opts.cg_variant = {"auto": _capi.PP_CG_AUTO, "reference": _capi.PP_CG_REFERENCE,
                                "single_reduction": _capi.PP_CG_SINGLE_REDUCTION,
                                "reduced": _capi.PP_CG_REDUCED}[variant]
        return self

    def solve(self, A, b, x):
        """BaseSolver::solve(A, b, x) (include/Solver/BaseSolver.h:16).  b, x:
        numpy (host in / host out, x0 in x) or CUDA float64 tensors (device-resident)."""
        if not isinstance(A, CSRMatrix):
            A = CSRMatrix(*A)
        ctx = A.ctx
        M = self.precond.h if self.precond is not None else None
        if self.precond is not None and not self.precond.isReady():
            self.precond.setup()
        L = _capi.lib()
        rep = _capi.SolveReport()
        if isinstance(b, np.ndarray):
            bb = _f64(b)
            if x.dtype != np.float64 or not x.flags.c_contiguous:
                raise TypeError("x must be a contiguous float64 array")
            st = L.pp_solve_host(ctx.h, A.h, M, C.byref(self.opts), _p(bb, C.c_double), _p(x, C.c_double),
                                 C.byref(rep))
        else:
            _before(ctx)
            pb, _ = _dev_ptr(b, PP_F64)
            px, _ = _dev_ptr(x, PP_F64)
            st = L.pp_solve(ctx.h, A.h, M, C.byref(self.opts), pb, px, C.byref(rep))
        self.report = rep
        self._ctx = ctx
        ctx.check(st)
        return x

    def getNumIter(self):
        return self.report.iters

    def getError(self):
        return self.report.final_error

    def getConvergence(self):
        return bool(self.report.converged)

    def history(self):
        L = _capi.lib()
        cnt = C.c_uint32()
        L.pp_solve_history(self._ctx.h, None, 0, C.byref(cnt))
        out = np.zeros(max(cnt.value, 1))
        L.pp_solve_history(self._ctx.h, _p(out, C.c_double), cnt.value, C.byref(cnt))
        return out[:cnt.value]


class CG(_Solver):
    """CG<T>(problemSize, precond, CGType, tol, maxIter) (include/Solver/CG.h:35-38)."""
    _solver = PP_SOLVER_CG

    def __init__(self, problemSize, precond=None, CGType=FlexibleCG, tol=1e-8, maxIter=1000, precision="fp64"):
        super().__init__(problemSize, precond, tol, maxIter, CGType == FlexibleCG, precision)


class GMRES(_Solver):
    """GMRES<T>(problemSize, precond, tolerance, GMRESType) + setRestart
    (include/Solver/GMRES.h:39-49).  maxIter is the reference's total budget
    (MAX_ITER_NUM_SOLVER = 500), divided by the restart length."""
    _solver = PP_SOLVER_GMRES

    def __init__(self, problemSize, precond=None, tolerance=1e-7, GMRESType=FlexibleGMRES, maxIter=500,
                 precision="fp64", reference_quirks=False):
        super().__init__(problemSize, precond, tolerance, maxIter, GMRESType == FlexibleGMRES, precision)
        self.opts.restart = max(1, problemSize - 1)
        self.opts.reference_quirks = int(reference_quirks)

    def setRestart(self, restart, m):
        self.opts.restart = m if restart else max(1, self.n - 1)


# ----------------------------------------------------------------------------- batched systems
class BatchedSystems:
    """Many independent SPD systems (block-diagonal CSR) solved together by ONE
    persistent, graph-captured IC(0)-PCG launch, one workgroup per system
    (no reference counterpart).  `systems` is a list of (rp, ci, v) CSR
    triples; right-hand sides and solutions are concatenated CUDA float64
    tensors.  Solves start from x0 = 0."""

    def __init__(self, systems, ctx=None, ordering=0):
        self.ctx = ctx or Context.default()
        sizes = [len(s[0]) - 1 for s in systems]
        self.row_off = np.zeros(len(systems) + 1, np.uint32)
        self.row_off[1:] = np.cumsum(sizes)
        rps, cis, vs, base_nnz = [np.zeros(1, np.uint32)], [], [], 0
        for s, (rp, ci, v) in enumerate(systems):
            rps.append(_u32(rp)[1:] + base_nnz)
            cis.append(_u32(ci) + self.row_off[s])
            vs.append(_f64(v))
            base_nnz += int(rp[-1])
        self.rp = np.concatenate(rps).astype(np.uint32)
        self.ci = np.concatenate(cis).astype(np.uint32)
        self.v = np.concatenate(vs)
        self.nsys, self.n, self.nnz = len(systems), int(self.row_off[-1]), len(self.ci)
        h = C.c_void_p()
        self.ctx.check(_capi.lib().pp_batch_create(self.ctx.h, self.nsys, _p(self.row_off, C.c_uint32),
                                                   _p(self.rp, C.c_uint32), _p(self.ci, C.c_uint32),
                                                   _p(self.v, C.c_double), C.byref(h)))
        self.h = h
        self.block = 256
        self.report = _capi.SolveReport()
        self.ordering = 0
        if ordering:
            self.set_ordering(ordering)

    def set_ordering(self, ordering):
        """1: multicolour order inside every system (IC(0) sweeps with one level
        per colour); 0: natural order.  Re-factors; b and x keep the caller's order."""
        self.ctx.check(_capi.lib().pp_batch_set_ordering(self.ctx.h, self.h, int(ordering)))
        self.ordering = int(ordering)

    def tune(self):
        """Pick the workgroup size (row tile of the SpMV and sweep steps) by timing."""
        bs = C.c_int()
        self.ctx.check(_capi.lib().pp_batch_tune(self.ctx.h, self.h, C.byref(bs)))
        self.block = bs.value
        return bs.value

    def set_values(self, values):
        """New values on the same pattern (numpy or CUDA tensor); re-factors IC(0)."""
        L = _capi.lib()
        if isinstance(values, np.ndarray):
            v = _f64(values)
            self.ctx.check(L.pp_batch_set_values(self.ctx.h, self.h, v.ctypes.data_as(C.c_void_p), 0))
        else:
            _before(self.ctx)
            p, _ = _dev_ptr(values.contiguous(), PP_F64)
            self.ctx.check(L.pp_batch_set_values(self.ctx.h, self.h, p, 1))
        self.values_version = getattr(self, "values_version", 0) + 1

    def _solve(self, fn, b, x, precision, tol, maxit, inner_tol):
        o = _capi.SolveOpts()
        L = _capi.lib()
        L.pp_solve_opts_default(C.byref(o), PP_SOLVER_CG)
        o.precision = PP_MIXED if precision == "mixed" else PP_FP64
        o.tol, o.maxit, o.inner_tol = tol, maxit, inner_tol
        its = np.zeros(self.nsys, np.int32)
        rep = _capi.SolveReport()
        _before(self.ctx)
        pb, _ = _dev_ptr(b, PP_F64)
        px, _ = _dev_ptr(x, PP_F64)
        self.ctx.check(fn(self.ctx.h, self.h, C.byref(o), pb, px, _p(its, C.c_int32), C.byref(rep)))
        self.report = rep
        return its

    def solve(self, b, x, precision="mixed", tol=1e-8, maxit=20000, inner_tol=1e-6):
        """x = A^{-1} b for every system; returns per-system iteration counts."""
        return self._solve(_capi.lib().pp_batch_solve, b, x, precision, tol, maxit, inner_tol)

    def solve_t(self, g, lam, precision="mixed", tol=1e-8, maxit=20000, inner_tol=1e-6):
        """lambda = A^{-T} g (the adjoint solve)."""
        return self._solve(_capi.lib().pp_batch_solve_t, g, lam, precision, tol, maxit, inner_tol)

    def outer_on_pattern(self, lam, x, scale=-1.0):
        """grad[k] = scale * lam[row(k)] * x[col(k)] on the batch pattern."""
        torch = _torch()
        g = torch.empty(self.nnz, dtype=torch.float64, device=lam.device)
        _before(self.ctx)
        self.ctx.check(_capi.lib().pp_batch_outer_on_pattern(self.ctx.h, self.h, scale, _dev_ptr(lam, PP_F64)[0],
                                                              _dev_ptr(x, PP_F64)[0], _dev_ptr(g, PP_F64)[0]))
        return g

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _capi.lib().pp_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


class AMSED(BasePrecondition):
    """AMSEDPrecondition<Low,High> (include/Preconditioner/AMSEDPrecondition.h:75-533):
    multilevel low-rank Schur correction for SPD matrices.  levels / ndom give
    the hierarchical interface decomposition (ndom leaf subdomains, a power of
    two), rank the low-rank correction per level; tolerance is the ILDLT/ICT
    drop tolerance of the block factors (0 = exact blocks)."""
    _kind, _type = _capi.PP_PC_AMSED, PreconditionAMSED

    _LOWRANK = {"auto": 0, "reference": 1, "deflation": 2, "subspace": 3}

    def __init__(self, A, levels=3, ndom=16, rank=32, tolerance=1e-3, lanczos_steps=0, exact=False,
                 lowrank="auto", eps=0.8, **kw):
        """lowrank: "reference" (default "auto"; AMSED: the reference's thick-restart
        Lanczos with H = mu / (1 - mu); GMSLR: its first Schur pairs with a
        diagonal core), "deflation" (AMSED: H = (eps - lam) / lam for mu > 1 - eps)
        or "subspace" (this build's one-run Lanczos / exact-subspace core)."""
        self._lanczos_steps = lanczos_steps
        super().__init__(A, tau=0.0 if exact else tolerance, levels=levels, ndom=ndom, rank=rank,
                         lowrank=self._LOWRANK[lowrank] if isinstance(lowrank, str) else int(lowrank), eps=eps,
                         **kw)

    def info_levels(self):
        nlev = C.c_uint32()
        sizes, ranks = (C.c_uint32 * 32)(), (C.c_uint32 * 32)()
        check(_capi.lib().pp_ml_info(self.h, C.byref(nlev), sizes, ranks, 32))
        return list(sizes)[:nlev.value], list(ranks)[:nlev.value]

    def permutation(self):
        perm = np.zeros(self.A.n, np.uint32)
        check(_capi.lib().pp_ml_permutation(self.h, _p(perm, C.c_uint32)))
        return perm


class GMSLR(AMSED):
    """GMSLRPrecondition<Low,High> (include/Preconditioner/GMSLRPrecondition.h:48-399):
    the nonsymmetric multilevel low-rank Schur preconditioner (ILUT blocks,
    lower coupling F, Arnoldi low-rank term)."""
    _kind, _type = _capi.PP_PC_GMSLR, PreconditionGMSLR

    def __init__(self, A, levels=3, ndom=16, rank=32, tolerance=1e-2, lfill=20, lanczos_steps=0, exact=False,
                 lowrank="auto", **kw):
        super().__init__(A, levels=levels, ndom=ndom, rank=rank, tolerance=tolerance, lanczos_steps=lanczos_steps,
                         exact=exact, lfil=lfill if not exact else 1 << 30, lowrank=lowrank, **kw)
