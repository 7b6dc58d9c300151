"""Differentiable sparse solves: torch.autograd.Function wrappers over the
C ABI (no reference counterpart).

For x = A^{-1} b and a loss L(x), the adjoint solve A^T lambda = dL/dx gives
    dL/db    = lambda
    dL/dA_ij = -lambda_i x_j      (on the sparsity pattern of A only).
Both the forward and the adjoint solve run in the HIP library; the pattern
outer product is a HIP kernel too.

The backward pass solves with the values of its own forward pass: when the
matrix received other values in between (a second forward on the same
pattern), it sets the saved forward values again before the adjoint solve,
and the matrix keeps them afterwards.
"""
import ctypes as C

import torch

from . import _capi


class _BatchedSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, batch, precision, tol):
        vals = values.detach().clone()
        batch.set_values(vals)
        x = torch.zeros_like(b)
        batch.solve(b.detach().contiguous(), x, precision=precision, tol=tol)
        ctx.batch, ctx.precision, ctx.tol = batch, precision, tol
        ctx.version = batch.values_version
        ctx.save_for_backward(x, vals)
        return x

    @staticmethod
    def backward(ctx, gx):
        x, vals = ctx.saved_tensors
        if ctx.batch.values_version != ctx.version:  # another forward set other values since
            ctx.batch.set_values(vals)
            ctx.version = ctx.batch.values_version
        lam = torch.zeros_like(x)
        ctx.batch.solve_t(gx.contiguous(), lam, precision=ctx.precision, tol=ctx.tol)
        gvals = ctx.batch.outer_on_pattern(lam, x, scale=-1.0)
        return gvals, lam, None, None, None


def batched_solve(values, b, batch, precision="mixed", tol=1e-8):
    """x = A(values)^{-1} b for a BatchedSystems pattern; differentiable in
    values (concatenated CSR values) and b."""
    return _BatchedSolve.apply(values, b, batch, precision, tol)


class _SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, A, make_precond, solver_kw):
        from . import CG, GMRES
        vals = values.detach().clone()
        A.set_values(vals)
        M = make_precond(A)
        M.setup()
        kw = dict(solver_kw)
        gmres = kw.pop("gmres", False)
        tol = kw.pop("tol", 1e-8)
        restart = kw.pop("restart", 30)
        if gmres:
            S = GMRES(A.n, M, tolerance=tol, **kw)
            S.setRestart(True, restart)
        else:
            S = CG(A.n, M, tol=tol, **kw)
        x = torch.zeros_like(b)
        S.solve(A, b.detach().contiguous(), x)
        ctx.A, ctx.make_precond, ctx.kw = A, make_precond, solver_kw
        ctx.version = A.values_version
        ctx.save_for_backward(x, vals)
        return x

    @staticmethod
    def backward(ctx, gx):
        x, vals = ctx.saved_tensors
        A = ctx.A
        if A.values_version != ctx.version:  # another forward set other values since
            A.set_values(vals)
            ctx.version = A.values_version
        L = _capi.lib()
        lam = torch.zeros_like(x)
        kw = dict(ctx.kw)
        M = ctx.make_precond(A)
        popts = M._opts
        o = _capi.SolveOpts()
        L.pp_solve_opts_default(C.byref(o), _capi.PP_SOLVER_GMRES if kw.get("gmres") else _capi.PP_SOLVER_CG)
        o.tol = kw.get("tol", 1e-8)
        rep = _capi.SolveReport()
        torch.cuda.current_stream().synchronize()
        A.ctx.check(L.pp_adjoint_solve(A.ctx.h, A.h, C.byref(popts), C.byref(o), C.c_void_p(gx.contiguous().data_ptr()),
                                       C.c_void_p(lam.data_ptr()), C.byref(rep)))
        g = torch.empty(A.getNNZnum(), dtype=torch.float64, device=x.device)
        A.ctx.check(L.pp_csr_outer_on_pattern(A.ctx.h, A.h, -1.0, C.c_void_p(lam.data_ptr()),
                                              C.c_void_p(x.data_ptr()), C.c_void_p(g.data_ptr())))
        A.ctx.sync()
        return g, lam, None, None, None


def sparse_solve(values, b, A, make_precond, **solver_kw):
    """x = A(values)^{-1} b on the pattern of the CSRMatrix A; differentiable in
    values and b.  make_precond(A) builds the preconditioner (e.g. IC0, ILU0)."""
    return _SparseSolve.apply(values, b, A, make_precond, solver_kw)
