// ref_gpu_harness.cpp -- the reference's OWN solvers (HOST::CG<T>,
// HOST::GMRES<T>, compiled unmodified from /root/reference by
// oracle/Makefile) driving the GPU preconditioners of
// include/prispack_amd_reference.hpp, which derive from the reference's own
// TriangularPrecondition<T>.  A C-ABI shim so that the GPU tests can run the
// integration a PrisPack maintainer would build, and compare it with the
// reference's CPU preconditioners on the same inputs.
//
// TEST INFRASTRUCTURE ONLY: the shipped library never links or loads this.
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "include/MatrixClass/CSRMatrix.h"
#include "include/Preconditioner/IncompleteCholesky.h"
#include "include/Preconditioner/IncompleteLDLT.h"
#include "include/Preconditioner/IncompleteLU.h"
#include "include/Solver/CG.h"
#include "include/Solver/GMRES.h"
#include "include/VectorClass/AutoAllocateVector.h"
#include "prispack_amd_reference.hpp"

namespace {

thread_local std::string g_err;

template <typename T>
std::shared_ptr<HOST::CSRMatrix<T>> make_matrix(uint32_t n, const uint32_t *rp, const uint32_t *ci,
                                                const double *v) {
    const uint32_t nnz = rp[n];
    auto a = std::make_shared<HOST::CSRMatrix<T>>(n, n, nnz, memoryBase);
    std::memcpy(a->getRowOffsetPtr(0), rp, (n + 1) * sizeof(uint32_t));
    std::memcpy(a->getColIndicesPtr(0), ci, nnz * sizeof(uint32_t));
    T *av = a->getCSRValuesPtr(0);
    for (uint32_t k = 0; k < nnz; ++k) av[k] = (T)v[k];
    return a;
}

// kind: 1 IncompleteCholesky(tau), 2 IncompleteLU(lfil, tau), 3 IncompleteLDLT(tau)
template <typename T>
std::shared_ptr<prispack_gpu::GPUTriangularPrecondition<T>> make_gpu(
    int kind, const std::shared_ptr<HOST::CSRMatrix<T>> &a, double tau, uint32_t lfil, int trsv) {
    if (kind == 1) return std::make_shared<prispack_gpu::IncompleteCholesky<T>>(a, tau, trsv);
    if (kind == 2) return std::make_shared<prispack_gpu::IncompleteLU<T>>(a, lfil, tau, trsv);
    return std::make_shared<prispack_gpu::IncompleteLDLT<T>>(a, tau, trsv);
}

template <typename T>
std::shared_ptr<TriangularPrecondition<T>> make_cpu(int kind, const std::shared_ptr<HOST::CSRMatrix<T>> &a,
                                                    double tau, uint32_t lfil) {
    if (kind == 1) return std::make_shared<HOST::IncompleteCholesky<T, T>>(a, tau);
    if (kind == 2) return std::make_shared<HOST::IncompleteLU<T, T>>(a, lfil, tau);
    return std::make_shared<HOST::IncompleteLDLT<T, T>>(a, tau);
}

template <typename T, typename F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// HOST::CG<T> with a GPU (gpu != 0) or the reference's CPU preconditioner;
// x holds x0 on entry and the solution on return.
template <typename T>
int run_cg(int gpu, int kind, uint32_t n, const uint32_t *rp, const uint32_t *ci, const double *v,
           const double *b, double *x, double tau, uint32_t lfil, int trsv, double tol, uint32_t maxit, int *conv,
           double *err) {
    return guarded<T>([&] {
        auto a = make_matrix<T>(n, rp, ci, v);
        std::shared_ptr<BasePrecondition<T>> M;
        if (gpu) M = make_gpu<T>(kind, a, tau, lfil, trsv);
        else M = make_cpu<T>(kind, a, tau, lfil);
        HOST::CG<T> cg(n, M, HOST::FlexibleCG, tol, maxit);
        HOST::AutoAllocateVector<T> vb(n, memoryBase), vx(n, memoryBase);
        for (uint32_t i = 0; i < n; ++i) {
            (*vb)[i] = (T)b[i];
            (*vx)[i] = (T)x[i];
        }
        cg.solve(*a, *vb, *vx);
        for (uint32_t i = 0; i < n; ++i) x[i] = (double)(*vx)[i];
        *conv = cg.getConvergence();
        *err = cg.getError();
        return (int)cg.getNumIter();
    });
}

}  // namespace

extern "C" {

const char *refgpu_error(void) { return g_err.c_str(); }

int refgpu_cg(int gpu, int kind, uint32_t n, const uint32_t *rp, const uint32_t *ci, const double *v,
              const double *b, double *x, double tau, uint32_t lfil, int trsv, double tol, uint32_t maxit, int *conv,
              double *err) {
    return run_cg<double>(gpu, kind, n, rp, ci, v, b, x, tau, lfil, trsv, tol, maxit, conv, err);
}

int refgpu_cg_f32(int gpu, int kind, uint32_t n, const uint32_t *rp, const uint32_t *ci, const double *v,
                  const double *b, double *x, double tau, uint32_t lfil, int trsv, double tol, uint32_t maxit,
                  int *conv, double *err) {
    return run_cg<float>(gpu, kind, n, rp, ci, v, b, x, tau, lfil, trsv, tol, maxit, conv, err);
}

// HOST::GMRES<double>(n, M, tol, FlexibleGMRES) + setRestart(true, restart)
int refgpu_gmres(int gpu, int kind, uint32_t n, const uint32_t *rp, const uint32_t *ci, const double *v,
                 const double *b, double *x, double tau, uint32_t lfil, int trsv, double tol, uint32_t restart,
                 int *conv, double *err) {
    return guarded<double>([&] {
        auto a = make_matrix<double>(n, rp, ci, v);
        std::shared_ptr<BasePrecondition<double>> M;
        if (gpu) M = make_gpu<double>(kind, a, tau, lfil, trsv);
        else M = make_cpu<double>(kind, a, tau, lfil);
        HOST::GMRES<double> g(n, M, tol, HOST::FlexibleGMRES);
        g.setRestart(true, restart);
        HOST::AutoAllocateVector<double> vb(n, memoryBase), vx(n, memoryBase);
        std::memcpy(vb->getRawValPtr(), b, n * sizeof(double));
        std::memcpy(vx->getRawValPtr(), x, n * sizeof(double));
        g.solve(*a, *vb, *vx);
        std::memcpy(x, vx->getRawValPtr(), n * sizeof(double));
        *conv = g.getConvergence();
        *err = g.getError();
        return (int)g.getNumIter();
    });
}

// Every entry of the preconditioner interface on one input, through the
// base-class pointer the reference's solvers hold (gpu != 0: the GPU
// adapter; 0: the reference's CPU class).  out receives, each of length n:
//   [0] MInvSolve(vec)                  [1] MInvSolve(vec, start) on a padded vector
//   [2] MInvSolve(in, out)              [3] MInvSolve(in, inStart, out, outStart)
//   [4] MSolveLower(vec)  [5] MSolveDiagonal(vec)  [6] MSolveUpper(vec)
// and info[0..2] = isReady, getPreconditionNonZeroCount, getPreconditionType.
int refgpu_interface(int gpu, int kind, uint32_t n, const uint32_t *rp, const uint32_t *ci, const double *v,
                     double tau, uint32_t lfil, int trsv, const double *x, double *out, uint32_t *info) {
    return guarded<double>([&] {
        auto a = make_matrix<double>(n, rp, ci, v);
        std::shared_ptr<TriangularPrecondition<double>> M;
        if (gpu) M = make_gpu<double>(kind, a, tau, lfil, trsv);
        else M = make_cpu<double>(kind, a, tau, lfil);
        M->setup();
        BasePrecondition<double> *base = M.get();
        const uint32_t pad = 3;
        HOST::AutoAllocateVector<double> v0(n, memoryBase), v1(n + 2 * pad, memoryBase), vin(n, memoryBase),
            vout(n, memoryBase), vin2(n + pad, memoryBase), vout2(n + 2 * pad, memoryBase);
        std::memcpy(v0->getRawValPtr(), x, n * sizeof(double));
        base->MInvSolve(*v0);
        std::memcpy(out, v0->getRawValPtr(), n * sizeof(double));
        std::memcpy(v1->getRawValPtr() + pad, x, n * sizeof(double));
        base->MInvSolve(*v1, pad);
        std::memcpy(out + n, v1->getRawValPtr() + pad, n * sizeof(double));
        std::memcpy(vin->getRawValPtr(), x, n * sizeof(double));
        base->MInvSolve(*vin, *vout);
        std::memcpy(out + 2 * n, vout->getRawValPtr(), n * sizeof(double));
        std::memcpy(vin2->getRawValPtr() + pad, x, n * sizeof(double));
        base->MInvSolve(*vin2, pad, *vout2, 2 * pad);
        std::memcpy(out + 3 * n, vout2->getRawValPtr() + 2 * pad, n * sizeof(double));
        for (int p = 0; p < 3; ++p) {
            HOST::AutoAllocateVector<double> w(n, memoryBase);
            std::memcpy(w->getRawValPtr(), x, n * sizeof(double));
            if (p == 0) M->MSolveLower(*w);
            else if (p == 1) M->MSolveDiagonal(*w);
            else M->MSolveUpper(*w);
            std::memcpy(out + (4 + p) * n, w->getRawValPtr(), n * sizeof(double));
        }
        info[0] = base->isReady();
        info[1] = base->getPreconditionNonZeroCount();
        info[2] = (uint32_t)base->getPreconditionType();
        return 0;
    });
}

}  // extern "C"
