// ref_harness.cpp -- a C-ABI shim over the UNMODIFIED reference sources
// (PrisPack, compiled from /root/reference by oracle/Makefile into
// oracle/_ref/).  It only constructs the reference's own classes from raw
// arrays and calls their public methods, so that tests/golden/make_golden.py
// can write golden vectors and bench.py can time the reference CPU path.
//
// TEST INFRASTRUCTURE ONLY: the shipped library never links or loads this.
#include <cstdint>
#include <cstring>
#include <memory>

#include "include/MatrixClass/CSRMatrix.h"
#include "include/MatrixClass/ModelProblem.h"
#include "include/Preconditioner/IncompleteCholesky.h"
#include "include/Preconditioner/IncompleteLDLT.h"
#include "include/Preconditioner/IncompleteLU.h"
#include "include/Solver/CG.h"
#include "include/Solver/GMRES.h"
#include "include/VectorClass/AutoAllocateVector.h"
#include "include/utils/TestTools/ReadMtxTools.h"
#include "rcm.hpp"

using Csr = HOST::CSRMatrix<double>;
using Prec = BasePrecondition<double>;

struct RefCsr { std::shared_ptr<Csr> a; };
struct RefPrec {
    int kind = 0;  // 1 IC, 2 ILUT, 3 ILDLT
    std::shared_ptr<Prec> p;
    std::shared_ptr<HOST::IncompleteCholesky<double, double>> ic;
    std::shared_ptr<HOST::IncompleteLU<double, double>> ilu;
    std::shared_ptr<HOST::IncompleteLDLT<double, double>> ildlt;
};

static void copy_csr_out(const Csr &m, uint32_t *rp, uint32_t *ci, double *v) {
    uint32_t n = m.getRowNum();
    const uint32_t *mrp = m.getRowOffsetPtr(0);
    uint32_t nnz = mrp[n];
    if (rp) std::memcpy(rp, mrp, (n + 1) * sizeof(uint32_t));
    if (nnz && ci) std::memcpy(ci, m.getColIndicesPtr(0), nnz * sizeof(uint32_t));
    if (nnz && v) std::memcpy(v, m.getCSRValuesPtr(0), nnz * sizeof(double));
}

extern "C" {

void *refh_csr(uint32_t n, const uint32_t *rp, const uint32_t *ci, const double *v) {
    auto *h = new RefCsr;
    uint32_t nnz = rp[n];
    h->a = std::make_shared<Csr>(n, n, nnz, memoryBase);
    std::memcpy(h->a->getRowOffsetPtr(0), rp, (n + 1) * sizeof(uint32_t));
    if (nnz) {
        std::memcpy(h->a->getColIndicesPtr(0), ci, nnz * sizeof(uint32_t));
        std::memcpy(h->a->getCSRValuesPtr(0), v, nnz * sizeof(double));
    }
    return h;
}

void *refh_laplacian(int nx, int ny, int nz, double ax, double ay, double az, double shift) {
    auto *h = new RefCsr;
    h->a = std::make_shared<Csr>();
    HOST::generateLaplacianCSR(*h->a, nx, ny, nz, ax, ay, az, shift);
    return h;
}

void refh_csr_free(void *h) { delete static_cast<RefCsr *>(h); }
uint32_t refh_csr_n(void *h) { return static_cast<RefCsr *>(h)->a->getRowNum(); }
uint32_t refh_csr_nnz(void *h) {
    auto &a = *static_cast<RefCsr *>(h)->a;
    return a.getRowOffsetPtr(0)[a.getRowNum()];
}
void refh_csr_get(void *h, uint32_t *rp, uint32_t *ci, double *v) {
    copy_csr_out(*static_cast<RefCsr *>(h)->a, rp, ci, v);
}

void refh_spmv(void *h, const double *x, double *y) {
    static_cast<RefCsr *>(h)->a->MatPVecWithRawPtr(x, y);
}
void refh_spmv_t_acc(void *h, const double *x, double *y) {
    static_cast<RefCsr *>(h)->a->transMatPVec(x, y);
}

void *refh_ict(void *ha, double tau) {
    auto *p = new RefPrec;
    p->kind = 1;
    p->ic = std::make_shared<HOST::IncompleteCholesky<double, double>>(static_cast<RefCsr *>(ha)->a, tau);
    p->ic->setup();
    p->p = p->ic;
    return p;
}
void *refh_ilut(void *ha, uint32_t lfil, double tau) {
    auto *p = new RefPrec;
    p->kind = 2;
    p->ilu = std::make_shared<HOST::IncompleteLU<double, double>>(static_cast<RefCsr *>(ha)->a, lfil, tau);
    p->ilu->setup();
    p->p = p->ilu;
    return p;
}
void *refh_ildlt(void *ha, double tau) {
    auto *p = new RefPrec;
    p->kind = 3;
    p->ildlt = std::make_shared<HOST::IncompleteLDLT<double, double>>(static_cast<RefCsr *>(ha)->a, tau);
    p->ildlt->setup();
    p->p = p->ildlt;
    return p;
}
void refh_prec_free(void *h) { delete static_cast<RefPrec *>(h); }
int refh_prec_ready(void *h) { return static_cast<RefPrec *>(h)->p->isReady(); }
uint32_t refh_prec_mnnz(void *h) { return static_cast<RefPrec *>(h)->p->getPreconditionNonZeroCount(); }

// factor accessors: which = 0 -> IC L^T, 1 -> ILU L, 2 -> ILU U
uint32_t refh_factor_nnz(void *h, int which) {
    auto *p = static_cast<RefPrec *>(h);
    const Csr *m = which == 0 ? p->ic->getTransL().get().get()
                 : which == 1 ? p->ilu->getL().get().get() : p->ilu->getU().get().get();
    return m->getRowOffsetPtr(0)[m->getRowNum()];
}
void refh_factor_get(void *h, int which, uint32_t *rp, uint32_t *ci, double *v) {
    auto *p = static_cast<RefPrec *>(h);
    const Csr *m = which == 0 ? p->ic->getTransL().get().get()
                 : which == 1 ? p->ilu->getL().get().get() : p->ilu->getU().get().get();
    copy_csr_out(*m, rp, ci, v);
}
void refh_ilu_diag(void *h, double *d, uint32_t n) {
    auto *p = static_cast<RefPrec *>(h);
    auto &dv = p->ilu->getDiagVals();
    for (uint32_t i = 0; i < n; ++i) d[i] = (*dv)[i];
}

static void with_vec(uint32_t n, double *vec, const std::function<void(HostVector<double> &)> &f) {
    HOST::AutoAllocateVector<double> v(n, memoryBase);
    std::memcpy(v->getRawValPtr(), vec, n * sizeof(double));
    f(*v);
    std::memcpy(vec, v->getRawValPtr(), n * sizeof(double));
}

void refh_prec_apply(void *h, double *vec, uint32_t n) {
    auto *p = static_cast<RefPrec *>(h);
    with_vec(n, vec, [&](HostVector<double> &v) { p->p->MInvSolve(v); });
}
// part: 0 lower, 1 diagonal, 2 upper
void refh_prec_part(void *h, int part, double *vec, uint32_t n) {
    auto *p = static_cast<RefPrec *>(h);
    TriangularPrecondition<double> *t = nullptr;
    if (p->kind == 1) t = p->ic.get();
    if (p->kind == 2) t = p->ilu.get();
    if (p->kind == 3) t = p->ildlt.get();
    with_vec(n, vec, [&](HostVector<double> &v) {
        if (part == 0) t->MSolveLower(v);
        else if (part == 1) t->MSolveDiagonal(v);
        else t->MSolveUpper(v);
    });
}

// FlexibleCG / NormalCG: returns iterations; conv and err via pointers.
int refh_cg(void *ha, void *hp, int flexible, const double *b, double *x, double tol, uint32_t maxit,
            int *conv, double *err) {
    auto &A = *static_cast<RefCsr *>(ha)->a;
    uint32_t n = A.getRowNum();
    std::shared_ptr<Prec> prec = hp ? static_cast<RefPrec *>(hp)->p : nullptr;
    HOST::CG<double> cg(n, prec, flexible ? HOST::FlexibleCG : HOST::NormalCG, tol, maxit);
    HOST::AutoAllocateVector<double> vb(n, memoryBase), vx(n, memoryBase);
    std::memcpy(vb->getRawValPtr(), b, n * sizeof(double));
    std::memcpy(vx->getRawValPtr(), x, n * sizeof(double));
    cg.solve(A, *vb, *vx);
    std::memcpy(x, vx->getRawValPtr(), n * sizeof(double));
    *conv = cg.getConvergence();
    *err = cg.getError();
    return cg.getNumIter();
}

int refh_gmres(void *ha, void *hp, int flexible, const double *b, double *x, double tol, uint32_t restart,
               int *conv, double *err) {
    auto &A = *static_cast<RefCsr *>(ha)->a;
    uint32_t n = A.getRowNum();
    std::shared_ptr<Prec> prec = hp ? static_cast<RefPrec *>(hp)->p : nullptr;
    HOST::GMRES<double> g(n, prec, tol, flexible ? HOST::FlexibleGMRES : HOST::NormalGMRES);
    g.setRestart(true, restart);
    HOST::AutoAllocateVector<double> vb(n, memoryBase), vx(n, memoryBase);
    std::memcpy(vb->getRawValPtr(), b, n * sizeof(double));
    std::memcpy(vx->getRawValPtr(), x, n * sizeof(double));
    g.solve(A, *vb, *vx);
    std::memcpy(x, vx->getRawValPtr(), n * sizeof(double));
    *conv = g.getConvergence();
    *err = g.getError();
    return g.getNumIter();
}

// HOST::ReadMtxTools<double>(binaryPath, Apath, base).loadMatrix(csr)
// (source/utils/TestTools/ReadMtxTools.cpp:11-46, 339-384).  The reader
// changes the working directory to dirname(binaryPath); callers pass
// absolute paths and restore their directory.
void *refh_read_mtx(const char *binary_path, const char *path, int base) {
    auto *h = new RefCsr;
    h->a = std::make_shared<Csr>();
    HOST::ReadMtxTools<double> rd(const_cast<char *>(binary_path), path, (UINT8)base);
    rd.loadMatrix(*h->a);
    return h;
}

// The reference's genrcm (external_packages/RCM/lib/rcm.cpp:999), called
// the way HOST::rcmReorderCSR calls it (MatrixReorderTools.h:59-80, which
// cannot be included here: its header also pulls in SuiteSparse's amd.h):
// the CSR shifted to 1-based indices, genrcm, the result shifted back.
// perm[k] = the row placed at position k.
void refh_rcm(void *h, uint32_t *perm) {
    auto &m = *static_cast<RefCsr *>(h)->a;
    const uint32_t dim = m.getRowNum();
    INT32 *rp = (INT32 *)m.getRowOffsetPtr(0);
    INT32 *ci = (INT32 *)m.getColIndicesPtr(0);
    const uint32_t nnz = m.getRowOffsetPtr(0)[dim];
    for (uint32_t i = 0; i <= dim; ++i) rp[i] += 1;
    for (uint32_t i = 0; i < nnz; ++i) ci[i] += 1;
    genrcm((int)dim, (int)nnz, rp, ci, (int *)perm);
    for (uint32_t i = 0; i <= dim; ++i) rp[i] -= 1;
    for (uint32_t i = 0; i < nnz; ++i) ci[i] -= 1;
    for (uint32_t i = 0; i < dim; ++i) perm[i] -= 1;
}

}  // extern "C"
