// prispack_amd_reference.hpp -- the GPU triangular preconditioners as
// subclasses of PrisPack's OWN preconditioner base, so that PrisPack's own
// solvers drive the GPU applies unchanged:
//
//   auto A = std::make_shared<HOST::CSRMatrix<double>>(...);
//   auto M = std::make_shared<prispack_gpu::IncompleteCholesky<double>>(A, 1e-3);
//   M->setup();                                     // factor + upload + sweep analysis
//   HOST::CG<double> cg(n, M, HOST::FlexibleCG, 1e-8, 1000);   // the reference's CG
//   cg.solve(*A, b, x);                             // every M^-1 r runs on the GPU
//
// Replaces (PrisPack file:line):
//   BasePrecondition<T>        include/Preconditioner/BasePreconditon.h:34-145
//                              (setup, the four MInvSolve overloads, isReady,
//                              getPreconditionType, nonzero counts)
//   TriangularPrecondition<T>  BasePreconditon.h:150-358 (MSolveLower /
//                              MSolveDiagonal / MSolveUpper through the
//                              protected MSolve*UsePtr hooks)
//   IncompleteCholesky(A, threshold)        include/Preconditioner/IncompleteCholesky.h:42
//   IncompleteLU(A, lfill, tolerance)       include/Preconditioner/IncompleteLU.h
//   IncompleteLDLT(A, tolerance)            include/Preconditioner/IncompleteLDLT.h
// and the solver hook it plugs into: CG(n, shared_ptr<BasePrecondition<T>>,
// CGType, tol, maxIter) (include/Solver/CG.h:35-36), GMRES(n, precond, tol,
// GMRESType) (include/Solver/GMRES.h:39-49).
//
// Build: this header is compiled INSIDE a PrisPack build.  The include path
// must reach PrisPack's root (for "include/Preconditioner/BasePreconditon.h")
// and this repository's include/ directory; link libprispack_amd.so.
//
// Vectors: host vectors (the reference's HostVector family) are staged
// through one device buffer of the factor's length: one host-to-device copy,
// one GPU apply (both sweeps, and D for ILDLT), one device-to-host copy per
// MInvSolve.  Vectors whose getLocation() is a device id are applied in place
// on the device.  T = double applies the FP64 factor; T = float applies the
// factor rounded to FP32 (FP32 sweeps with FP32 accumulation), as the
// reference's <float, float> instantiation does with its float factor.
// Errors throw std::runtime_error with the library's message (the reference
// throws std exceptions from THROW_EXCEPTION; nothing here calls exit()).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "include/MatrixClass/CSRMatrix.h"
#include "include/Preconditioner/BasePreconditon.h"
#include "prispack_amd.h"

namespace prispack_gpu {

template <typename T>
struct DeviceType;
template <>
struct DeviceType<double> {
    static constexpr int value = PP_F64;
};
template <>
struct DeviceType<float> {
    static constexpr int value = PP_F32;
};

template <typename T>
class GPUTriangularPrecondition : public TriangularPrecondition<T> {
public:
    GPUTriangularPrecondition(const std::shared_ptr<HOST::CSRMatrix<T>> &matA, const pp_precond_opts &opts,
                              PreconditionType_t type, int device = 0)
        : m_matA(matA), m_opts(opts), m_device(device) {
        if (!matA) throw std::invalid_argument("prispack_gpu: NULL matrix");
        this->m_precondType = type;
        this->m_ArowNum = matA->getRowNum();
        this->m_AcolNum = matA->getColNum();
        this->m_Annz = matA->getRowOffsetPtr(0)[matA->getRowNum()];
        if (DeviceType<T>::value == PP_F32) m_opts.make_f32 = 1;
    }

    ~GPUTriangularPrecondition() override {
        if (m_stage) pp_free(m_ctx, m_stage);
        pp_precond_destroy(m_M);
        pp_csr_destroy(m_A);
        pp_ctx_destroy(m_ctx);
    }
    GPUTriangularPrecondition(const GPUTriangularPrecondition &) = delete;
    GPUTriangularPrecondition &operator=(const GPUTriangularPrecondition &) = delete;

    // BasePreconditon.h:102: factor on the host, upload, sweep analysis; a
    // second call does nothing, as the reference's (IncompleteCholesky.cpp:105)
    void setup() override {
        if (this->m_isReady == PRECONDITION_READY) return;
        const uint32_t n = this->m_ArowNum, nnz = this->m_Annz;
        const uint32_t *rp = m_matA->getRowOffsetPtr(0);
        if (!m_ctx) check_global(pp_ctx_create(m_device, &m_ctx));
        std::vector<double> v(m_matA->getCSRValuesPtr(0), m_matA->getCSRValuesPtr(0) + nnz);
        check(pp_csr_create(m_ctx, n, this->m_AcolNum, rp, m_matA->getColIndicesPtr(0), v.data(), &m_A));
        check(pp_precond_create(m_ctx, m_A, &m_opts, &m_M));
        check(pp_malloc(m_ctx, (size_t)n * sizeof(T), &m_stage));
        check(pp_precond_setup(m_ctx, m_M));
        pp_precond_info info{};
        check(pp_precond_info_get(m_M, &info));
        this->m_Mnnz = info.mnnz;
        this->m_isReady = PRECONDITION_READY;
    }

    // vec = M^{-1} vec (BasePreconditon.h:108)
    void MInvSolve(BaseVector<T> &vec) override { apply(vec, 0, vec, 0); }
    // vec[start : start + n) = M^{-1} vec[start : start + n) (BasePreconditon.h:114)
    void MInvSolve(BaseVector<T> &vec, UINT32 resStartIdx) override { apply(vec, resStartIdx, vec, resStartIdx); }
    // vecOUT = M^{-1} vecIN, vecOUT grown to n when shorter (BasePreconditon.h:119)
    void MInvSolve(BaseVector<T> &vecIN, BaseVector<T> &vecOUT) override {
        if (&vecIN != &vecOUT && vecOUT.getLength() < this->m_ArowNum)
            vecOUT.resize(this->m_ArowNum, RESERVE_NO_DATA);
        apply(vecIN, 0, vecOUT, 0);
    }
    // vecOUT[outStart : +n) = M^{-1} vecIN[inStart : +n) (BasePreconditon.h:126)
    void MInvSolve(BaseVector<T> &vecIN, UINT32 inStartIdx, BaseVector<T> &vecOUT, UINT32 outStartIdx) override {
        apply(vecIN, inStartIdx, vecOUT, outStartIdx);
    }

    pp_precond *handle() const { return m_M; }
    pp_ctx *context() const { return m_ctx; }

protected:
    // host pointers of the reference's TriangularPrecondition (BasePreconditon.h:153-157)
    void MSolveLowerUsePtr(T *vec) override { part(0, vec); }
    void MSolveDiagonalUsePtr(T *vec) override { part(1, vec); }
    void MSolveUpperUsePtr(T *vec) override { part(2, vec); }

private:
    void check(int st) const {
        if (st != PP_OK) throw std::runtime_error(std::string("prispack_gpu: ") + pp_last_error(m_ctx));
    }
    static void check_global(int st) {
        if (st != PP_OK) throw std::runtime_error(std::string("prispack_gpu: ") + pp_global_error());
    }
    void require_ready() const {
        if (this->m_isReady != PRECONDITION_READY)
            throw std::logic_error("prispack_gpu: preconditioner used before setup()");
    }

    void apply(BaseVector<T> &in, UINT32 inStart, BaseVector<T> &out, UINT32 outStart) {
        require_ready();
        const uint32_t n = this->m_ArowNum;
        const size_t bytes = (size_t)n * sizeof(T);
        if ((uint64_t)inStart + n > in.getLength() || (uint64_t)outStart + n > out.getLength())
            throw std::out_of_range("prispack_gpu: MInvSolve slice exceeds the vector");
        T *src = in.getRawValPtr() + inStart, *dst = out.getRawValPtr() + outStart;
        const bool din = in.getLocation() >= 0, dout = out.getLocation() >= 0;
        const int dt = DeviceType<T>::value;
        if (din && dout) {
            check(pp_precond_apply(m_ctx, m_M, dt, src, dst));
            check(pp_ctx_sync(m_ctx));
            return;
        }
        check(din ? pp_memcpy_d2d(m_ctx, m_stage, src, bytes) : pp_memcpy_h2d(m_ctx, m_stage, src, bytes));
        check(pp_precond_apply(m_ctx, m_M, dt, m_stage, m_stage));
        check(dout ? pp_memcpy_d2d(m_ctx, dst, m_stage, bytes) : pp_memcpy_d2h(m_ctx, dst, m_stage, bytes));
        check(pp_ctx_sync(m_ctx));
    }

    void part(int which, T *hostvec) {
        require_ready();
        const size_t bytes = (size_t)this->m_ArowNum * sizeof(T);
        check(pp_memcpy_h2d(m_ctx, m_stage, hostvec, bytes));
        check(pp_precond_apply_part(m_ctx, m_M, which, DeviceType<T>::value, m_stage));
        check(pp_memcpy_d2h(m_ctx, hostvec, m_stage, bytes));
    }

    std::shared_ptr<HOST::CSRMatrix<T>> m_matA;  // shallow-held, as the reference holds its matrix
    pp_precond_opts m_opts;
    int m_device = 0;
    pp_ctx *m_ctx = nullptr;
    pp_csr *m_A = nullptr;
    pp_precond *m_M = nullptr;
    void *m_stage = nullptr;
};

inline pp_precond_opts make_opts(int kind, double tau, uint32_t lfil, int trsv, int ordering) {
    pp_precond_opts o;
    pp_precond_opts_default(&o);
    o.kind = kind;
    o.tau = tau;
    o.lfil = lfil;
    o.trsv = trsv;
    o.ordering = ordering;
    return o;
}

// IncompleteCholesky(matA, threshold): the reference's threshold IC factor,
// bit for bit, applied on the GPU.
template <typename T>
class IncompleteCholesky : public GPUTriangularPrecondition<T> {
public:
    IncompleteCholesky(const std::shared_ptr<HOST::CSRMatrix<T>> &matA, FLOAT64 threshold, int trsv = PP_TRSV_AUTO,
                       int device = 0)
        : GPUTriangularPrecondition<T>(matA, make_opts(PP_PC_ICT, threshold, 0, trsv, PP_ORDER_NATURAL),
                                       PreconditionIC, device) {}
};

// IncompleteLU(matA, lfill, tolerance): the reference's ILUT factor.
template <typename T>
class IncompleteLU : public GPUTriangularPrecondition<T> {
public:
    IncompleteLU(const std::shared_ptr<HOST::CSRMatrix<T>> &matA, UINT32 lfill, FLOAT64 tolerance,
                 int trsv = PP_TRSV_AUTO, int device = 0)
        : GPUTriangularPrecondition<T>(matA, make_opts(PP_PC_ILUT, tolerance, lfill, trsv, PP_ORDER_NATURAL),
                                       PreconditionILUT, device) {}
};

// IncompleteLDLT(matA, tolerance): the reference's threshold LDL^T factor.
template <typename T>
class IncompleteLDLT : public GPUTriangularPrecondition<T> {
public:
    IncompleteLDLT(const std::shared_ptr<HOST::CSRMatrix<T>> &matA, FLOAT64 tolerance, int trsv = PP_TRSV_AUTO,
                   int device = 0)
        : GPUTriangularPrecondition<T>(matA, make_opts(PP_PC_ILDLT, tolerance, 0, trsv, PP_ORDER_NATURAL),
                                       PreconditionILDLT, device) {}
};

}  // namespace prispack_gpu
