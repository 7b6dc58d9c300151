// prispack_amd.hpp -- C++ drop-in classes over the C ABI in prispack_amd.h.
//
// The class names, constructor arguments, virtual interfaces and getters
// mirror PrisPack's host classes so that code written against them compiles
// against these with a namespace change:
//   BaseVector<T>        include/VectorClass/BaseVector.h:29-126
//   HOST::CSRMatrix<T>   include/MatrixClass/CSRMatrix.h:19-230 (accessors)
//   BasePrecondition<T>  include/Preconditioner/BasePreconditon.h:34-145
//   TriangularPrecondition<T> (MSolveLower/Diagonal/Upper) BasePreconditon.h:200-303
//   BaseSolver<T>        include/Solver/BaseSolver.h:9-17
//   CG<T>                include/Solver/CG.h:35-58
//   GMRES<T>             include/Solver/GMRES.h:39-70
// Work runs on one AMD GPU through a pp_ctx (one device, one HIP stream).
// Vectors may live on the host (HostVector) or the device (DeviceVector):
// MInvSolve and solve() accept either, staging host data through the device.
// Errors throw prispack::Error (a std::runtime_error) carrying the C ABI's
// status and message; nothing calls exit().
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "prispack_amd.h"

namespace prispack {

enum PreconditionType {  // BasePreconditon.h:12-19
    PreconditionNone,
    PreconditionIC,
    PreconditionILUT,
    PreconditionILDLT,
    PreconditionGMSLR,
    PreconditionAMSED
};
typedef PreconditionType PreconditionType_t;

enum memoryType_t { memoryBase, memoryAligned, memoryPageLocked, memoryDeviceSync };  // BaseVector.h:19-24
enum CGType_t { NormalCG, FlexibleCG };
enum GMRESType_t { NormalGMRES, FlexibleGMRES };
enum Precision { PrecisionFP64 = PP_FP64, PrecisionMixed = PP_MIXED };

class Error : public std::runtime_error {
public:
    Error(int status, const std::string &msg) : std::runtime_error(msg), m_status(status) {}
    int status() const { return m_status; }

private:
    int m_status;
};

// One device and one HIP stream.  Every object below holds a shared_ptr to
// the context it was made on; calls on one context are serialised by the
// caller.
class Context {
public:
    explicit Context(int device = 0) {
        pp_ctx *h = nullptr;
        const int st = pp_ctx_create(device, &h);
        if (st != PP_OK) throw Error(st, std::string("pp_ctx_create: ") + pp_global_error());
        m_h = h;
        m_device = device;
    }
    ~Context() { pp_ctx_destroy(m_h); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;

    pp_ctx *handle() const { return m_h; }
    int device() const { return m_device; }
    void check(int st) const {
        if (st != PP_OK) throw Error(st, pp_last_error(m_h));
    }
    void synchronize() const { check(pp_ctx_sync(m_h)); }

    static std::shared_ptr<Context> defaultContext() {
        static std::shared_ptr<Context> ctx = std::make_shared<Context>(0);
        return ctx;
    }

private:
    pp_ctx *m_h = nullptr;
    int m_device = 0;
};

// ---------------------------------------------------------------- vectors
template <typename T>
class BaseVector {
public:
    virtual ~BaseVector() = default;
    virtual uint32_t getLength() const = 0;
    virtual T *getRawValPtr() = 0;
    virtual const T *getRawValPtr() const = 0;
    virtual memoryType_t getMemoryType() const = 0;
    virtual int getLocation() const = 0;  // -1 host, >= 0 device id
    virtual void resize(uint32_t length) = 0;
};

template <typename T>
class HostVector : public BaseVector<T> {
public:
    HostVector() = default;
    explicit HostVector(uint32_t length) : m_v(length, T(0)) {}
    explicit HostVector(std::vector<T> v) : m_v(std::move(v)) {}
    uint32_t getLength() const override { return (uint32_t)m_v.size(); }
    T *getRawValPtr() override { return m_v.data(); }
    const T *getRawValPtr() const override { return m_v.data(); }
    memoryType_t getMemoryType() const override { return memoryBase; }
    int getLocation() const override { return -1; }
    void resize(uint32_t length) override { m_v.assign(length, T(0)); }
    T &operator[](uint32_t i) { return m_v[i]; }
    const T &operator[](uint32_t i) const { return m_v[i]; }
    std::vector<T> &data() { return m_v; }
    const std::vector<T> &data() const { return m_v; }

private:
    std::vector<T> m_v;
};

// Device buffer (DeviceVector.cuh:17-251): copies in and out are
// stream-ordered and synchronous for the host.
template <typename T>
class DeviceVector : public BaseVector<T> {
public:
    explicit DeviceVector(uint32_t length = 0, std::shared_ptr<Context> ctx = Context::defaultContext())
        : m_ctx(std::move(ctx)) {
        resize(length);
    }
    DeviceVector(const HostVector<T> &h, std::shared_ptr<Context> ctx = Context::defaultContext())
        : m_ctx(std::move(ctx)) {
        resize(h.getLength());
        copyFromHost(h);
    }
    ~DeviceVector() override { release(); }
    DeviceVector(const DeviceVector &) = delete;
    DeviceVector &operator=(const DeviceVector &) = delete;
    DeviceVector(DeviceVector &&o) noexcept : m_ctx(o.m_ctx), m_p(o.m_p), m_n(o.m_n) {
        o.m_p = nullptr;
        o.m_n = 0;
    }

    uint32_t getLength() const override { return m_n; }
    T *getRawValPtr() override { return m_p; }
    const T *getRawValPtr() const override { return m_p; }
    memoryType_t getMemoryType() const override { return memoryDeviceSync; }
    int getLocation() const override { return m_ctx->device(); }
    void resize(uint32_t length) override {
        release();
        m_n = length;
        if (length) {
            void *p = nullptr;
            m_ctx->check(pp_malloc(m_ctx->handle(), sizeof(T) * length, &p));
            m_p = static_cast<T *>(p);
            m_ctx->check(pp_memset_zero(m_ctx->handle(), m_p, sizeof(T) * length));
        }
    }
    void copyFromHost(const BaseVector<T> &h) {
        if (h.getLength() != m_n) throw Error(PP_ERR_INVALID, "DeviceVector::copyFromHost: length mismatch");
        if (m_n) m_ctx->check(pp_memcpy_h2d(m_ctx->handle(), m_p, h.getRawValPtr(), sizeof(T) * m_n));
    }
    void copyToHost(BaseVector<T> &h) const {
        if (h.getLength() != m_n) h.resize(m_n);
        if (m_n) m_ctx->check(pp_memcpy_d2h(m_ctx->handle(), h.getRawValPtr(), m_p, sizeof(T) * m_n));
    }
    const std::shared_ptr<Context> &context() const { return m_ctx; }

private:
    void release() {
        if (m_p) pp_free(m_ctx->handle(), m_p);
        m_p = nullptr;
        m_n = 0;
    }
    std::shared_ptr<Context> m_ctx;
    T *m_p = nullptr;
    uint32_t m_n = 0;
};

// ---------------------------------------------------------------- matrices
// Host CSR with the reference's accessors (0-based, sorted columns).
template <typename T>
class CSRMatrix {
public:
    CSRMatrix() = default;
    CSRMatrix(uint32_t nrow, uint32_t ncol, std::vector<uint32_t> rowOffset, std::vector<uint32_t> colIndices,
              std::vector<T> values)
        : m_nrow(nrow), m_ncol(ncol), m_rp(std::move(rowOffset)), m_ci(std::move(colIndices)),
          m_v(std::move(values)) {
        if (m_rp.size() != (size_t)nrow + 1 || m_ci.size() != m_rp.back() || m_v.size() != m_rp.back())
            throw Error(PP_ERR_INVALID, "CSRMatrix: array sizes do not match rowOffset");
    }
    uint32_t getRowNum() const { return m_nrow; }
    uint32_t getColNum() const { return m_ncol; }
    uint32_t getNNZnum() const { return m_rp.empty() ? 0 : m_rp.back(); }
    const uint32_t *getRowOffsetPtr() const { return m_rp.data(); }
    const uint32_t *getColIndicesPtr() const { return m_ci.data(); }
    const T *getCSRValuesPtr() const { return m_v.data(); }
    T *getCSRValuesPtr() { return m_v.data(); }
    // value of entry (i, j), 0 if not stored (BaseMatrix::getValue)
    T getValue(uint32_t i, uint32_t j) const {
        for (uint32_t k = m_rp[i]; k < m_rp[i + 1]; ++k)
            if (m_ci[k] == j) return m_v[k];
        return T(0);
    }

private:
    uint32_t m_nrow = 0, m_ncol = 0;
    std::vector<uint32_t> m_rp, m_ci;
    std::vector<T> m_v;
};

// generateLaplacianCSR (ModelProblem.cpp:68-133): 5-point (nz == 1) or
// 7-point operator with centered convection and a diagonal shift
inline CSRMatrix<double> generateLaplacianCSR(int nx, int ny, int nz, double ax = 0.0, double ay = 0.0,
                                              double az = 0.0, double shift = 0.0) {
    uint32_t n = 0, nnz = 0;
    int st = pp_generate_laplacian(nx, ny, nz, ax, ay, az, shift, &n, &nnz, nullptr, nullptr, nullptr);
    if (st != PP_OK) throw Error(st, pp_global_error());
    std::vector<uint32_t> rp(n + 1), ci(nnz);
    std::vector<double> v(nnz);
    st = pp_generate_laplacian(nx, ny, nz, ax, ay, az, shift, &n, &nnz, rp.data(), ci.data(), v.data());
    if (st != PP_OK) throw Error(st, pp_global_error());
    return CSRMatrix<double>(n, n, std::move(rp), std::move(ci), std::move(v));
}

// DEVICE::DeviceCSRMatrix (DeviceCSRMatrix.cuh:19-93): FP64 values on the
// device (an FP32 copy is made on first mixed-precision use).
class DeviceCSRMatrix {
public:
    explicit DeviceCSRMatrix(const CSRMatrix<double> &A, std::shared_ptr<Context> ctx = Context::defaultContext())
        : m_ctx(std::move(ctx)) {
        pp_csr *h = nullptr;
        m_ctx->check(pp_csr_create(m_ctx->handle(), A.getRowNum(), A.getColNum(), A.getRowOffsetPtr(),
                                   A.getColIndicesPtr(), A.getCSRValuesPtr(), &h));
        m_h = h;
    }
    ~DeviceCSRMatrix() { pp_csr_destroy(m_h); }
    DeviceCSRMatrix(const DeviceCSRMatrix &) = delete;
    DeviceCSRMatrix &operator=(const DeviceCSRMatrix &) = delete;

    uint32_t getRowNum() const { return dims()[0]; }
    uint32_t getColNum() const { return dims()[1]; }
    uint32_t getNNZnum() const { return dims()[2]; }
    pp_csr *handle() const { return m_h; }
    const std::shared_ptr<Context> &context() const { return m_ctx; }
    // new values on the same pattern (host or device array)
    void setValues(const double *values, bool onDevice) {
        m_ctx->check(pp_csr_set_values(m_ctx->handle(), m_h, values, onDevice ? 1 : 0));
    }
    // y = A x (CSRMatrix::MatPVec, CSRMatrix.h:360-395)
    void MatPVec(const DeviceVector<double> &x, DeviceVector<double> &y) const {
        MatPVec(1.0, x, 0.0, y);
    }
    // y = alpha A x + beta y (CSRMatrix.h:423-447)
    void MatPVec(double alpha, const DeviceVector<double> &x, double beta, DeviceVector<double> &y) const {
        if (x.getLength() != getColNum() || y.getLength() != getRowNum())
            throw Error(PP_ERR_INVALID, "MatPVec: vector lengths do not match the matrix");
        m_ctx->check(pp_spmv(m_ctx->handle(), m_h, PP_F64, alpha, x.getRawValPtr(), beta, y.getRawValPtr()));
    }
    // y = A^T x (CSRMatrix::transMatPVec, CSRMatrix.h:453-486)
    void transMatPVec(const DeviceVector<double> &x, DeviceVector<double> &y) const {
        if (x.getLength() != getRowNum() || y.getLength() != getColNum())
            throw Error(PP_ERR_INVALID, "transMatPVec: vector lengths do not match the matrix");
        m_ctx->check(pp_spmv_t(m_ctx->handle(), m_h, PP_F64, 1.0, x.getRawValPtr(), 0.0, y.getRawValPtr()));
    }

private:
    std::vector<uint32_t> dims() const {
        uint32_t r = 0, c = 0, z = 0;
        m_ctx->check(pp_csr_dims(m_h, &r, &c, &z));
        return {r, c, z};
    }
    std::shared_ptr<Context> m_ctx;
    pp_csr *m_h = nullptr;
};

// ---------------------------------------------------------------- preconditioners
template <typename T>
class BasePrecondition {
public:
    virtual ~BasePrecondition() = default;
    virtual void setup() = 0;
    virtual uint8_t isReady() const = 0;
    // vec = M^{-1} vec
    virtual void MInvSolve(BaseVector<T> &vec) = 0;
    // vec[start : start + n) = M^{-1} vec[start : start + n)
    virtual void MInvSolve(BaseVector<T> &vec, uint32_t resStartIdx) = 0;
    // out = M^{-1} in (out resized to n)
    virtual void MInvSolve(BaseVector<T> &vecIN, BaseVector<T> &vecOUT) = 0;
    // out[outStart : +n) = M^{-1} in[inStart : +n)
    virtual void MInvSolve(BaseVector<T> &vecIN, uint32_t inStartIdx, BaseVector<T> &vecOUT,
                           uint32_t outStartIdx) = 0;
    virtual const PreconditionType_t &getPreconditionType() const = 0;
    virtual uint32_t getPreconditionNonZeroCount() const = 0;
    virtual uint32_t getOriginalNonZeroCount() const = 0;
    double getPreconditionFillinRatio() const {
        return getPreconditionNonZeroCount() / static_cast<double>(getOriginalNonZeroCount());
    }
};

// A preconditioner built and applied on the device.  Host vectors are staged
// through a device buffer of length n.
class DevicePrecondition : public BasePrecondition<double> {
public:
    ~DevicePrecondition() override { pp_precond_destroy(m_h); }
    DevicePrecondition(const DevicePrecondition &) = delete;
    DevicePrecondition &operator=(const DevicePrecondition &) = delete;

    void setup() override { m_ctx->check(pp_precond_setup(m_ctx->handle(), m_h)); }
    // numeric re-factorization from the matrix's current values (pp_precond_refactor)
    void refactor() { m_ctx->check(pp_precond_refactor(m_ctx->handle(), m_h)); }
    uint8_t isReady() const override { return (uint8_t)info().ready; }
    void MInvSolve(BaseVector<double> &vec) override { MInvSolve(vec, 0, vec, 0); }
    void MInvSolve(BaseVector<double> &vec, uint32_t resStartIdx) override {
        MInvSolve(vec, resStartIdx, vec, resStartIdx);
    }
    void MInvSolve(BaseVector<double> &vecIN, BaseVector<double> &vecOUT) override {
        if (&vecIN != &vecOUT && vecOUT.getLength() != n()) vecOUT.resize(n());
        MInvSolve(vecIN, 0, vecOUT, 0);
    }
    void MInvSolve(BaseVector<double> &vecIN, uint32_t inStartIdx, BaseVector<double> &vecOUT,
                   uint32_t outStartIdx) override {
        const uint32_t len = n();
        if ((uint64_t)inStartIdx + len > vecIN.getLength() || (uint64_t)outStartIdx + len > vecOUT.getLength())
            throw Error(PP_ERR_INVALID, "MInvSolve: the slice exceeds the vector");
        const bool din = vecIN.getLocation() >= 0, dout = vecOUT.getLocation() >= 0;
        if (din && dout) {
            m_ctx->check(pp_precond_apply(m_ctx->handle(), m_h, PP_F64, vecIN.getRawValPtr() + inStartIdx,
                                          vecOUT.getRawValPtr() + outStartIdx));
            return;
        }
        double *s = stage();
        if (din)
            m_ctx->check(pp_memcpy_d2d(m_ctx->handle(), s, vecIN.getRawValPtr() + inStartIdx, 8ull * len));
        else
            m_ctx->check(pp_memcpy_h2d(m_ctx->handle(), s, vecIN.getRawValPtr() + inStartIdx, 8ull * len));
        m_ctx->check(pp_precond_apply(m_ctx->handle(), m_h, PP_F64, s, s));
        if (dout)
            m_ctx->check(pp_memcpy_d2d(m_ctx->handle(), vecOUT.getRawValPtr() + outStartIdx, s, 8ull * len));
        else
            m_ctx->check(pp_memcpy_d2h(m_ctx->handle(), vecOUT.getRawValPtr() + outStartIdx, s, 8ull * len));
    }
    const PreconditionType_t &getPreconditionType() const override { return m_type; }
    uint32_t getPreconditionNonZeroCount() const override { return info().mnnz; }
    uint32_t getOriginalNonZeroCount() const override { return info().annz; }

    pp_precond *handle() const { return m_h; }
    const std::shared_ptr<Context> &context() const { return m_ctx; }
    pp_precond_info info() const {
        pp_precond_info i{};
        m_ctx->check(pp_precond_info_get(m_h, &i));
        return i;
    }

protected:
    DevicePrecondition(const std::shared_ptr<DeviceCSRMatrix> &A, const pp_precond_opts &o, PreconditionType_t t)
        : m_A(A), m_ctx(A->context()), m_type(t) {
        pp_precond *h = nullptr;
        m_ctx->check(pp_precond_create(m_ctx->handle(), A->handle(), &o, &h));
        m_h = h;
    }
    static pp_precond_opts opts(int kind) {
        pp_precond_opts o;
        pp_precond_opts_default(&o);
        o.kind = kind;
        return o;
    }
    uint32_t n() const { return m_A->getRowNum(); }
    double *stage() {
        if (!m_stage) m_stage.reset(new DeviceVector<double>(n(), m_ctx));
        return m_stage->getRawValPtr();
    }

    std::shared_ptr<DeviceCSRMatrix> m_A;  // shallow-held, as the reference holds its matrix
    std::shared_ptr<Context> m_ctx;
    pp_precond *m_h = nullptr;
    PreconditionType_t m_type;
    std::unique_ptr<DeviceVector<double>> m_stage;
};

// TriangularPrecondition: the three parts of M^{-1} = U^{-1} D^{-1} L^{-1}
class DeviceTriangularPrecondition : public DevicePrecondition {
public:
    void MSolveLower(BaseVector<double> &vec) { part(0, vec); }
    void MSolveDiagonal(BaseVector<double> &vec) { part(1, vec); }
    void MSolveUpper(BaseVector<double> &vec) { part(2, vec); }

protected:
    using DevicePrecondition::DevicePrecondition;
    void part(int which, BaseVector<double> &vec) {
        if (vec.getLength() < n()) throw Error(PP_ERR_INVALID, "MSolve: vector shorter than the factor");
        if (vec.getLocation() >= 0) {
            m_ctx->check(pp_precond_apply_part(m_ctx->handle(), m_h, which, PP_F64, vec.getRawValPtr()));
            return;
        }
        double *s = stage();
        m_ctx->check(pp_memcpy_h2d(m_ctx->handle(), s, vec.getRawValPtr(), 8ull * n()));
        m_ctx->check(pp_precond_apply_part(m_ctx->handle(), m_h, which, PP_F64, s));
        m_ctx->check(pp_memcpy_d2h(m_ctx->handle(), vec.getRawValPtr(), s, 8ull * n()));
    }
};

// Sweep schedule and row ordering shared by the triangular factors.
struct TriangularOptions {
    int trsv = PP_TRSV_AUTO;         // PP_TRSV_LEVELS / PP_TRSV_SYNCFREE
    int ordering = PP_ORDER_NATURAL;  // PP_ORDER_MULTICOLOR for IC(0)/ILU(0)
    bool keepFP32 = false;            // also keep an FP32 factor (mixed solves)
};

// IncompleteCholesky<Low,High>(A, threshold) (IncompleteCholesky.h:22-60)
class DeviceIncompleteCholesky : public DeviceTriangularPrecondition {
public:
    DeviceIncompleteCholesky(const std::shared_ptr<DeviceCSRMatrix> &A, double threshold,
                             TriangularOptions t = TriangularOptions())
        : DeviceTriangularPrecondition(A, make(PP_PC_ICT, threshold, 0, t), PreconditionIC) {}

protected:
    DeviceIncompleteCholesky(const std::shared_ptr<DeviceCSRMatrix> &A, const pp_precond_opts &o)
        : DeviceTriangularPrecondition(A, o, PreconditionIC) {}
    static pp_precond_opts make(int kind, double tau, uint32_t lfil, const TriangularOptions &t) {
        pp_precond_opts o = opts(kind);
        o.tau = tau;
        if (lfil) o.lfil = lfil;
        o.trsv = t.trsv;
        o.ordering = t.ordering;
        o.make_f32 = t.keepFP32 ? 1 : 0;
        return o;
    }
    friend class DeviceIncompleteLU;
    friend class DeviceIncompleteLDLT;
    friend class DeviceIC0;
    friend class DeviceILU0;
};

// IncompleteLU<Low,High>(A, lfill, tolerance) (IncompleteLU.h)
class DeviceIncompleteLU : public DeviceTriangularPrecondition {
public:
    DeviceIncompleteLU(const std::shared_ptr<DeviceCSRMatrix> &A, uint32_t lfill, double tolerance,
                       TriangularOptions t = TriangularOptions())
        : DeviceTriangularPrecondition(A, DeviceIncompleteCholesky::make(PP_PC_ILUT, tolerance, lfill, t),
                                       PreconditionILUT) {}
};

// IncompleteLDLT<Low,High>(A, tolerance) (IncompleteLDLT.h)
class DeviceIncompleteLDLT : public DeviceTriangularPrecondition {
public:
    DeviceIncompleteLDLT(const std::shared_ptr<DeviceCSRMatrix> &A, double tolerance,
                         TriangularOptions t = TriangularOptions())
        : DeviceTriangularPrecondition(A, DeviceIncompleteCholesky::make(PP_PC_ILDLT, tolerance, 0, t),
                                       PreconditionILDLT) {}
};

// zero-fill IC(0) and ILU(0) (no reference counterpart; reported as IC / ILUT)
class DeviceIC0 : public DeviceTriangularPrecondition {
public:
    explicit DeviceIC0(const std::shared_ptr<DeviceCSRMatrix> &A, TriangularOptions t = TriangularOptions())
        : DeviceTriangularPrecondition(A, DeviceIncompleteCholesky::make(PP_PC_IC0, 0.0, 0, t), PreconditionIC) {}
};
class DeviceILU0 : public DeviceTriangularPrecondition {
public:
    explicit DeviceILU0(const std::shared_ptr<DeviceCSRMatrix> &A, TriangularOptions t = TriangularOptions())
        : DeviceTriangularPrecondition(A, DeviceIncompleteCholesky::make(PP_PC_ILU0, 0.0, 0, t),
                                       PreconditionILUT) {}
};

// Multilevel low-rank Schur preconditioners (AMSEDPrecondition.h:75-533,
// GMSLRPrecondition.h:48-399): nlev levels of a hierarchical interface
// decomposition into ndom leaf subdomains, rank-`rank` correction per level.
struct MultilevelOptions {
    uint32_t levels = 3, ndom = 16, rank = 32;
    double tolerance = -1.0;       // block factor drop tolerance (< 0: 1e-3 AMSED, 1e-2 GMSLR)
    uint32_t lfill = 20;           // GMSLR ILUT fill per row
    uint32_t krylovSteps = 0;      // Lanczos / Arnoldi steps (0: the low-rank form's default)
    int lowrank = PP_LOWRANK_AUTO; // PP_LOWRANK_*: the reference's restarted Lanczos / Schur selection by default
    double eps = 0.8;              // AMSED deflation bound (PP_LOWRANK_DEFLATION)
};

class DeviceAMSED : public DevicePrecondition {
public:
    DeviceAMSED(const std::shared_ptr<DeviceCSRMatrix> &A, MultilevelOptions m = MultilevelOptions())
        : DevicePrecondition(A, make(PP_PC_AMSED, m, 1e-3), PreconditionAMSED) {}

protected:
    DeviceAMSED(const std::shared_ptr<DeviceCSRMatrix> &A, const pp_precond_opts &o, PreconditionType_t t)
        : DevicePrecondition(A, o, t) {}
    static pp_precond_opts make(int kind, const MultilevelOptions &m, double tau) {
        pp_precond_opts o = opts(kind);
        o.levels = m.levels;
        o.ndom = m.ndom;
        o.rank = m.rank;
        o.tau = m.tolerance >= 0.0 ? m.tolerance : tau;
        o.lfil = m.lfill;
        o.lanczos_steps = m.krylovSteps;
        o.lowrank = m.lowrank;
        o.eps = m.eps;
        return o;
    }
};

class DeviceGMSLR : public DeviceAMSED {
public:
    DeviceGMSLR(const std::shared_ptr<DeviceCSRMatrix> &A, MultilevelOptions m = MultilevelOptions())
        : DeviceAMSED(A, make(PP_PC_GMSLR, m, 1e-2), PreconditionGMSLR) {}
};

// ---------------------------------------------------------------- solvers
template <typename T>
class BaseSolver {
public:
    virtual ~BaseSolver() = default;
    virtual void solve(const CSRMatrix<T> &matA, const HostVector<T> &b, HostVector<T> &x) = 0;
};

// Shared state of DeviceCG / DeviceGMRES.  solve(CSRMatrix, HostVector,
// HostVector) is the drop-in (host in, host out); the DeviceCSRMatrix /
// DeviceVector overload keeps everything resident.
class DeviceSolver : public BaseSolver<double> {
public:
    void solve(const CSRMatrix<double> &matA, const HostVector<double> &b, HostVector<double> &x) override {
        if (b.getLength() != matA.getRowNum()) throw Error(PP_ERR_INVALID, "solve: b does not match A");
        if (x.getLength() != matA.getRowNum()) x.resize(matA.getRowNum());
        std::shared_ptr<DeviceCSRMatrix> dA = deviceMatrix(matA);
        run(*dA, [&](pp_ctx *c, pp_csr *h, pp_precond *M, pp_solve_report *r) {
            return pp_solve_host(c, h, M, &m_opts, b.getRawValPtr(), x.getRawValPtr(), r);
        });
    }
    void solve(const DeviceCSRMatrix &A, const DeviceVector<double> &b, DeviceVector<double> &x) {
        if (b.getLength() != A.getRowNum() || x.getLength() != A.getRowNum())
            throw Error(PP_ERR_INVALID, "solve: vector lengths do not match A");
        run(A, [&](pp_ctx *c, pp_csr *h, pp_precond *M, pp_solve_report *r) {
            return pp_solve(c, h, M, &m_opts, b.getRawValPtr(), x.getRawValPtr(), r);
        });
    }
    int32_t getNumIter() const { return m_rep.iters; }
    double getError() const { return m_rep.final_error; }
    bool getConvergence() const { return m_rep.converged != 0; }
    const pp_solve_report &report() const { return m_rep; }
    PreconditionType_t getPrecondtionType() const {
        return m_precond ? m_precond->getPreconditionType() : PreconditionNone;
    }
    void setPrecision(Precision p) { m_opts.precision = p; }
    pp_solve_opts &options() { return m_opts; }

protected:
    DeviceSolver(uint32_t n, std::shared_ptr<DevicePrecondition> precond, int solver, bool flexible, double tol,
                 uint32_t maxIter)
        : m_n(n), m_precond(std::move(precond)) {
        pp_solve_opts_default(&m_opts, solver);
        m_opts.flexible = flexible ? 1 : 0;
        m_opts.tol = tol;
        m_opts.maxit = maxIter;
    }
    template <typename F>
    void run(const DeviceCSRMatrix &A, F call) {
        if (A.getRowNum() != m_n) throw Error(PP_ERR_INVALID, "solve: matrix size differs from problemSize");
        pp_precond *M = nullptr;
        if (m_precond) {
            if (!m_precond->isReady()) m_precond->setup();
            M = m_precond->handle();
        }
        m_rep = pp_solve_report{};
        const auto &ctx = A.context();
        ctx->check(call(ctx->handle(), A.handle(), M, &m_rep));
    }
    // the preconditioner's device matrix when it holds matA's values (the
    // usual case: built from the same A), else a fresh upload
    std::shared_ptr<DeviceCSRMatrix> deviceMatrix(const CSRMatrix<double> &matA) {
        if (m_cached && m_cachedSrc == &matA) return m_cached;
        m_cached = std::make_shared<DeviceCSRMatrix>(matA);
        m_cachedSrc = &matA;
        return m_cached;
    }

    uint32_t m_n;
    std::shared_ptr<DevicePrecondition> m_precond;
    pp_solve_opts m_opts{};
    pp_solve_report m_rep{};
    std::shared_ptr<DeviceCSRMatrix> m_cached;
    const CSRMatrix<double> *m_cachedSrc = nullptr;
};

// CG<T>(problemSize, tol, maxIter) / (problemSize, precond, CGType, tol, maxIter)
class DeviceCG : public DeviceSolver {
public:
    DeviceCG(uint32_t problemSize, double tol, uint32_t maxIter)
        : DeviceSolver(problemSize, nullptr, PP_SOLVER_CG, true, tol, maxIter) {}
    DeviceCG(uint32_t problemSize, const std::shared_ptr<DevicePrecondition> &precond, CGType_t type, double tol,
             uint32_t maxIter)
        : DeviceSolver(problemSize, precond, PP_SOLVER_CG, type == FlexibleCG, tol, maxIter) {}
};

// GMRES<T>(problemSize[, precond, [tolerance,] GMRESType]) + setRestart.
// Without setRestart the restart length is n - 1, as in the reference.
class DeviceGMRES : public DeviceSolver {
public:
    explicit DeviceGMRES(uint32_t problemSize)
        : DeviceSolver(problemSize, nullptr, PP_SOLVER_GMRES, true, 1e-7, 500) {
        m_opts.restart = problemSize > 1 ? problemSize - 1 : 1;
    }
    DeviceGMRES(uint32_t problemSize, const std::shared_ptr<DevicePrecondition> &precond, GMRESType_t type)
        : DeviceGMRES(problemSize, precond, 1e-7, type) {}
    DeviceGMRES(uint32_t problemSize, const std::shared_ptr<DevicePrecondition> &precond, double tolerance,
                GMRESType_t type)
        : DeviceSolver(problemSize, precond, PP_SOLVER_GMRES, type == FlexibleGMRES, tolerance, 500) {
        m_opts.restart = problemSize > 1 ? problemSize - 1 : 1;
    }
    void setRestart(bool restart, uint32_t m) {
        m_opts.restart = restart ? m : (m_n > 1 ? m_n - 1 : 1);
    }
    // reproduce the reference's converged-cycle back-substitution quirk
    void setReferenceQuirks(bool on) { m_opts.reference_quirks = on ? 1 : 0; }
};

}  // namespace prispack
