/*
 * prispack_amd.h -- C ABI of the MI355X (gfx950) preconditioned-Krylov hot path.
 *
 * Plain pointers and sizes only.  Vectors are device pointers (hipMalloc'd by
 * the caller, by pp_malloc, or borrowed from a framework such as torch) with
 * an element type PP_F64 / PP_F32.  Matrices use 32-bit unsigned CSR indices,
 * 0-based, rows sorted by column, FP64 values (an FP32 copy is made on demand
 * for mixed precision).  All work is ordered on the context's HIP stream.
 * Every entry point returns a status (PP_OK = 0); pp_last_error(ctx) gives
 * the message.  Nothing here calls exit().
 *
 * Each entry point names the reference interface it replaces (PrisPack,
 * file:line).
 */
#ifndef PRISPACK_AMD_H
#define PRISPACK_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PP_ABI_VERSION 1

/* ---- status codes (replace config/debug.h:96-135 exit()/THROW paths) ---- */
enum {
    PP_OK = 0,
    PP_ERR_INVALID = 1,      /* bad argument / shape mismatch                    */
    PP_ERR_HIP = 2,          /* HIP runtime error                                 */
    PP_ERR_NOT_SPD = 3,      /* IC pivot <= 0 (IncompleteCholesky.cpp:189-193)    */
    PP_ERR_ZERO_PIVOT = 4,   /* ILUT/ILDLT zero pivot (IncompleteLU.cpp:199,
                                IncompleteLDLT.cpp:131)                           */
    PP_ERR_ZERO_ROW = 5,     /* ILUT zero row (IncompleteLU.cpp:89)               */
    PP_ERR_NOT_READY = 6,    /* preconditioner used before setup                  */
    PP_ERR_BREAKDOWN = 7,    /* Krylov breakdown (zero/NaN denominator)           */
    PP_ERR_NOT_FINITE = 8,   /* NaN/Inf in the right-hand side or the iterates    */
    PP_ERR_NOMEM = 9,        /* device allocation failed                          */
    PP_ERR_UNSUPPORTED = 10
};

enum { PP_F64 = 0, PP_F32 = 1 };

typedef struct pp_ctx pp_ctx;
typedef struct pp_csr pp_csr;
typedef struct pp_precond pp_precond;

/* ---- context: one per device + stream (replaces the per-call cudaSetDevice
 *      and StreamController of include/CUDA/StreamController.cuh:19-77) ---- */
int pp_ctx_create(int device, pp_ctx **out);
/* Borrow an existing hipStream_t (e.g. torch.cuda.current_stream().cuda_stream).
 * NULL restores the context's own stream. */
int pp_ctx_set_stream(pp_ctx *ctx, void *hip_stream);
void *pp_ctx_stream(pp_ctx *ctx);
int pp_ctx_sync(pp_ctx *ctx);
void pp_ctx_destroy(pp_ctx *ctx);
const char *pp_last_error(const pp_ctx *ctx);
/* message of the last failure that had no context (e.g. pp_ctx_create) */
const char *pp_global_error(void);
int pp_device_count(int *count);

/* ---- device memory (DeviceVector / SyncDeviceVector storage,
 *      include/VectorClass/DeviceVector.cuh:17-251) ---- */
int pp_malloc(pp_ctx *ctx, size_t bytes, void **dptr);
int pp_free(pp_ctx *ctx, void *dptr);
int pp_memcpy_h2d(pp_ctx *ctx, void *dst, const void *src, size_t bytes);
int pp_memcpy_d2h(pp_ctx *ctx, void *dst, const void *src, size_t bytes);
int pp_memcpy_d2d(pp_ctx *ctx, void *dst, const void *src, size_t bytes);
int pp_memset_zero(pp_ctx *ctx, void *dst, size_t bytes);   /* fill kernel, graph-safe */
/* dst = src by a streaming kernel (16-byte non-temporal loads and stores;
 * 16-byte aligned device pointers; groups 0 = 64 workgroups per CU).  No
 * reference counterpart: the rate a kernel that only moves bytes reaches,
 * which the SpMV roofline is compared with. */
int pp_copy_stream(pp_ctx *ctx, const void *src, void *dst, size_t bytes, uint32_t groups);

/* ---- device CSR (DEVICE::DeviceCSRMatrix, include/MatrixClass/DeviceCSRMatrix.cuh:19-93;
 *      host storage HOST::CSRMatrix, include/MatrixClass/CSRMatrix.h:19-31) ---- */
/* Host arrays are copied to the device. */
int pp_csr_create(pp_ctx *ctx, uint32_t nrow, uint32_t ncol, const uint32_t *rowptr,
                  const uint32_t *colidx, const double *values, pp_csr **out);
/* Replace the values (same pattern).  on_device != 0: values is a device pointer. */
int pp_csr_set_values(pp_ctx *ctx, pp_csr *A, const double *values, int on_device);
int pp_csr_dims(const pp_csr *A, uint32_t *nrow, uint32_t *ncol, uint32_t *nnz);
/* Bytes one y = A x (beta = 0) streams through HBM in the library's storage:
 * values, column indices (16-bit block-relative offsets where a row block's
 * column span fits, else 32-bit), row pointers, row-block table, x and y.
 * The algorithmic traffic the SpMV roofline is priced on.  No reference
 * counterpart. */
int pp_csr_spmv_bytes(const pp_csr *A, int dtype, uint64_t *bytes);
/* SpMV load variant of one precision (dtype PP_F64 / PP_F32), chosen by
 * timing on large matrices (FP64 when the matrix is created, FP32 when its
 * FP32 copy is made): 0-7 the CSR-stream kernel with bit 0 non-temporal
 * operand loads and output stores, bit 1 16-bit column offsets, bit 2
 * batched loads; 8 the sliced-column copy (one row per thread, entry j of
 * all rows contiguous, 16-bit column offsets from the row), which large
 * matrices with at most 16 entries per row, little padding and near-diagonal
 * columns also hold.  Every variant gives the same result bit for bit. */
int pp_csr_spmv_variant(const pp_csr *A, int dtype, int *variant);
/* Pins the SpMV load variant of one precision instead of the timed choice
 * (reproducible profiling, identical copies of one matrix).  PP_ERR_INVALID
 * for a value outside 0..8, for 16-bit offsets on a matrix that has none, or
 * for 8 on a matrix without the sliced-column copy. */
int pp_csr_set_spmv_variant(pp_csr *A, int dtype, int variant);
/* Device pointers of the stored arrays (values64 FP64; values32 NULL until an
 * FP32 copy exists). */
int pp_csr_device_ptrs(const pp_csr *A, const uint32_t **rowptr, const uint32_t **colidx,
                       const double **values64, const float **values32);
void pp_csr_destroy(pp_csr *A);

/* ---- SpMV (CSRMatrix::MatPVec / MatPVec(alpha,..,beta), CSRMatrix.h:360-447) ----
 * y = alpha * A x + beta * y  (beta == 0 never reads y).  dtype of x, y. */
int pp_spmv(pp_ctx *ctx, const pp_csr *A, int dtype, double alpha, const void *x,
            double beta, void *y);
/* y = alpha * A^T x + beta * y   (CSRMatrix::transMatPVec, CSRMatrix.h:453-486;
 * gather over an explicit transpose built on first use, no atomics). */
int pp_spmv_t(pp_ctx *ctx, pp_csr *A, int dtype, double alpha, const void *x, double beta,
              void *y);

/* ---- BLAS-1 (HostVector::add/scale/innerProduct/norm_2,
 *      include/VectorClass/BaseVector.h:279-530).  Dots accumulate in FP64 for
 *      PP_F64 and FP32 for PP_F32; reductions are deterministic (fixed order). */
int pp_dot(pp_ctx *ctx, int dtype, uint32_t n, const void *x, const void *y, double *result);
int pp_nrm2(pp_ctx *ctx, int dtype, uint32_t n, const void *x, double *result);
/* z = a x + b y (z may alias x or y) */
int pp_axpby(pp_ctx *ctx, int dtype, uint32_t n, double a, const void *x, double b,
             const void *y, void *z);
int pp_scale(pp_ctx *ctx, int dtype, uint32_t n, double a, void *x);
/* dtype conversion f64 <-> f32 (dst_dtype != src_dtype) */
int pp_convert(pp_ctx *ctx, uint32_t n, int src_dtype, const void *src, int dst_dtype, void *dst);

/* ---- preconditioners (BasePrecondition / TriangularPrecondition,
 *      include/Preconditioner/BasePreconditon.h:34-359) ---- */
enum {
    PP_PC_NONE = 0,
    PP_PC_ICT = 1,     /* IncompleteCholesky(A, tau)        IncompleteCholesky.cpp:96-295 */
    PP_PC_ILUT = 2,    /* IncompleteLU(A, lfil, tau)        IncompleteLU.cpp:38-297 */
    PP_PC_ILDLT = 3,   /* IncompleteLDLT(A, tau)            IncompleteLDLT.cpp:40-233 */
    PP_PC_IC0 = 4,     /* zero-fill IC  (no reference counterpart) */
    PP_PC_ILU0 = 5,    /* zero-fill ILU (no reference counterpart) */
    PP_PC_AMSED = 6,   /* AMSEDPrecondition                  AMSEDPrecondition.cpp */
    PP_PC_GMSLR = 7    /* GMSLRPrecondition                  GMSLRDPrecondition.cpp */
};

/* SpTRSV schedule */
enum {
    PP_TRSV_AUTO = 0,
    PP_TRSV_LEVELS = 1,    /* one launch per level (graph-captured)               */
    PP_TRSV_SYNCFREE = 2,  /* one launch per sweep, per-row ready flags            */
    PP_TRSV_BLOCK = 3,     /* both sweeps in one workgroup, the vector in LDS, a
                              barrier per level (n * 8 bytes must fit in LDS;
                              larger systems use LEVELS).  AUTO times it against
                              LEVELS at setup when it fits                         */
    PP_TRSV_PIPE = 4       /* BLOCK's storage; one wavefront computes the levels
                              with no barrier while the workgroup's other waves
                              stage the level data in LDS ahead of it (for many
                              narrow levels).  AUTO times it with the others      */
};

/* Row ordering of the factor (no reference counterpart; the reference factors
 * in the given order).  MULTICOLOR factors P A P^T with rows grouped by a
 * greedy distance-1 colouring, so IC(0)/ILU(0) sweeps have at most as many
 * dependency levels as colours (2 for 5/7-point stencils).  A solve on the
 * matrix the factor was built from runs entirely in the colour order (P A P^T
 * is formed once); a direct apply permutes in and out.  It changes the
 * preconditioner (and the iteration count). */
enum { PP_ORDER_NATURAL = 0, PP_ORDER_MULTICOLOR = 1, PP_ORDER_RCM = 2 };
/* PP_ORDER_RCM factors P A P^T with P the reverse Cuthill-McKee order of A's
 * pattern (pp_reorder_rcm; the ordering HOST::rcmReorderCSR computes,
 * include/utils/ExternalTools/MatrixReorderTools.h:59-80): a narrower band and
 * less fill for the threshold factors; solves run in that order like
 * MULTICOLOR's. */

typedef struct {
    int kind;            /* PP_PC_* */
    double tau;          /* drop tolerance (ICT/ILUT/ILDLT) */
    uint32_t lfil;       /* ILUT fill per row */
    int trsv;            /* PP_TRSV_* */
    int make_f32;        /* also keep an FP32 copy of the factor (mixed precision) */
    int ordering;        /* PP_ORDER_* */
    /* AMSED/GMSLR */
    uint32_t levels;     /* multilevel depth */
    uint32_t ndom;       /* subdomains at the first level */
    uint32_t rank;       /* low-rank correction rank */
    double eps;          /* AMSED deflation bound (m_eigValCorrectBound) */
    uint32_t lanczos_steps; /* AMSED/GMSLR Krylov steps (0: the form's default; the
                               deflated pairs at most for PP_LOWRANK_DEFLATION,
                               AMSED_MSTEPS = 30 by default) */
    int32_t lowrank;     /* PP_LOWRANK_* */
    uint32_t reserved[6];
} pp_precond_opts;

/* Low-rank correction form of AMSED / GMSLR:
 * REFERENCE (AUTO) -- AMSED: thick-restart (Krylov-Schur) generalised Lanczos
 *   with m = 2 rank, k = 4 rank steps and at most three restarts, H = mu/(1-mu)
 *   on the rank largest Ritz values (AMSEDBasic, AMSEDPrecondition.cpp:
 *   1010-1231); GMSLR: Arnoldi, the first rank entries of the real Schur form
 *   in its own order with a diagonal core lambda/(1-lambda)
 *   (GMSLRDPrecondition.cpp:1008-1032);
 * DEFLATION -- AMSED only: the same restarted Lanczos, H = (eps-lam)/lam with
 *   lam = 1-mu for every mu > 1-eps (AMSEDEigenvalueDeflation, :1233-1265);
 * SUBSPACE -- no reference counterpart: AMSED one Lanczos run of
 *   max(2 rank, rank+16) steps, H = theta/(1-theta) on the largest Ritz values in
 *   [0,1); GMSLR the largest real eigenvalues below 1, an orthonormal basis of
 *   their eigenvectors and the full core (I-R)^-1 - I, exact on that subspace. */
enum { PP_LOWRANK_AUTO = 0, PP_LOWRANK_REFERENCE = 1, PP_LOWRANK_DEFLATION = 2, PP_LOWRANK_SUBSPACE = 3 };

typedef struct {
    int kind;
    int ready;              /* isReady() */
    uint32_t n;
    uint32_t annz;          /* getOriginalNonZeroCount() */
    uint32_t mnnz;          /* getPreconditionNonZeroCount() */
    double fill_ratio;      /* getPreconditionFillinRatio() */
    uint32_t levels_lower;  /* SpTRSV dependency depth (AMSED / GMSLR: of the first
                               level's subdomain-block factor, its schedule in trsv) */
    uint32_t levels_upper;
    double setup_ms;        /* host factorization + analysis + upload */
    uint32_t fail_row;      /* row of a failed pivot */
    int32_t sweep_variant;  /* staged-sweep loads chosen at setup: bit 0 non-temporal,
                               bit 1 16-bit dependency offsets, bit 2 batched loads;
                               8: wide contiguous levels in sliced-column form */
    int32_t fused_levels;   /* 1: last forward and first backward level share one launch */
    int32_t trsv;           /* the SpTRSV schedule in use (PP_TRSV_AUTO resolved at setup) */
    float phase_ms[4];      /* setup phases: ordering, factorization, sweep analysis + upload,
                               schedule choice + load-variant timing */
    uint32_t device_refactors; /* pp_precond_refactor calls that factored on the device */
    int32_t reduced_form;   /* operators of the reduced CG (IC(0) in a two-colour order,
                               built by the first solve that runs it): 0 none, 1 built
                               on the host, 2 built on the device, 3 refreshed on the
                               device after new values */
    float reduced_form_ms;  /* the last build or refresh of those operators */
    int32_t setup_on_device; /* 1: setup() ran on the device (IC(0) / ILU(0) in multicolour
                               order; the host copy of the factor is built when asked for) */
} pp_precond_info;

void pp_precond_opts_default(pp_precond_opts *o);
int pp_precond_create(pp_ctx *ctx, pp_csr *A, const pp_precond_opts *opts, pp_precond **out);
/* setup(): factorization on the host, level analysis, upload (BasePreconditon.h:102) */
int pp_precond_setup(pp_ctx *ctx, pp_precond *M);
/* Numeric re-factorization from the matrix's current values (pp_csr_set_values
 * keeps the pattern).  No reference counterpart: the reference's setup() does
 * nothing once ready (IncompleteCholesky.cpp:105) and a new object has to be
 * made.  IC(0) and ILU(0) (any ordering, any schedule but PP_TRSV_BLOCK/PIPE) are
 * factored on the device: one launch per dependency level, one thread per
 * row, the host factorization's operation order (the factor is the same bit
 * for bit), written into the existing sweeps (graphs captured with them stay
 * valid); the symbolic maps are built by the first call.  Other kinds run the
 * full setup() again.  A failing pivot returns PP_ERR_NOT_SPD /
 * PP_ERR_ZERO_PIVOT with pp_precond_info.fail_row, and the preconditioner is
 * then not ready until a successful call. */
int pp_precond_refactor(pp_ctx *ctx, pp_precond *M);
/* Pins the level-sweep variant of both precisions instead of the timed choice
 * (0..7 the staged kernel's loads, 8 the sliced-column form of the wide
 * contiguous levels, kept by factors of at least 2^20 rows, the other levels
 * staged with the default loads); every variant gives the same result bit for
 * bit.  PP_ERR_INVALID outside 0..8 or for 8 when no level has that form;
 * PP_ERR_NOT_READY before setup.  No reference counterpart. */
int pp_precond_set_sweep_variant(pp_precond *M, int variant);
/* y = M^{-1} x (MInvSolve in->out, BasePreconditon.h:113-128); y may equal x */
int pp_precond_apply(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y);
/* part: 0 lower (MSolveLower), 1 diagonal (MSolveDiagonal), 2 upper (MSolveUpper),
 * in place (TriangularPrecondition, BasePreconditon.h:200-303) */
int pp_precond_apply_part(pp_ctx *ctx, pp_precond *M, int part, int dtype, void *x);
int pp_precond_info_get(const pp_precond *M, pp_precond_info *info);
/* Bytes one M^{-1} x streams through HBM in the library's sweep storage (both
 * sweeps: values, 16/32-bit dependency indices, row pointers, level row map
 * where read, diagonal, block tables, vectors).  in_solve != 0: as applied
 * inside a solve on the factor's own matrix (no colour-order gather/scatter).
 * The algorithmic traffic the sweep roofline is priced on.  Triangular
 * factors only; no reference counterpart. */
int pp_precond_apply_bytes(const pp_precond *M, int dtype, int in_solve, uint64_t *bytes);
/* Copy the host factor out (for inspection / parity tests).
 * which: 0 L^T (IC, ILDLT strict upper), 1 L (ILU), 2 U (ILU).
 * Call with NULL arrays to get nnz; d receives the diagonal (ILU 1/u_ii, ILDLT d). */
int pp_precond_factor(const pp_precond *M, int which, uint32_t *nnz, uint32_t *rowptr,
                      uint32_t *colidx, double *values, double *d);
void pp_precond_destroy(pp_precond *M);

/* Host-only factorization (no device needed): the setup() step of
 * IncompleteCholesky / IncompleteLU / IncompleteLDLT (and IC(0), ILU(0)) on
 * host arrays.  Returns PP_OK or the factorization failure; the handle is
 * created in both cases (the partial factor of a failure is kept). */
typedef struct pp_hfactor pp_hfactor;
int pp_factor_host(const pp_precond_opts *opts, uint32_t n, const uint32_t *rowptr, const uint32_t *colidx,
                   const double *values, pp_hfactor **out);
/* which: 0 L^T, 1 L, 2 U; NULL arrays to query nnz; d as in pp_precond_factor */
int pp_hfactor_get(const pp_hfactor *F, int which, uint32_t *nnz, uint32_t *rowptr, uint32_t *colidx,
                   double *values, double *d);
int pp_hfactor_info(const pp_hfactor *F, uint32_t *mnnz, uint32_t *fail_row, uint32_t *levels_lower,
                    uint32_t *levels_upper);
void pp_hfactor_destroy(pp_hfactor *F);

/* Multilevel (AMSED / GMSLR) introspection: number of levels, rows per level
 * and the low-rank correction rank of each level (AMSEDPrecondition.h:417-424). */
int pp_ml_info(const pp_precond *M, uint32_t *nlev, uint32_t *level_sizes, uint32_t *ranks, uint32_t cap);
/* the hierarchical-interface-decomposition ordering: new position k holds row perm[k] */
int pp_ml_permutation(const pp_precond *M, uint32_t *perm);
/* W (n x r, col-major) = V (n x k, col-major) * Y (k x r, col-major), FP64 on the
 * matrix cores (device pointers): the tall-skinny product of the low-rank
 * correction (cublasDgemm at AMSEDPrecondition.cpp:1851). */
int pp_gemm_tall_f64(pp_ctx *ctx, uint32_t n, uint32_t k, uint32_t r, const double *V, const double *Y, double *W);

/* ---- solvers (BaseSolver::solve, include/Solver/BaseSolver.h:9-17;
 *      CG, source/Solver/CG.cpp:33-83; GMRES, source/Solver/GMRES.cpp:59-221) ---- */
enum { PP_SOLVER_CG = 0, PP_SOLVER_GMRES = 1 };
enum { PP_FP64 = 0, PP_MIXED = 1 };
enum { PP_LAUNCH_GRAPH = 0, PP_LAUNCH_STREAM = 1 };
/* CG iteration: AUTO = the reference's FlexibleCG / NormalCG for FP64 solves;
 * for the FP32 inner solves of PP_MIXED the reduced CG when it applies, else
 * the single-reduction PCG;
 * REFERENCE = FlexibleCG / NormalCG everywhere (CG.cpp:33-83);
 * SINGLE_REDUCTION = Chronopoulos-Gear PCG everywhere (one SpMV carrying the
 * three inner products, one fused update; no reference counterpart);
 * REDUCED = IC(0) of a two-colour ordering (PP_ORDER_MULTICOLOR on a matrix
 * whose colour blocks are diagonal, e.g. 5- and 7-point stencils) applied in
 * Schur-complement form: the split-preconditioned operator L^-1 A L^-T is
 * the identity on the first colour and the scaled Schur complement on the
 * second (Eisenstat 1981), so CG runs on the second colour's unknowns with
 * no triangular sweep -- the same Krylov iteration as IC(0)-PCG (an FP64
 * solve then confirms the true residual in an outer step).  No reference
 * counterpart; PP_ERR_UNSUPPORTED when the structure is absent. */
enum { PP_CG_AUTO = 0, PP_CG_REFERENCE = 1, PP_CG_SINGLE_REDUCTION = 2, PP_CG_REDUCED = 3 };

typedef struct {
    int solver;          /* PP_SOLVER_* */
    int flexible;        /* 1: FlexibleCG / FlexibleGMRES (default); 0: Normal* */
    int precision;       /* PP_FP64 or PP_MIXED (FP32 inner, FP64 residual/refinement) */
    double tol;          /* relative residual target */
    uint32_t maxit;      /* CG: maxIter; GMRES: total budget (reference default 500) */
    uint32_t restart;    /* GMRES(m) */
    int reference_quirks;/* GMRES: reproduce the reference's converged-cycle
                            back-substitution with dim = j (GMRES.cpp:194-201) */
    uint32_t check_every;/* iterations between host convergence polls */
    double inner_tol;    /* mixed: relative reduction asked of each inner solve */
    uint32_t max_outer;  /* mixed: refinement steps */
    int record_history;  /* keep per-iteration ||r||/||b|| */
    int launch;          /* PP_LAUNCH_STREAM (pp_solve_opts_default): the iteration
                            body's kernels are launched on the stream each time;
                            PP_LAUNCH_GRAPH: the body is captured once in a hipGraph
                            and replayed (consecutive graph launches leave 8-14 us
                            between bodies on MI355X, stream launches none) */
    int cg_variant;      /* PP_CG_* */
    uint32_t reserved[6];
} pp_solve_opts;

typedef struct {
    int32_t iters;          /* getNumIter(), reference counting convention */
    int32_t converged;      /* getConvergence() */
    double final_error;     /* getError(): the recursive residual estimate */
    double true_relres;     /* ||b - A x|| / ||b|| recomputed in FP64 at exit (a
                               refinement solve's last FP64 residual check) */
    int32_t outer_iters;    /* mixed: refinement steps taken */
    int32_t inner_iters;    /* mixed: total inner iterations */
    double solve_ms;        /* stream time, setup excluded */
    int32_t status;         /* PP_OK or the failure */
    int32_t nhist;
    int32_t cg_variant;     /* CG: the iteration that ran (PP_CG_REFERENCE,
                               _SINGLE_REDUCTION or _REDUCED; for PP_MIXED the
                               inner solves'); GMRES: PP_CG_AUTO */
    int32_t inner_fp64_steps; /* refinement steps whose inner solve ran in FP64
                               (PP_MIXED reduced CG: the steps after an FP32
                               step fell short of its own residual test on the
                               true residual) */
    uint32_t reserved[6];
} pp_solve_report;

void pp_solve_opts_default(pp_solve_opts *o, int solver);
/* Device-resident solve: b, x are FP64 device pointers; x holds x0 on entry. */
int pp_solve(pp_ctx *ctx, pp_csr *A, pp_precond *M, const pp_solve_opts *opts,
             const double *b, double *x, pp_solve_report *rep);
/* Host in / host out (the BaseSolver::solve drop-in): uploads b and x0,
 * solves, downloads x. */
int pp_solve_host(pp_ctx *ctx, pp_csr *A, pp_precond *M, const pp_solve_opts *opts,
                  const double *b_host, double *x_host, pp_solve_report *rep);
/* pp_solve without any allocation: no hipMalloc, hipFree or other allocating
 * call, no graph instantiation.  No counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  FP64 CG only (FlexibleCG /
 * NormalCG, or PP_CG_SINGLE_REDUCTION) with no preconditioner or a triangular
 * one (ICT, ILUT, ILDLT, IC(0), ILU(0), either ordering); the iteration
 * kernels are launched on the stream (opts->launch is ignored) and no residual
 * history is kept.  The work vectors live in one fixed device array of
 * 128 MiB per GPU (held from the first context on that GPU until the process
 * ends): four n-vectors (five for the single-reduction CG), the final
 * residual, and b and x in colour order for a multicolour factor, each
 * rounded up to 256 bytes -- n <= 2,396,745 for a colour-ordered FlexibleCG
 * solve, n <= 3,355,443 in natural order.  Larger systems fail with
 * PP_ERR_INVALID before any work; other solver kinds with PP_ERR_UNSUPPORTED;
 * a multicolour factor whose matrix values changed since setup with
 * PP_ERR_NOT_READY.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two contexts or threads) race. */
int pp_solve_noalloc(pp_ctx *ctx, pp_csr *A, pp_precond *M, const pp_solve_opts *opts,
                     const double *b, double *x, pp_solve_report *rep);
/* residual history of the last solve on this context (nhist entries) */
int pp_solve_history(pp_ctx *ctx, double *out, uint32_t cap, uint32_t *count);

/* ---- adjoint (no reference counterpart): solve A^T lambda = g with the same
 * Krylov/preconditioner machinery on the explicit transpose. ---- */
int pp_adjoint_solve(pp_ctx *ctx, pp_csr *A, const pp_precond_opts *popts,
                     const pp_solve_opts *opts, const double *g, double *lambda,
                     pp_solve_report *rep);
/* dL/dA on the pattern of A: grad[k] = -lambda[row(k)] * x[col(k)] */
int pp_csr_outer_on_pattern(pp_ctx *ctx, const pp_csr *A, double scale, const double *lambda,
                            const double *x, double *grad_values);

/* ---- batched solves (no reference counterpart): nsys independent CSR systems
 * stored back to back (block-diagonal), solved by one persistent launch of
 * IC(0)-PCG with one workgroup per system.  Offsets are host arrays. ---- */
typedef struct pp_batch pp_batch;
int pp_batch_create(pp_ctx *ctx, uint32_t nsys, const uint32_t *row_off, const uint32_t *rowptr,
                    const uint32_t *colidx, const double *values, pp_batch **out);
int pp_batch_set_values(pp_ctx *ctx, pp_batch *B, const double *values, int on_device);
/* PP_ORDER_MULTICOLOR: solve every system as P_s A_s P_s^T with its rows
 * grouped by a greedy colouring (one sweep level per colour instead of one per
 * anti-diagonal); b and x stay in the caller's order.  Re-factors.  It changes
 * the preconditioner and the iteration counts.  Default PP_ORDER_NATURAL. */
int pp_batch_set_ordering(pp_ctx *ctx, pp_batch *B, int ordering);
/* b, x: device, concatenated; every system starts from x0 = 0 and x receives
 * the solutions.  precision PP_FP64 or PP_MIXED.
 * iters_out (host, nsys) receives per-system iteration counts (may be NULL). */
int pp_batch_solve(pp_ctx *ctx, pp_batch *B, const pp_solve_opts *opts, const double *b,
                   double *x, int32_t *iters_out, pp_solve_report *rep);
/* transpose-solve for the adjoint of every system at once */
int pp_batch_solve_t(pp_ctx *ctx, pp_batch *B, const pp_solve_opts *opts, const double *g,
                     double *lambda, int32_t *iters_out, pp_solve_report *rep);
int pp_batch_tune(pp_ctx *ctx, pp_batch *B, int *chosen_block);
/* grad[k] = scale * lambda[row(k)] * x[col(k)] over the batch's pattern (device) */
int pp_batch_outer_on_pattern(pp_ctx *ctx, const pp_batch *B, double scale, const double *lambda,
                              const double *x, double *grad_values);
void pp_batch_destroy(pp_batch *B);

/* ---- model problems (ModelProblem.cpp:68-133 / MatrixTools.h:16-62 and the
 * operators the reference lacks).  Two calls: first with NULL arrays for nnz. ---- */
int pp_generate_laplacian(int nx, int ny, int nz, double ax, double ay, double az, double shift,
                          uint32_t *n, uint32_t *nnz, uint32_t *rowptr, uint32_t *colidx,
                          double *values);
int pp_generate_upwind(int nx, int ny, int nz, double bx, double by, double bz, uint32_t *n,
                       uint32_t *nnz, uint32_t *rowptr, uint32_t *colidx, double *values);
int pp_generate_vardiff2d(int side, const double *kappa, uint32_t *n, uint32_t *nnz,
                          uint32_t *rowptr, uint32_t *colidx, double *values);
/* ---- MatrixMarket input (HOST::ReadMtxTools, source/utils/TestTools/ReadMtxTools.cpp;
 * no device needed).  pp_mtx_read takes a "%%MatrixMarket matrix coordinate real"
 * file with general, symmetric or skew-symmetric storage (hermitian real data
 * is taken as stored, as the reference does) and applies the reference
 * reader's rules: indices minus `base` (1 for files written by the format's
 * convention), entries of magnitude <= 10 DBL_EPSILON dropped, symmetric and
 * skew-symmetric storage expanded with the mirrored (negated) entry of every
 * off-diagonal entry, rows in file order then sorted by column with
 * duplicates kept (transCOO2CSR, include/MatrixClass/MatrixTools.h:16-62).
 * Other fields (complex, integer, pattern) and dense files are
 * PP_ERR_UNSUPPORTED; pp_global_error() gives the message. ---- */
typedef struct pp_mtx pp_mtx;
int pp_mtx_read(const char *path, int base, pp_mtx **out);
int pp_mtx_dims(const pp_mtx *m, uint32_t *nrow, uint32_t *ncol, uint32_t *nnz);
/* copies the CSR arrays out (nrow + 1, nnz, nnz entries; any may be NULL) */
int pp_mtx_csr(const pp_mtx *m, uint32_t *rowptr, uint32_t *colidx, double *values);
void pp_mtx_destroy(pp_mtx *m);
/* a dense "%%MatrixMarket matrix array real" file (right-hand sides,
 * ReadMtxTools::loadRightHand): nrow x ncol values, column-major; values NULL
 * to query the sizes */
int pp_mtx_read_dense(const char *path, uint32_t *nrow, uint32_t *ncol, double *values);

/* ---- reverse Cuthill-McKee (HOST::rcmReorderCSR, MatrixReorderTools.h:59-80,
 * over the vendored genrcm of external_packages/RCM): perm[k] = the row placed
 * at position k, for a structurally symmetric pattern (its stored entries, the
 * diagonal included, are the adjacency lists).  The same permutation as the
 * reference's. ---- */
int pp_reorder_rcm(uint32_t n, const uint32_t *rowptr, const uint32_t *colidx, uint32_t *perm);

/* seeded SplitMix64 inputs: kind 0 uniform(-1,1), 1 normal(0,1) */
int pp_fill_random(int kind, uint64_t seed, uint64_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif /* PRISPACK_AMD_H */
