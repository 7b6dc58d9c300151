// pp_reduced_dev.hip -- new values of an existing reduced form on the device.
//
// The reduced CG's operators (KA, KB, KF, KE, Dt, Dt^-1/2; pp_reduced.cpp)
// depend on the colour-ordered matrix's pattern only through their own
// patterns, which mirror its rows: KA row i is row i of the first colour
// without its diagonal, KE the same with 1/d_i first, KB and KF row q are
// row n0 + q of the second colour without its diagonal and with one more
// entry last.  When the matrix keeps its pattern and receives new values (a
// numeric re-factorization), these kernels recompute every value in place,
// one thread per row, with the host construction's formulas and operation
// order (no contraction), so the operators are the same bit for bit and the
// captured iteration graphs, which hold their addresses, stay valid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "pp_internal.h"

namespace pp {
namespace {

constexpr int kB = 256;

unsigned grid_of(uint32_t n) { return std::max(1u, std::min<uint32_t>((n + kB - 1) / kB, 65535u)); }

// d[i] = a_ii (every row holds one); fail when a first-colour pivot is not positive
__global__ void k_rr_diag(uint32_t n, uint32_t n0, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                          const double *__restrict__ av, double *d, int *fail) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        double v = 0.0;
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) {
                v = av[k];
                break;
            }
        d[i] = v;
        if (!(v > 0.0)) atomicOr(fail, 1);
    }
}

// Second colour, row i = n0 + q: Dt_q = a_ii - sum_j a_ij^2 / a_jj over the
// first-colour neighbours in row order; KB = -a_ij / sqrt(a_jj Dt_q) then
// a_ii / Dt_q, KF = -a_ij / a_jj * Dt_q^-1/2 then Dt_q^-1/2 (reduced_host).
__global__ void k_rr_second(uint32_t n0, uint32_t n1, const uint32_t *__restrict__ rp,
                            const uint32_t *__restrict__ ci, const double *__restrict__ av,
                            const double *__restrict__ d, const uint32_t *__restrict__ kb_rp, double *kb_v,
                            const uint32_t *__restrict__ kf_rp, double *kf_v, double *dt, double *s1, int *fail) {
#pragma clang fp contract(off)
    for (uint32_t q = blockIdx.x * kB + threadIdx.x; q < n1; q += gridDim.x * kB) {
        const uint32_t i = n0 + q;
        double t = d[i];
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] < n0) t -= av[k] * av[k] / d[ci[k]];
        dt[q] = t;
        if (!(t > 0.0 && isfinite(t))) {
            atomicOr(fail, 1);
            continue;
        }
        const double rs = 1.0 / sqrt(t);
        uint32_t pb = kb_rp[q], pf = kf_rp[q];
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const uint32_t j = ci[k];
            if (j == i) continue;
            const double a = av[k];
            kb_v[pb++] = -a / sqrt(d[j] * t);
            kf_v[pf++] = -a / d[j] * rs;
        }
        kb_v[pb] = d[i] / t;
        kf_v[pf] = rs;
        s1[q] = rs;
    }
}

// First colour, row i: KE = 1/a_ii then -a_ij / a_ii, KA = a_ij / sqrt(a_ii Dt_j)
__global__ void k_rr_first(uint32_t n0, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                           const double *__restrict__ av, const double *__restrict__ d,
                           const double *__restrict__ dt, const uint32_t *__restrict__ ka_rp, double *ka_v,
                           const uint32_t *__restrict__ ke_rp, double *ke_v) {
#pragma clang fp contract(off)
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n0; i += gridDim.x * kB) {
        uint32_t pa = ka_rp[i], pe = ke_rp[i];
        ke_v[pe++] = 1.0 / d[i];
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const uint32_t j = ci[k];
            if (j == i) continue;
            const double a = av[k];
            ka_v[pa++] = a / sqrt(d[i] * dt[j - n0]);
            ke_v[pe++] = -a / d[i];
        }
    }
}

}  // namespace

bool reduced_refresh(pp_ctx *ctx, ReducedForm *R, pp_csr *As) {
    const uint32_t n = As->nrow, n0 = R->n0, n1 = R->n1;
    PP_CHECK(n == n0 + n1 && R->KA && R->KB && R->KF && R->KE, PP_ERR_INVALID, "reduced form does not match");
    if (!R->dwork) R->dwork = dalloc<double>(n);
    double *d = R->dwork;
    if (!R->dfail) R->dfail = dalloc<int>(1);
    int *fail = R->dfail;
    launch_fill_zero(ctx, fail, sizeof(int));
    k_rr_diag<<<grid_of(n), kB, 0, ctx->stream>>>(n, n0, As->rp, As->ci, As->v64, d, fail);
    k_rr_second<<<grid_of(n1), kB, 0, ctx->stream>>>(n0, n1, As->rp, As->ci, As->v64, d, R->KB->rp, R->KB->v64,
                                                     R->KF->rp, R->KF->v64, R->dt_64, R->s1_64, fail);
    k_rr_first<<<grid_of(n0), kB, 0, ctx->stream>>>(n0, As->rp, As->ci, As->v64, d, R->dt_64, R->KA->rp,
                                                    R->KA->v64, R->KE->rp, R->KE->v64);
    PP_HIP(hipGetLastError());
    int bad = 0;
    PP_HIP(hipMemcpyAsync(&bad, fail, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    if (bad) return false;
    // the copies: sliced-column forms, FP32 forms, Dt and Dt^-1/2 in FP32
    if (R->ka.ok) ell_fill_values(ctx, R->ka, PP_F64, R->KA->rp, R->KA->v64);
    if (R->kb.ok) ell_fill_values(ctx, R->kb, PP_F64, R->KB->rp, R->KB->v64);  // its diagonal too
    for (pp_csr *K : {R->KA, R->KB, R->KF, R->KE}) {
        if (K->ell.ok) ell_fill_values(ctx, K->ell, PP_F64, K->rp, K->v64);
        K->f32_dirty = true;  // the next FP32 use converts again (csr_need_f32)
        if (K->v32) {
            launch_convert(ctx, K->nnz, PP_F64, K->v64, PP_F32, K->v32);
            if (K->ell.ok && K->ell.v32) ell_fill_values(ctx, K->ell, PP_F32, K->rp, K->v32);
            K->f32_dirty = false;
        }
        ++K->version;
    }
    for (EllOp *E : {&R->ka, &R->kb}) {
        if (E->ok && E->v32) launch_convert(ctx, E->width * E->n, PP_F64, E->v64, PP_F32, E->v32);
        if (E->ok && E->dg32) launch_convert(ctx, E->n, PP_F64, E->dg64, PP_F32, E->dg32);
    }
    if (R->s1_32) launch_convert(ctx, n1, PP_F64, R->s1_64, PP_F32, R->s1_32);
    if (R->dt_32) launch_convert(ctx, n1, PP_F64, R->dt_64, PP_F32, R->dt_32);
    PP_HIP(hipGetLastError());
    // the largest absolute row sum, from the matrix's host copy of its values
    double anorm = 0.0;
    {
        std::mutex mu;
        const uint32_t *rp = As->h_rp.data();
        const double *hv = As->h_v.data();
        parallel_for(n, 1u << 15, [&](size_t i0, size_t i1) {
            double m = 0.0;
            for (size_t i = i0; i < i1; ++i) {
                double t = 0.0;
                for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) t += std::fabs(hv[k]);
                m = std::max(m, t);
            }
            std::lock_guard<std::mutex> g(mu);
            anorm = std::max(anorm, m);
        });
    }
    R->anorm = anorm;
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return true;
}

}  // namespace pp
