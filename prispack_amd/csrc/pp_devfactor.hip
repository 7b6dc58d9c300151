// pp_devfactor.hip -- numeric IC(0) / ILU(0) factorization on the device.
//
// A preconditioner set up once keeps its symbolic structure (the zero-fill
// pattern, its dependency levels, and where every factor entry lands in the
// sweeps' level-ordered storage).  When the matrix receives new values, a
// second setup() runs the numeric factorization here, one launch per
// dependency level (the colours of a multicolour order), one thread per row,
// and scatters the factor into the sweeps through those maps -- no host
// factorization and no upload.  Every row's operation sequence is the host
// factorization's (pp_factor.cpp: factor_ic0, factor_ilu0), so the factor is
// the same bit for bit.
#include <hip/hip_runtime.h>

#include "pp_internal.h"

namespace pp {
namespace {

constexpr int kB = 256;

// IC(0) rows of one level: L strict lower (rows sorted by column, entries
// rp/ci), initial values a_ij gathered from the matrix, then
//   l_ij = (a_ij - sum_{k<j} l_ik l_jk) / l_jj   (the sum in the order of row j's entries)
//   l_ii = sqrt(a_ii - sum_j l_ij^2)
// as factor_ic0 computes them.  A non-positive pivot records the lowest such
// row in *fail and is replaced by 1, as the host does.
__global__ __launch_bounds__(kB) void k_ic0_level(const uint32_t *__restrict__ rows, uint32_t nrows,
                                                  const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                                                  const uint32_t *__restrict__ src, const uint32_t *__restrict__ dsrc,
                                                  const double *__restrict__ av, double *lv, double *diag, int *fail) {
#pragma clang fp contract(off)
    const uint32_t q = blockIdx.x * kB + threadIdx.x;
    if (q >= nrows) return;
    const uint32_t i = rows[q];
    const uint32_t i0 = rp[i], i1 = rp[i + 1];
    for (uint32_t t = i0; t < i1; ++t) {
        const uint32_t k = ci[t];
        double s = av[src[t]];
        uint32_t p = i0;
        for (uint32_t u = rp[k]; u < rp[k + 1]; ++u) {
            const uint32_t c = ci[u];
            while (p < t && ci[p] < c) ++p;
            if (p < t && ci[p] == c) s -= lv[p] * lv[u];
        }
        lv[t] = s / diag[k];
    }
    double s = av[dsrc[i]];
    for (uint32_t t = i0; t < i1; ++t) s -= lv[t] * lv[t];
    if (!(s > 0.0)) {
        atomicMin(fail, (int)i);
        s = 1.0;
    }
    diag[i] = sqrt(s);
}

// ILU(0) rows of one level in IKJ form on the matrix's own pattern (rows
// sorted): w = a_i, then for every k < i in the row, w_ik /= u_kk and
// w_ij -= w_ik u_kj over row k's upper part, as factor_ilu0 does; D_i = 1/u_ii.
// A pivot below 1e-15 in magnitude records the lowest such row in *fail.
__global__ __launch_bounds__(kB) void k_ilu0_level(const uint32_t *__restrict__ rows, uint32_t nrows,
                                                   const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                                                   const uint32_t *__restrict__ dpos, const double *__restrict__ av,
                                                   double *w, double *dinv, int *fail) {
#pragma clang fp contract(off)
    const uint32_t q = blockIdx.x * kB + threadIdx.x;
    if (q >= nrows) return;
    const uint32_t i = rows[q];
    const uint32_t i0 = rp[i], i1 = rp[i + 1];
    for (uint32_t t = i0; t < i1; ++t) w[t] = av[t];
    for (uint32_t t = i0; t < i1; ++t) {
        const uint32_t k = ci[t];
        if (k >= i) break;
        const double wt = w[t] / w[dpos[k]];
        w[t] = wt;
        uint32_t p = i0;
        for (uint32_t u = dpos[k] + 1; u < rp[k + 1]; ++u) {
            const uint32_t c = ci[u];
            while (p < i1 && ci[p] < c) ++p;
            if (p < i1 && ci[p] == c) w[p] -= wt * w[u];
        }
    }
    const double piv = w[dpos[i]];
    if (fabs(piv) < 1e-15) atomicMin(fail, (int)i);
    dinv[i] = 1.0 / piv;
}

template <class T>
__global__ void k_gather_map(uint32_t n, const uint32_t *__restrict__ map, const double *__restrict__ in, T *out) {
    for (uint32_t q = blockIdx.x * kB + threadIdx.x; q < n; q += gridDim.x * kB) out[q] = (T)in[map[q]];
}

__global__ void k_to_index(uint32_t n, const double *__restrict__ in, uint32_t *out) {
    for (uint32_t q = blockIdx.x * kB + threadIdx.x; q < n; q += gridDim.x * kB) out[q] = (uint32_t)in[q];
}

unsigned grid_of(uint32_t n) { return std::max(1u, std::min<uint32_t>((n + kB - 1) / kB, 65535u)); }

}  // namespace

void launch_ic0_level(pp_ctx *ctx, const uint32_t *rows, uint32_t nrows, const uint32_t *rp, const uint32_t *ci,
                      const uint32_t *src, const uint32_t *dsrc, const double *av, double *lv, double *diag,
                      int *fail) {
    if (!nrows) return;
    k_ic0_level<<<(nrows + kB - 1) / kB, kB, 0, ctx->stream>>>(rows, nrows, rp, ci, src, dsrc, av, lv, diag, fail);
    PP_HIP(hipGetLastError());
}

void launch_ilu0_level(pp_ctx *ctx, const uint32_t *rows, uint32_t nrows, const uint32_t *rp, const uint32_t *ci,
                       const uint32_t *dpos, const double *av, double *w, double *dinv, int *fail) {
    if (!nrows) return;
    k_ilu0_level<<<(nrows + kB - 1) / kB, kB, 0, ctx->stream>>>(rows, nrows, rp, ci, dpos, av, w, dinv, fail);
    PP_HIP(hipGetLastError());
}

void launch_gather_map(pp_ctx *ctx, uint32_t n, const uint32_t *map, const double *in, int dtype, void *out) {
    if (!n) return;
    if (dtype == PP_F64) k_gather_map<double><<<grid_of(n), kB, 0, ctx->stream>>>(n, map, in, (double *)out);
    else k_gather_map<float><<<grid_of(n), kB, 0, ctx->stream>>>(n, map, in, (float *)out);
    PP_HIP(hipGetLastError());
}

void launch_to_index(pp_ctx *ctx, uint32_t n, const double *in, uint32_t *out) {
    if (!n) return;
    k_to_index<<<grid_of(n), kB, 0, ctx->stream>>>(n, in, out);
    PP_HIP(hipGetLastError());
}

}  // namespace pp
