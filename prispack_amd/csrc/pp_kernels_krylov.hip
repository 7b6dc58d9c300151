// pp_kernels_krylov.hip -- tall-skinny GEMV kernels of the GMRES orthogonalisation
// (classical Gram-Schmidt, applied twice) and mixed-precision glue kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pp_internal.h"

namespace pp {

namespace {
constexpr int kGemvBlocks = 2048;
constexpr int kMaxCols = 257;

__device__ inline double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
}  // namespace

// Buffer of GMRES coefficients on the device: [0,320) h, [320, 640) y or a copy of h, and
// the per-column partials of the GEMV^T behind it.
double *gmres_coeff_buffer(pp_ctx *ctx) {
    return (double *)ws_get(ctx, kWsGmresCoeff, (size_t)(640 + (size_t)kMaxCols * kGemvBlocks) * sizeof(double));
}

// partial[c][block] = sum over the block's rows of V[r, c] * w[r].  The
// columns go in groups of kGemvGroup: one pass over the block's rows
// accumulates the whole group in registers (every column's loads in flight
// together, no barrier per column), then one reduction per group.  Each
// column is summed in the same order as a column-at-a-time pass (the
// thread's rows in stride order, then the wave and the four waves), so the
// result does not depend on the grouping.
template <int NG, class TV>
__global__ __launch_bounds__(256) void k_gemv_t_partial(uint32_t n, uint32_t ncol, const TV *__restrict__ V,
                                                        size_t ld, const TV *__restrict__ w, double *partial,
                                                        const int *stop) {
    __shared__ double sh[NG][4];
    if (stop && *stop) return;
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = blockIdx.x * chunk, r1 = min(n, r0 + chunk);
    for (uint32_t c0 = 0; c0 < ncol; c0 += NG) {
        const uint32_t nc = min((uint32_t)NG, ncol - c0);
        double acc[NG];
#pragma unroll
        for (int c = 0; c < NG; ++c) acc[c] = 0.0;
        for (uint32_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
            const double wr = w[r];
            // every slot's load is unconditional (slots past the group's
            // last column reread that column, a cache hit, and their sums are
            // never used), so the loads issue back to back instead of one
            // behind each branch
            double vv[NG];
#pragma unroll
            for (int c = 0; c < NG; ++c) vv[c] = (double)V[ld * (c0 + min((uint32_t)c, nc - 1)) + r];
#pragma unroll
            for (int c = 0; c < NG; ++c) acc[c] += vv[c] * wr;
        }
#pragma unroll
        for (int c = 0; c < NG; ++c)
            if ((uint32_t)c < nc) {
                const double v = wsum(acc[c]);
                if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = v;
            }
        __syncthreads();
        if (threadIdx.x < nc)
            partial[(size_t)(c0 + threadIdx.x) * gridDim.x + blockIdx.x] =
                ((sh[threadIdx.x][0] + sh[threadIdx.x][1]) + sh[threadIdx.x][2]) + sh[threadIdx.x][3];
        __syncthreads();
    }
}

// h[c] = sum_b partial[c][b] in block order (deterministic)
__global__ __launch_bounds__(256) void k_gemv_t_reduce(uint32_t nb, const double *partial, double *h,
                                                       const int *stop) {
    __shared__ double sh[4];
    if (stop && *stop) return;
    const uint32_t c = blockIdx.x;
    double acc = 0.0;
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) acc += partial[(size_t)c * nb + b];
    acc = wsum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) h[c] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

template <class TV>
void launch_gemv_t(pp_ctx *ctx, uint32_t n, uint32_t ncol, const TV *V, size_t ld, const TV *w, double *dest) {
    PP_CHECK(ncol <= kMaxCols, PP_ERR_UNSUPPORTED, "too many columns");
    double *buf = gmres_coeff_buffer(ctx);
    double *partial = buf + 640;
    // the column group no wider than the columns (4, 8 or 16 slots)
    if (ncol <= 4)
        k_gemv_t_partial<4, TV><<<kGemvBlocks, 256, 0, ctx->stream>>>(n, ncol, V, ld, w, partial, ctx->stop_word);
    else if (ncol <= 8)
        k_gemv_t_partial<8, TV><<<kGemvBlocks, 256, 0, ctx->stream>>>(n, ncol, V, ld, w, partial, ctx->stop_word);
    else k_gemv_t_partial<16, TV><<<kGemvBlocks, 256, 0, ctx->stream>>>(n, ncol, V, ld, w, partial, ctx->stop_word);
    k_gemv_t_reduce<<<ncol, 256, 0, ctx->stream>>>(kGemvBlocks, partial, dest ? dest : buf, ctx->stop_word);
    PP_HIP(hipGetLastError());
}

// w[r] -= sum_c V[r, c] h[c] (the sum in FP64 for either storage type)
template <class TV>
__global__ __launch_bounds__(256) void k_gemv_sub(uint32_t n, uint32_t ncol, const TV *__restrict__ V,
                                                  size_t ld, const double *__restrict__ h, TV *__restrict__ w,
                                                  const int *stop) {
    if (stop && *stop) return;
    constexpr int kU = 4;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double acc = 0.0;
        // the columns four at a time: the four loads first (past the last
        // column they reread it), then the products in column order
        for (uint32_t c0 = 0; c0 < ncol; c0 += kU) {
            double vv[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) vv[u] = (double)V[ld * min(c0 + u, ncol - 1) + r];
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (c0 + u < ncol) acc += vv[u] * h[c0 + u];
        }
        w[r] = (TV)((double)w[r] - acc);
    }
}

template <class TV>
void launch_gemv_sub(pp_ctx *ctx, uint32_t n, uint32_t ncol, const TV *V, size_t ld, const double *h, TV *w) {
    unsigned g = std::min<uint32_t>((n + 255) / 256, 8192);
    k_gemv_sub<TV><<<g, 256, 0, ctx->stream>>>(n, ncol, V, ld, h, w, ctx->stop_word);
    PP_HIP(hipGetLastError());
}

// The first pass's update and the second pass's products in one sweep over
// V (ncol <= kFuseCols): w[r] -= sum_c V[r, c] h[c] exactly as k_gemv_sub
// forms it, then partial[c][block] of V[r, c] * w[r] with the new w exactly
// as k_gemv_t_partial forms it (same rows per workgroup, same order), so
// the results are those of the two separate launches, with one read of V
// fewer.  Each thread keeps its row's ncol basis entries in registers
// between the two uses.
constexpr int kFuseCols = 32;
template <int NC, class TV>
__global__ __launch_bounds__(256) void k_gemv_sub_t(uint32_t n, uint32_t ncol, const TV *__restrict__ V,
                                                    size_t ld, const double *__restrict__ h,
                                                    TV *__restrict__ w, double *partial, const int *stop) {
    __shared__ double sh[NC][4];
    if (stop && *stop) return;
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = blockIdx.x * chunk, r1 = min(n, r0 + chunk);
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    for (uint32_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        double v[NC];
        double a = 0.0;
        // loads unconditional (past the last column they reread it), so they
        // issue back to back; the products only over the real columns
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = (double)V[ld * min((uint32_t)c, ncol - 1) + r];
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if ((uint32_t)c < ncol) a += v[c] * h[c];
        const TV wt = (TV)((double)w[r] - a);
        w[r] = wt;
        const double wr = (double)wt;
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if ((uint32_t)c < ncol) acc[c] += v[c] * wr;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
        if ((uint32_t)c < ncol) {
            const double t = wsum(acc[c]);
            if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = t;
        }
    __syncthreads();
    if (threadIdx.x < ncol)
        partial[(size_t)threadIdx.x * gridDim.x + blockIdx.x] =
            ((sh[threadIdx.x][0] + sh[threadIdx.x][1]) + sh[threadIdx.x][2]) + sh[threadIdx.x][3];
}

// w -= V h1, then dest = V^T w: fused where the columns fit the registers
template <class TV>
void launch_gemv_sub_t(pp_ctx *ctx, uint32_t n, uint32_t ncol, const TV *V, size_t ld, const double *h1, TV *w,
                       double *dest) {
    if (ncol > (uint32_t)kFuseCols) {
        launch_gemv_sub(ctx, n, ncol, V, ld, h1, w);
        launch_gemv_t(ctx, n, ncol, V, ld, w, dest);
        return;
    }
    double *buf = gmres_coeff_buffer(ctx);
    double *partial = buf + 640;
    // register slots for the columns: 8, 16 or 32
    if (ncol <= 8)
        k_gemv_sub_t<8, TV><<<kGemvBlocks, 256, 0, ctx->stream>>>(n, ncol, V, ld, h1, w, partial, ctx->stop_word);
    else if (ncol <= 16)
        k_gemv_sub_t<16, TV><<<kGemvBlocks, 256, 0, ctx->stream>>>(n, ncol, V, ld, h1, w, partial, ctx->stop_word);
    else k_gemv_sub_t<32, TV><<<kGemvBlocks, 256, 0, ctx->stream>>>(n, ncol, V, ld, h1, w, partial, ctx->stop_word);
    k_gemv_t_reduce<<<ncol, 256, 0, ctx->stream>>>(kGemvBlocks, partial, dest ? dest : buf, ctx->stop_word);
    PP_HIP(hipGetLastError());
}
template void launch_gemv_t<double>(pp_ctx *, uint32_t, uint32_t, const double *, size_t, const double *, double *);
template void launch_gemv_t<float>(pp_ctx *, uint32_t, uint32_t, const float *, size_t, const float *, double *);
template void launch_gemv_sub<double>(pp_ctx *, uint32_t, uint32_t, const double *, size_t, const double *,
                                      double *);
template void launch_gemv_sub<float>(pp_ctx *, uint32_t, uint32_t, const float *, size_t, const double *, float *);
template void launch_gemv_sub_t<double>(pp_ctx *, uint32_t, uint32_t, const double *, size_t, const double *,
                                        double *, double *);
template void launch_gemv_sub_t<float>(pp_ctx *, uint32_t, uint32_t, const float *, size_t, const double *, float *,
                                       double *);


// x[r] += sum_c Z[r, c] y[c] (Z in FP32 or FP64), columns in ascending order
// like DenseMatrix::MatVec (DenseMatrix.cpp:208-224).  A column whose
// coefficient is zero adds nothing and is not read: the columns past a
// converged cycle's last step hold whatever the early-exit kernels left.
template <class T>
__global__ __launch_bounds__(256) void k_gemv_add(uint32_t n, uint32_t ncol, const T *__restrict__ Z, size_t ld,
                                                  const double *__restrict__ y, double *__restrict__ x) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (uint32_t c = 0; c < ncol; ++c)
            if (y[c] != 0.0) acc += (double)Z[ld * c + r] * y[c];
        x[r] += acc;
    }
}

void launch_gemv_add(pp_ctx *ctx, uint32_t n, uint32_t ncol, const float *Z32, const double *Z64, size_t ld,
                     const double *y, double *x) {
    unsigned g = std::min<uint32_t>((n + 255) / 256, 8192);
    if (Z32) k_gemv_add<float><<<g, 256, 0, ctx->stream>>>(n, ncol, Z32, ld, y, x);
    else k_gemv_add<double><<<g, 256, 0, ctx->stream>>>(n, ncol, Z64, ld, y, x);
    PP_HIP(hipGetLastError());
}

__global__ void k_scale_convert(uint32_t n, const double *s, double a, float *d) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        d[i] = (float)(s[i] * a);
}

void launch_scale_convert(pp_ctx *ctx, uint32_t n, const double *src, double scale, float *dst) {
    unsigned g = std::min<uint32_t>((n + 255) / 256, 4096);
    k_scale_convert<<<g, 256, 0, ctx->stream>>>(n, src, scale, dst);
    PP_HIP(hipGetLastError());
}

__global__ void k_axpy_f32_f64(uint32_t n, double a, const float *x, double *y) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[i] += a * (double)x[i];
}

void launch_axpy_f32_to_f64(pp_ctx *ctx, uint32_t n, double a, const float *x, double *y) {
    unsigned g = std::min<uint32_t>((n + 255) / 256, 4096);
    k_axpy_f32_f64<<<g, 256, 0, ctx->stream>>>(n, a, x, y);
    PP_HIP(hipGetLastError());
}

// first == -1: the GMRES coefficient buffer; else device scalar slots.
void read_slots(pp_ctx *ctx, int first, int count, double *out) {
    const double *src = first < 0 ? gmres_coeff_buffer(ctx) : ctx->scal + first;
    PP_CHECK(count <= 320, PP_ERR_INVALID, "too many slots");
    PP_HIP(hipMemcpyAsync(ctx->pinned, src, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    std::copy(ctx->pinned, ctx->pinned + count, out);
}

}  // namespace pp

// ---------------------------------------------------------------------------
// Device-resident least-squares of FlexibleGMRES(m): the Hessenberg column,
// the Givens rotations (GMRES.cpp:72-101), the convergence test and the
// back substitution (GMRES.cpp:103-119) run in single-thread kernels on the
// stream, so an Arnoldi cycle is issued without any host round trip; the
// stop word ends the cycle's remaining kernels once the estimate converges.
// ---------------------------------------------------------------------------
namespace pp {

GmresDev gmres_state(pp_ctx *ctx, uint32_t mm) {
    const size_t ld = mm + 1;
    const size_t total = ld * mm + 5 * ld + kGmMisc;
    double *base = (double *)ws_get(ctx, kWsGmresState, total * sizeof(double));
    GmresDev st;
    st.mm = mm;
    st.H = base;
    st.c = st.H + ld * mm;
    st.s = st.c + ld;
    st.g = st.s + ld;
    st.h1 = st.g + ld;
    st.h2 = st.h1 + ld;
    st.misc = st.h2 + ld;
    return st;
}

namespace {

__device__ void dev_givens(const GmresDev &st, uint32_t j) {
#pragma clang fp contract(off)
    const uint32_t ld = st.mm + 1;
    double *H = st.H;
    for (uint32_t i = 0; i < j; ++i) {
        const double h1 = H[(size_t)j * ld + i], h2 = H[(size_t)j * ld + i + 1];
        H[(size_t)j * ld + i + 1] = st.c[i] * h2 - st.s[i] * h1;
        H[(size_t)j * ld + i] = st.c[i] * h1 + st.s[i] * h2;
    }
    const double h1 = H[(size_t)j * ld + j], h2 = H[(size_t)j * ld + j + 1];
    if (fabs(h1) > fabs(h2)) {
        const double t = h2 / h1;
        st.c[j] = 1.0 / sqrt(fabs(1.0 + t * t));
        st.s[j] = st.c[j] * t;
    } else {
        const double t = h1 / h2;
        st.s[j] = 1.0 / sqrt(fabs(1.0 + t * t));
        st.c[j] = st.s[j] * t;
    }
    H[(size_t)j * ld + j] = st.c[j] * h1 + st.s[j] * h2;
    H[(size_t)j * ld + j + 1] = 0.0;
    st.g[j + 1] = -st.s[j] * st.g[j];
    st.g[j] = st.c[j] * st.g[j];
}

__global__ void k_gmres_cycle_init(GmresDev st, double beta, double bnorm, double tol, uint32_t hist_base) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (uint32_t i = 0; i < st.mm; ++i) st.g[i] = 0.0;  // g[mm] keeps its value (GMRES.cpp:160)
    st.g[0] = beta;
    st.misc[GM_M] = st.mm;
    st.misc[GM_STATUS] = 0.0;
    st.misc[GM_CONVCYC] = 0.0;
    st.misc[GM_NREC] = 0.0;
    st.misc[GM_HBASE] = (double)hist_base;
    st.misc[GM_BNORM] = bnorm;
    st.misc[GM_TOL] = tol;
}

__global__ void k_gmres_step(GmresDev st, uint32_t j, int quirks, const double *wn2, int *ctl, double *hist,
                             uint32_t hist_idx, uint32_t hist_cap) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (ctl[C_STOP]) return;
    const uint32_t ld = st.mm + 1;
    const double hn = sqrt(*wn2);
    for (uint32_t i = 0; i <= j; ++i) st.H[(size_t)j * ld + i] = st.h1[i] + st.h2[i];
    st.H[(size_t)j * ld + j + 1] = hn;
    const double bnorm = st.misc[GM_BNORM];
    if (!(hn == hn)) {
        st.misc[GM_STATUS] = PP_ERR_NOT_FINITE;
        st.misc[GM_M] = j;
        ctl[C_STOP] = 1;
        return;
    }
    if (fabs(hn) == 0.0) {  // happy breakdown (GMRES.cpp:183-186)
        st.misc[GM_M] = quirks ? j : j + 1;
        if (!quirks) {
            dev_givens(st, j);
            const double rnorm = fabs(st.g[j + 1]);
            st.misc[GM_RNORM] = rnorm;
            st.misc[GM_RELERR] = rnorm / bnorm;
            st.misc[GM_CONVCYC] = 1.0;
        }
        ctl[C_STOP] = 1;
        return;
    }
    st.misc[GM_INVH] = 1.0 / hn;
    dev_givens(st, j);
    const double rnorm = fabs(st.g[j + 1]);
    const double relerr = rnorm / bnorm;
    st.misc[GM_RNORM] = rnorm;
    st.misc[GM_RELERR] = relerr;
    // hist_idx counts from the cycle's first step (a captured cycle is replayed unchanged)
    const uint32_t hidx = (uint32_t)st.misc[GM_HBASE] + hist_idx;
    if (hist && hidx < hist_cap) hist[hidx] = relerr;
    st.misc[GM_NREC] += 1.0;
    if (relerr < st.misc[GM_TOL]) {
        st.misc[GM_M] = quirks ? j : j + 1;  // the reference drops column j
        st.misc[GM_CONVCYC] = 1.0;
        ctl[C_STOP] = 1;
    }
}

// v_{j+1} = (1/h_{j+1,j}) w, evaluated like HostVector::add with b = 0
template <class TV>
__global__ void k_gmres_next_basis(uint32_t n, const double *misc, const TV *__restrict__ w, TV *__restrict__ v,
                                   const int *stop) {
#pragma clang fp contract(off)
    if (stop && *stop) return;
    const double a = misc[GM_INVH];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double t1 = a * (double)w[i];
        const double t2 = 0.0 * (double)w[i];
        v[i] = (TV)(t1 + t2);
    }
}

__global__ void k_gmres_backsolve(GmresDev st, int quirks, double *y) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t ld = st.mm + 1;
    const uint32_t m = (uint32_t)st.misc[GM_M];
    for (int i = (int)m - 1; i >= 0; --i) {
        double t = st.g[i];
        for (uint32_t j = i + 1; j < m; ++j) t -= st.H[(size_t)j * ld + i] * st.g[j];
        st.g[i] = t / st.H[(size_t)i * ld + i];
    }
    st.g[m] = 0.0;
    for (uint32_t i = 0; i <= st.mm; ++i) y[i] = (!quirks && i >= m) ? 0.0 : st.g[i];
}

}  // namespace

void launch_gmres_cycle_init(pp_ctx *ctx, const GmresDev &st, double beta, double bnorm, double tol,
                             uint32_t hist_base) {
    k_gmres_cycle_init<<<1, 64, 0, ctx->stream>>>(st, beta, bnorm, tol, hist_base);
    PP_HIP(hipGetLastError());
}

void launch_gmres_step(pp_ctx *ctx, const GmresDev &st, uint32_t j, int quirks, double *hist, uint32_t hist_idx,
                       uint32_t hist_cap) {
    k_gmres_step<<<1, 64, 0, ctx->stream>>>(st, j, quirks, ctx->scal + S_TMP2, ctx->ctl, hist, hist_idx, hist_cap);
    PP_HIP(hipGetLastError());
}

template <class TV>
void launch_gmres_next_basis(pp_ctx *ctx, const GmresDev &st, uint32_t n, const TV *w, TV *vnext) {
    const unsigned g = std::min<uint32_t>((n + 255) / 256, 4096);
    k_gmres_next_basis<TV><<<g, 256, 0, ctx->stream>>>(n, st.misc, w, vnext, ctx->stop_word);
    PP_HIP(hipGetLastError());
}
template void launch_gmres_next_basis<double>(pp_ctx *, const GmresDev &, uint32_t, const double *, double *);
template void launch_gmres_next_basis<float>(pp_ctx *, const GmresDev &, uint32_t, const float *, float *);

void launch_gmres_backsolve(pp_ctx *ctx, const GmresDev &st, int quirks, double *y) {
    k_gmres_backsolve<<<1, 64, 0, ctx->stream>>>(st, quirks, y);
    PP_HIP(hipGetLastError());
}

}  // namespace pp
