// pp_kernels_krylov.hip -- tall-skinny GEMV kernels of the GMRES orthogonalisation
// (classical Gram-Schmidt, applied twice) and mixed-precision glue kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pp_internal.h"

namespace pp {

namespace {
constexpr int kGemvBlocks = 512;
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
    return (double *)ws_get(ctx, 31, (size_t)(640 + (size_t)kMaxCols * kGemvBlocks) * sizeof(double));
}

// partial[c][block] = sum over the block's rows of V[r, c] * w[r]
__global__ __launch_bounds__(256) void k_gemv_t_partial(uint32_t n, uint32_t ncol, const double *__restrict__ V,
                                                        size_t ld, const double *__restrict__ w, double *partial) {
    __shared__ double sh[4];
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = blockIdx.x * chunk, r1 = min(n, r0 + chunk);
    for (uint32_t c = 0; c < ncol; ++c) {
        const double *col = V + ld * c;
        double acc = 0.0;
        for (uint32_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) acc += col[r] * w[r];
        acc = wsum(acc);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) partial[(size_t)c * gridDim.x + blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
        __syncthreads();
    }
}

// h[c] = sum_b partial[c][b] in block order (deterministic)
__global__ __launch_bounds__(256) void k_gemv_t_reduce(uint32_t nb, const double *partial, double *h) {
    __shared__ double sh[4];
    const uint32_t c = blockIdx.x;
    double acc = 0.0;
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) acc += partial[(size_t)c * nb + b];
    acc = wsum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) h[c] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

void launch_gemv_t(pp_ctx *ctx, uint32_t n, uint32_t ncol, const double *V, size_t ld, const double *w, int) {
    PP_CHECK(ncol <= kMaxCols, PP_ERR_UNSUPPORTED, "too many columns");
    double *buf = gmres_coeff_buffer(ctx);
    double *partial = buf + 640;
    k_gemv_t_partial<<<kGemvBlocks, 256, 0, ctx->stream>>>(n, ncol, V, ld, w, partial);
    k_gemv_t_reduce<<<ncol, 256, 0, ctx->stream>>>(kGemvBlocks, partial, buf);
    PP_HIP(hipGetLastError());
}

// w[r] -= sum_c V[r, c] h[c]
__global__ __launch_bounds__(256) void k_gemv_sub(uint32_t n, uint32_t ncol, const double *__restrict__ V,
                                                  size_t ld, const double *__restrict__ h, double *__restrict__ w) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (uint32_t c = 0; c < ncol; ++c) acc += V[ld * c + r] * h[c];
        w[r] -= acc;
    }
}

void launch_gemv_sub(pp_ctx *ctx, uint32_t n, uint32_t ncol, const double *V, size_t ld, const double *h, double *w) {
    unsigned g = std::min<uint32_t>((n + 255) / 256, 8192);
    k_gemv_sub<<<g, 256, 0, ctx->stream>>>(n, ncol, V, ld, h, w);
    PP_HIP(hipGetLastError());
}

// x[r] += sum_c Z[r, c] y[c] (Z in FP32 or FP64), columns in ascending order
// like DenseMatrix::MatVec (DenseMatrix.cpp:208-224).
template <class T>
__global__ __launch_bounds__(256) void k_gemv_add(uint32_t n, uint32_t ncol, const T *__restrict__ Z, size_t ld,
                                                  const double *__restrict__ y, double *__restrict__ x) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (uint32_t c = 0; c < ncol; ++c) acc += (double)Z[ld * c + r] * y[c];
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
