// pp_multilevel.hip -- AMSED (SPD) / GMSLR (nonsymmetric) multilevel low-rank
// Schur-complement preconditioners on the device.
//
// With the hierarchical interface decomposition every level l splits the
// remaining matrix as A_l = [B_l E_l; F_l C_l] with block-diagonal B_l and
// C_l = A_{l+1}.  The apply (AMSEDPrecondition::levelCInvRecursiveSolve,
// source/Preconditioner/AMSEDPrecondition.cpp:1494-1587) is, on the device:
//     x1 = B^{-1} x1 ;  x2 -= F x1 ;  s = x2 ;  x2 = C^{-1} x2 (recursion)
//     x2 += W (H o (W^T s)) ;  x1 -= B^{-1} (E x2)
// B^{-1}: one level-scheduled / sync-free sweep pair over all blocks of the
// level at once (the blocks are independent, so their rows share levels).
// The low-rank term W H W^T comes from a generalised Lanczos process in the
// C-inner product whose C^{-1} is the already-built deeper recursion
// (AMSEDPrecondition.cpp:1275-1350, classic form without restarts); the
// tall-skinny product W = V Y runs on the FP64 matrix cores.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "pp_ml.h"

namespace pp {

void precond_apply(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y);

namespace {

constexpr int kLrBlocks = 256;

template <class T>
__global__ void k_gather(uint32_t n, const uint32_t *perm, const T *x, T *xp) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) xp[k] = x[perm[k]];
}
template <class T>
__global__ void k_scatter(uint32_t n, const uint32_t *perm, const T *xp, T *y) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) y[perm[k]] = xp[k];
}

// part[c][b] = sum over block b's rows of W[:, c] * s (accumulated in T
// within a thread, in FP64 across threads)
template <class T>
__global__ __launch_bounds__(256) void k_lr_partial(uint32_t n, uint32_t r, const T *__restrict__ W,
                                                    const T *__restrict__ s, double *part) {
    __shared__ double sh[4];
    const uint32_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t r0 = blockIdx.x * chunk, r1 = min(n, r0 + chunk);
    for (uint32_t c = 0; c < r; ++c) {
        T acct = 0;
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) acct += W[(size_t)c * n + i] * s[i];
        double acc = (double)acct;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) part[(size_t)c * gridDim.x + blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
        __syncthreads();
    }
}

// h[c] = H[c] * sum_b part[c][b]
__global__ __launch_bounds__(256) void k_lr_reduce(uint32_t nb, const double *part, const double *H, double *h) {
    __shared__ double sh[4];
    const uint32_t c = blockIdx.x;
    double acc = 0.0;
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) acc += part[(size_t)c * nb + b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) h[c] = H[c] * (((sh[0] + sh[1]) + sh[2]) + sh[3]);
}

// h = M t with M dense r x r (column-major); one workgroup
__global__ __launch_bounds__(64) void k_lr_dense(uint32_t r, const double *__restrict__ M, const double *__restrict__ t,
                                                 double *h) {
    for (uint32_t i = threadIdx.x; i < r; i += blockDim.x) {
        double acc = 0.0;
        for (uint32_t j = 0; j < r; ++j) acc += M[(size_t)j * r + i] * t[j];
        h[i] = acc;
    }
}

// x[i] += sum_c W[i, c] h[c] (W, x in T; h in FP64)
template <class T>
__global__ __launch_bounds__(256) void k_lr_add(uint32_t n, uint32_t r, const T *__restrict__ W,
                                                const double *__restrict__ h, T *__restrict__ x) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T acc = 0;
        for (uint32_t c = 0; c < r; ++c) acc += W[(size_t)c * n + i] * (T)h[c];
        x[i] += acc;
    }
}

__global__ void k_to_f32(size_t n, const double *s, float *d) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = (float)s[i];
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// W (n x r) = V (n x k) * Y (k x r), column-major, FP64 MFMA 16x16x4: one
// wavefront per 16 x 16 output tile, K in steps of 4.  Operand layout of
// v_mfma_f64_16x16x4f64: lane l holds A[l%16][l/16], B[l/16][l%16] and
// D[l/16 + 4*reg][l%16] (the FP64 output interleaves rows by 4, unlike the
// 32-bit 16x16 forms).
__global__ __launch_bounds__(256) void k_gemm_tall_f64(uint32_t n, uint32_t k, uint32_t r, const double *__restrict__ V,
                                                       const double *__restrict__ Y, double *__restrict__ W) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t ntr = (n + 15) / 16, ntc = (r + 15) / 16;
    if (wave >= ntr * ntc) return;
    const uint32_t tr = wave % ntr, tc = wave / ntr;
    const uint32_t i = lane & 15, q = lane >> 4;
    const uint32_t row = tr * 16 + i, col = tc * 16 + i;
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t k0 = 0; k0 < k; k0 += 4) {
        const uint32_t kk = k0 + q;
        const double a = (row < n && kk < k) ? V[(size_t)kk * n + row] : 0.0;
        const double b = (kk < k && col < r) ? Y[(size_t)col * k + kk] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    const uint32_t ocol = tc * 16 + i;
#pragma unroll
    for (int reg = 0; reg < 4; ++reg) {
        const uint32_t orow = tr * 16 + q + 4 * reg;
        if (orow < n && ocol < r) W[(size_t)ocol * n + orow] = acc[reg];
    }
}

void check_status(pp_ctx *ctx, int st) {
    if (st != PP_OK) throw Error(st, pp_last_error(ctx));
}

pp_csr *make_csr(pp_ctx *ctx, const HostCsr &M, uint32_t ncol) {
    pp_csr *A = nullptr;
    std::vector<uint32_t> ci(M.ci);
    std::vector<double> v(M.v);
    if (ci.empty()) {
        ci.push_back(0);
        v.push_back(0.0);
    }
    check_status(ctx, pp_csr_create(ctx, M.n, ncol, M.rp.data(), ci.data(), v.data(), &A));
    return A;
}

pp_precond *make_factor(pp_ctx *ctx, pp_csr *A, int kind, double tau, uint32_t lfil, int trsv) {
    pp_precond_opts o;
    pp_precond_opts_default(&o);
    o.kind = kind;
    o.tau = tau;
    o.lfil = lfil;
    o.trsv = trsv;
    pp_precond *P = nullptr;
    check_status(ctx, pp_precond_create(ctx, A, &o, &P));
    int st = pp_precond_setup(ctx, P);
    if (st != PP_OK) {
        std::string msg = pp_last_error(ctx);
        pp_precond_destroy(P);
        throw Error(st, msg);
    }
    return P;
}

template <class T>
struct MlBuf;
template <>
struct MlBuf<double> {
    static double *save(MlLevel &L) { return L.save; }
    static double *t(MlLevel &L) { return L.t; }
    static const double *W(MlLevel &L) { return L.W; }
};
template <>
struct MlBuf<float> {
    static float *save(MlLevel &L) { return L.save32; }
    static float *t(MlLevel &L) { return L.t32; }
    static const float *W(MlLevel &L) { return L.W32; }
};

// x (device, length n - lev_ptr[level+1]) <- C_level^{-1} x; level -1 is A.
// T = float: every operand in FP32 (factors, couplings, low-rank basis), the
// low-rank coefficients in FP64.
template <class T>
void rec_solve(pp_ctx *ctx, MlState &S, int level, T *x) {
    constexpr int DT = sizeof(T) == 8 ? PP_F64 : PP_F32;
    const int next = level + 1;
    if (next == (int)S.hid.nlev - 1) {
        precond_apply(ctx, S.last, DT, x, x);
        return;
    }
    MlLevel &L = S.lev[next];
    T *x1 = x, *x2 = x + L.nb;
    T *save = MlBuf<T>::save(L), *t = MlBuf<T>::t(L);
    precond_apply(ctx, L.Bpc, DT, x1, x1);
    launch_spmv(ctx, L.F, DT, -1.0, x1, 1.0, x2, -1, nullptr, false);  // x2 -= F x1
    // symmetric: S^{-1} ~ C^{-1} + W H W^T (on the input of C^{-1});
    // nonsymmetric: S^{-1} ~ (I + W H W^T) C^{-1} (on its output)
    if (L.rank && S.symmetric)
        PP_HIP(hipMemcpyAsync(save, x2, (size_t)L.nrest * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
    rec_solve<T>(ctx, S, next, x2);
    if (L.rank && !S.symmetric)
        PP_HIP(hipMemcpyAsync(save, x2, (size_t)L.nrest * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
    if (L.rank) {
        double *part = L.h + 64 * 1024;
        k_lr_partial<T><<<kLrBlocks, 256, 0, ctx->stream>>>(L.nrest, L.rank, MlBuf<T>::W(L), save, part);
        if (L.Md) {  // nonsymmetric: h = ((I - R)^{-1} - I) U^T s
            double *tt = L.h + 32 * 1024;
            k_lr_reduce<<<L.rank, 256, 0, ctx->stream>>>(kLrBlocks, part, L.Hd, tt);
            k_lr_dense<<<1, 64, 0, ctx->stream>>>(L.rank, L.Md, tt, L.h);
        } else {
            k_lr_reduce<<<L.rank, 256, 0, ctx->stream>>>(kLrBlocks, part, L.Hd, L.h);
        }
        const unsigned g = std::min<uint32_t>((L.nrest + 255) / 256, 8192);
        k_lr_add<T><<<std::max(1u, g), 256, 0, ctx->stream>>>(L.nrest, L.rank, MlBuf<T>::W(L), L.h, x2);  // x2 += W h
        PP_HIP(hipGetLastError());
    }
    launch_spmv(ctx, L.E, DT, 1.0, x2, 0.0, t, -1, nullptr, false);  // t = E x2
    precond_apply(ctx, L.Bpc, DT, t, t);
    launch_axpby(ctx, DT, L.nb, 1.0, x1, -1.0, t, x1);  // x1 -= t
}

double dot_dev(pp_ctx *ctx, uint32_t n, const double *a, const double *b) {
    launch_dot(ctx, PP_F64, n, a, b, S_TMP1);
    return read_slot(ctx, S_TMP1);
}

// Generalised Lanczos for E^T B^{-1} E w = theta C w at level l (symmetric
// case), C^{-1} by the recursion below l.  Full reorthogonalisation against
// the stored basis.  Returns the Ritz values (ascending) and sets W, H.
void lowrank_lanczos(pp_ctx *ctx, MlState &S, int l, uint32_t rank, uint32_t steps) {
    MlLevel &L = S.lev[l];
    const uint32_t n = L.nrest;
    uint32_t k = std::min<uint32_t>(std::min<uint32_t>(steps, n), 250);
    if (rank == 0 || n == 0 || k == 0) return;
    double *V = dalloc<double>((size_t)n * (k + 1));
    double *Z = dalloc<double>((size_t)n * (k + 1));
    double *u = dalloc<double>(n), *w = dalloc<double>(n), *tb = dalloc<double>(std::max(1u, L.nb));
    std::vector<double> alpha, beta;
    try {
        // start vector: seeded uniform(-1, 1)
        std::vector<double> u0(n);
        PP_CHECK(pp_fill_random(0, 12345 + l, n, u0.data()) == PP_OK, PP_ERR_INVALID, "rng");
        PP_HIP(hipMemcpy(u, u0.data(), n * 8, hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(w, u, n * 8, hipMemcpyDeviceToDevice));
        rec_solve(ctx, S, l, w);
        double b0 = dot_dev(ctx, n, w, u);
        PP_CHECK(b0 > 0 && std::isfinite(b0), PP_ERR_BREAKDOWN, "Lanczos start vector has non-positive C-norm");
        b0 = std::sqrt(b0);
        launch_axpby(ctx, PP_F64, n, 1.0 / b0, w, 0.0, w, V);
        launch_axpby(ctx, PP_F64, n, 1.0 / b0, u, 0.0, u, Z);
        uint32_t done = 0;
        for (uint32_t j = 0; j < k; ++j) {
            double *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n;
            // u = F B^{-1} E v_j  (F = E^T)
            launch_spmv(ctx, L.E, PP_F64, 1.0, vj, 0.0, tb, -1, nullptr, false);
            precond_apply(ctx, L.Bpc, PP_F64, tb, tb);
            launch_spmv(ctx, L.F, PP_F64, 1.0, tb, 0.0, u, -1, nullptr, false);
            if (j > 0) launch_axpby(ctx, PP_F64, n, 1.0, u, -beta[j - 1], Z + (size_t)(j - 1) * n, u);
            const double a = dot_dev(ctx, n, u, vj);
            alpha.push_back(a);
            launch_axpby(ctx, PP_F64, n, 1.0, u, -a, zj, u);
            // u -= Z (V^T u), twice for stability
            for (int pass = 0; pass < 2; ++pass) {
                launch_gemv_t(ctx, n, j + 1, V, n, u, nullptr);
                launch_gemv_sub(ctx, n, j + 1, Z, n, gmres_coeff_buffer(ctx), u);
            }
            done = j + 1;
            PP_HIP(hipMemcpyAsync(w, u, n * 8, hipMemcpyDeviceToDevice, ctx->stream));
            rec_solve(ctx, S, l, w);
            double bt = dot_dev(ctx, n, w, u);
            if (!(bt > 1e-28) || j + 1 == k) break;  // C-norm collapsed (AMSEDPrecondition.cpp:1334-1339)
            bt = std::sqrt(bt);
            beta.push_back(bt);
            launch_axpby(ctx, PP_F64, n, 1.0 / bt, w, 0.0, w, V + (size_t)(j + 1) * n);
            launch_axpby(ctx, PP_F64, n, 1.0 / bt, u, 0.0, u, Z + (size_t)(j + 1) * n);
        }
        const uint32_t m = done;
        std::vector<double> d(alpha.begin(), alpha.begin() + m), e(beta.begin(), beta.begin() + (m ? m - 1 : 0));
        std::vector<double> Yall;
        tridiag_eig(d, e, Yall);
        // largest Ritz values in [0, 1): theta/(1 - theta) (the basic correction)
        std::vector<uint32_t> pick;
        for (int i = (int)m - 1; i >= 0 && pick.size() < rank; --i)
            if (d[i] > 0.0 && d[i] < 1.0 - 1e-12) pick.push_back((uint32_t)i);
        L.rank = (uint32_t)pick.size();
        L.theta.assign(d.begin(), d.end());
        if (L.rank) {
            std::vector<double> Y((size_t)m * L.rank);
            L.Hhost.resize(L.rank);
            for (uint32_t c = 0; c < L.rank; ++c) {
                const double th = d[pick[c]];
                L.Hhost[c] = th / (1.0 - th);
                std::copy(Yall.begin() + (size_t)pick[c] * m, Yall.begin() + (size_t)(pick[c] + 1) * m,
                          Y.begin() + (size_t)c * m);
            }
            double *Yd = dalloc<double>(Y.size());
            PP_HIP(hipMemcpy(Yd, Y.data(), Y.size() * 8, hipMemcpyHostToDevice));
            L.W = dalloc<double>((size_t)n * L.rank);
            launch_gemm_tall_f64(ctx, n, m, L.rank, V, Yd, L.W);
            PP_HIP(hipStreamSynchronize(ctx->stream));
            dfree(Yd);
            L.Hd = dalloc<double>(L.rank);
            PP_HIP(hipMemcpy(L.Hd, L.Hhost.data(), L.rank * 8, hipMemcpyHostToDevice));
            L.save = dalloc<double>(n);
            L.h = dalloc<double>(64 * 1024 + (size_t)L.rank * kLrBlocks);
        }
    } catch (...) {
        dfree(V); dfree(Z); dfree(u); dfree(w); dfree(tb);
        throw;
    }
    dfree(V); dfree(Z); dfree(u); dfree(w); dfree(tb);
}

// ---------------------------------------------------------------------------
// The reference's low-rank setup of AMSED: thick-restart (Krylov-Schur)
// generalised Lanczos, AMSEDPrecondition.cpp:1010-1273, 1276-1345, 1382-1470.
//   m = 2 rank steps (AMSED_LANCZOS_EXPAND_COEFFICIENT), extended to
//   k = 4 rank (AMSED_LANCZOS_MAX_COEFFICIENT) before each of at most three
//   restarts; a restart keeps the m largest Ritz pairs (V <- V Y_m), turns
//   diag(theta_m) bordered by beta_k Y(k, 1:m) back into a tridiagonal matrix
//   by the reference's chain of Givens rotations (applied to the kept basis
//   too), and continues the three-term recurrence from the old residual
//   vector; it stops early when |beta_k Y(k, k)| <= 1e-30.
// The low-rank term from the final Ritz pairs mu (largest first):
//   PP_LOWRANK_REFERENCE  rank pairs, H = mu / (1 - mu)        (AMSEDBasic)
//   PP_LOWRANK_DEFLATION  the pairs with mu > 1 - eps (at most lanczos_cap),
//                         H = (eps - lam) / lam, lam = 1 - mu  (AMSEDEigenvalueDeflation)
// The basis V (C-orthonormal) and Z = C V live on the device; the products
// V Y and V Q are MFMA GEMMs, the small tridiagonal work is on the host.
// ---------------------------------------------------------------------------
namespace {

constexpr double kAmsedExitZero = 1e-15;          // AMSED_EXIT_ZERO
constexpr uint32_t kAmsedMsteps = 30;             // AMSED_MSTEPS: the deflated pairs at most
constexpr double kAmsedHighExitZero = 1e-30;      // AMSED_HIGH_PRECISION_EXIT_ZERO

// Lanczos steps j = from .. to-1 of the generalised problem
// E^T B^{-1} E v = theta C v (expandReorthogonalLanczos): alpha_j = D[j],
// beta_j = E[j]; returns the number of steps whose vectors exist.
uint32_t lanczos_expand(pp_ctx *ctx, MlState &S, int l, double *V, double *Z, std::vector<double> &D,
                        std::vector<double> &E, uint32_t from, uint32_t to, double *u, double *w, double *tb) {
    MlLevel &L = S.lev[l];
    const uint32_t n = L.nrest;
    uint32_t actual = from;
    for (uint32_t j = from; j < to; ++j) {
        if (j > 0 && E[j - 1] <= kAmsedExitZero) return actual;
        ++actual;
        double *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n;
        launch_spmv(ctx, L.E, PP_F64, 1.0, vj, 0.0, tb, -1, nullptr, false);
        precond_apply(ctx, L.Bpc, PP_F64, tb, tb);
        launch_spmv(ctx, L.F, PP_F64, 1.0, tb, 0.0, u, -1, nullptr, false);
        if (j > 0) launch_axpby(ctx, PP_F64, n, 1.0, u, -E[j - 1], Z + (size_t)(j - 1) * n, u);
        D[j] = dot_dev(ctx, n, u, vj);
        launch_axpby(ctx, PP_F64, n, 1.0, u, -D[j], zj, u);
        // full reorthogonalisation u -= Z (V^T u), twice
        for (int pass = 0; pass < 2; ++pass) {
            launch_gemv_t(ctx, n, j + 1, V, n, u, nullptr);
            launch_gemv_sub(ctx, n, j + 1, Z, n, gmres_coeff_buffer(ctx), u);
        }
        PP_HIP(hipMemcpyAsync(w, u, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        rec_solve(ctx, S, l, w);  // w = C^{-1} u
        double bt = dot_dev(ctx, n, w, u);
        if (!(bt >= 0.0)) {  // approximate B^{-1} and C^{-1} can make it negative: stop
            E[j] = kAmsedExitZero;
            break;
        }
        bt = std::sqrt(bt);
        E[j] = bt;
        if (bt <= kAmsedExitZero) break;
        launch_axpby(ctx, PP_F64, n, 1.0 / bt, w, 0.0, w, V + (size_t)(j + 1) * n);
        launch_axpby(ctx, PP_F64, n, 1.0 / bt, u, 0.0, u, Z + (size_t)(j + 1) * n);
    }
    return actual;
}

// Columns c1, c2 of the column-major rows x ? matrix Q: [c1 c2] <- [c1 c + c2 s, -c1 s + c2 c]
void rotate_columns(std::vector<double> &Q, uint32_t rows, uint32_t c1, uint32_t c2, double c, double s) {
    double *a = Q.data() + (size_t)c1 * rows, *b = Q.data() + (size_t)c2 * rows;
    for (uint32_t r = 0; r < rows; ++r) {
        const double x = a[r], y = b[r];
        a[r] = x * c + y * s;
        b[r] = -1.0 * x * s + y * c;
    }
}

// diag(mu[0..m)) bordered by the row z back to a tridiagonal (D, E) by the
// reference's rotation chain (convertDiagonal2TriDiagonal), each rotation
// also applied to the columns of Q (rows x m); z ends with its weight in its
// last entry.
void diag_to_tridiag(const std::vector<double> &mu, std::vector<double> &z, uint32_t m, std::vector<double> &D,
                     std::vector<double> &E, std::vector<double> &Q, uint32_t rows) {
    for (uint32_t i = 0; i < m; ++i) D[i] = mu[i];
    for (uint32_t i = 0; i + 1 < m; ++i) E[i] = 0.0;
    for (uint32_t i = 0; i + 1 < m; ++i) {
        const double zi = z[i], zi1 = z[i + 1];
        const double delta = std::sqrt(zi * zi + zi1 * zi1);
        double c = zi1 * 1.0 / delta, s = -1.0 * zi / delta;
        z[i + 1] = delta;
        z[i] = 0;
        double cc = c * c, ss = s * s, cs = c * s;
        {
            const double ai = D[i], bi = E[i], ai1 = D[i + 1], t = 2 * cs * bi;
            D[i] = cc * ai + t + ss * ai1;
            E[i] = -1.0 * cs * ai + (cc - ss) * bi + cs * ai1;
            D[i + 1] = ss * ai - t + cc * ai1;
        }
        rotate_columns(Q, rows, i, i + 1, c, s);
        if (i > 0) {
            double gamma = -1.0 * s * E[i - 1];
            E[i - 1] *= c;
            for (uint32_t r = i; r >= 1; --r) {
                const double ar1 = D[r - 1], br1 = E[r - 1], br = E[r], ar = D[r];
                const double sigma = std::sqrt(gamma * gamma + br * br);
                c = br * 1.0 / sigma;
                s = -1.0 * gamma / sigma;
                cc = c * c;
                ss = s * s;
                cs = c * s;
                const double t = 2 * cs * br1;
                D[r - 1] = cc * ar1 + t + ss * ar;
                E[r - 1] = -1.0 * cs * ar1 + (cc - ss) * br1 + cs * ar;
                D[r] = ss * ar1 - t + cc * ar;
                E[r] = sigma;
                if (r > 1) {
                    gamma = -1.0 * s * E[r - 2];
                    E[r - 2] *= c;
                }
                rotate_columns(Q, rows, r - 1, r, c, s);
            }
        }
    }
}

// eigenpairs of the tridiagonal (D[0..k), E[0..k-1)): mu ascending, Y k x k
void tridiag_pairs(const std::vector<double> &D, const std::vector<double> &E, uint32_t k, std::vector<double> &mu,
                   std::vector<double> &Y) {
    mu.assign(D.begin(), D.begin() + k);
    std::vector<double> e(E.begin(), E.begin() + (k ? k - 1 : 0));
    tridiag_eig(mu, e, Y);
}

}  // namespace

void lowrank_lanczos_restarted(pp_ctx *ctx, MlState &S, int l, uint32_t rank, int mode, double eps,
                               uint32_t corr_cap) {
    MlLevel &L = S.lev[l];
    const uint32_t n = L.nrest;
    if (rank == 0 || n < 2) return;
    // m and k of the reference, kept within the level's size
    const uint32_t kk = std::min<uint32_t>(4 * rank, n), mm = std::min<uint32_t>(2 * rank, kk);
    const uint32_t restarts = 3;
    double *V = dalloc<double>((size_t)n * (kk + 1));
    double *Z = dalloc<double>((size_t)n * (kk + 1));
    double *T = dalloc<double>((size_t)n * mm);
    double *u = dalloc<double>(n), *w = dalloc<double>(n), *tb = dalloc<double>(std::max(1u, L.nb));
    double *Qd = dalloc<double>((size_t)(kk + 1) * (kk + 1));
    std::vector<double> D(kk + 1, 0.0), E(kk + 1, 0.0), mu, Y;
    try {
        // start: a seeded random z1, v1 = C^{-1} z1 scaled to v1^T C v1 = 1
        std::vector<double> u0(n);
        PP_CHECK(pp_fill_random(0, 12345 + l, n, u0.data()) == PP_OK, PP_ERR_INVALID, "rng");
        PP_HIP(hipMemcpy(u, u0.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(w, u, (size_t)n * 8, hipMemcpyDeviceToDevice));
        rec_solve(ctx, S, l, w);
        double b0 = dot_dev(ctx, n, w, u);
        PP_CHECK(b0 > 0 && std::isfinite(b0), PP_ERR_BREAKDOWN, "Lanczos start vector has non-positive C-norm");
        b0 = std::sqrt(b0);
        launch_axpby(ctx, PP_F64, n, 1.0 / b0, w, 0.0, w, V);
        launch_axpby(ctx, PP_F64, n, 1.0 / b0, u, 0.0, u, Z);
        uint32_t actual = lanczos_expand(ctx, S, l, V, Z, D, E, 0, mm, u, w, tb);
        bool have_pairs = false;
        for (uint32_t rs = 0; rs < restarts; ++rs) {
            if (actual < mm) {  // ended early: the space is invariant
                tridiag_pairs(D, E, actual, mu, Y);
                have_pairs = true;
                break;
            }
            actual = lanczos_expand(ctx, S, l, V, Z, D, E, mm, kk, u, w, tb);
            const double betak = E[actual - 1];
            tridiag_pairs(D, E, actual, mu, Y);
            have_pairs = true;
            const uint32_t last = actual - 1;
            if (std::fabs(betak) * std::fabs(Y[(size_t)last * actual + last]) <= kAmsedHighExitZero ||
                rs == restarts - 1 || actual <= mm)
                break;
            // keep the mm largest pairs (largest first), V <- V Y_m Q, Z likewise
            std::vector<double> mud(mm), Qm((size_t)actual * mm), zrow(mm);
            for (uint32_t c = 0; c < mm; ++c) {
                const uint32_t src = actual - 1 - c;
                mud[c] = mu[src];
                std::copy(Y.begin() + (size_t)src * actual, Y.begin() + (size_t)(src + 1) * actual,
                          Qm.begin() + (size_t)c * actual);
                zrow[c] = Y[(size_t)src * actual + last];
            }
            diag_to_tridiag(mud, zrow, mm, D, E, Qm, actual);
            // the rotations gather the border into its last entry: beta_m = z_m beta_k
            E[mm - 1] = zrow[mm - 1] * betak;
            PP_HIP(hipMemcpy(Qd, Qm.data(), Qm.size() * 8, hipMemcpyHostToDevice));
            launch_gemm_tall_f64(ctx, n, actual, mm, V, Qd, T);
            PP_HIP(hipMemcpyAsync(V, T, (size_t)n * mm * 8, hipMemcpyDeviceToDevice, ctx->stream));
            launch_gemm_tall_f64(ctx, n, actual, mm, Z, Qd, T);
            PP_HIP(hipMemcpyAsync(Z, T, (size_t)n * mm * 8, hipMemcpyDeviceToDevice, ctx->stream));
            PP_HIP(hipMemcpyAsync(V + (size_t)mm * n, V + (size_t)actual * n, (size_t)n * 8, hipMemcpyDeviceToDevice,
                                  ctx->stream));
            PP_HIP(hipMemcpyAsync(Z + (size_t)mm * n, Z + (size_t)actual * n, (size_t)n * 8, hipMemcpyDeviceToDevice,
                                  ctx->stream));
            actual = mm;
        }
        if (!have_pairs) tridiag_pairs(D, E, actual, mu, Y);
        // the low-rank term from the largest Ritz pairs
        std::vector<uint32_t> pick;
        std::vector<double> hval;
        if (mode == PP_LOWRANK_DEFLATION) {
            const double para = 1.0 - eps;
            for (uint32_t i = 0; i < actual; ++i) {
                const double m_i = mu[actual - 1 - i];
                if (m_i <= para || pick.size() >= corr_cap) break;
                const double lam = 1.0 - m_i;
                pick.push_back(actual - 1 - i);
                hval.push_back((eps - lam) / lam);
            }
        } else {
            for (uint32_t i = 0; i < std::min(actual, rank); ++i) {
                const double m_i = mu[actual - 1 - i];
                if (!(m_i < 1.0 - 1e-12)) continue;  // mu >= 1 would make mu / (1 - mu) infinite or negative
                pick.push_back(actual - 1 - i);
                hval.push_back(m_i / (1.0 - m_i));
            }
        }
        L.rank = (uint32_t)pick.size();
        L.theta.assign(mu.begin(), mu.end());
        if (L.rank) {
            std::vector<double> Ysel((size_t)actual * L.rank);
            for (uint32_t c = 0; c < L.rank; ++c)
                std::copy(Y.begin() + (size_t)pick[c] * actual, Y.begin() + (size_t)(pick[c] + 1) * actual,
                          Ysel.begin() + (size_t)c * actual);
            PP_HIP(hipMemcpy(Qd, Ysel.data(), Ysel.size() * 8, hipMemcpyHostToDevice));
            L.W = dalloc<double>((size_t)n * L.rank);
            launch_gemm_tall_f64(ctx, n, actual, L.rank, V, Qd, L.W);
            L.Hhost = hval;
            L.Hd = dalloc<double>(L.rank);
            PP_HIP(hipMemcpy(L.Hd, L.Hhost.data(), L.rank * 8, hipMemcpyHostToDevice));
            L.save = dalloc<double>(n);
            L.h = dalloc<double>(64 * 1024 + (size_t)L.rank * kLrBlocks);
        }
        PP_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        dfree(V); dfree(Z); dfree(T); dfree(u); dfree(w); dfree(tb); dfree(Qd);
        throw;
    }
    dfree(V); dfree(Z); dfree(T); dfree(u); dfree(w); dfree(tb); dfree(Qd);
}

// I - R for the r x r column-major R
std::vector<double> identity_minus(const std::vector<double> &R, uint32_t r) {
    std::vector<double> A(R.size());
    for (size_t i = 0; i < R.size(); ++i) A[i] = -R[i];
    for (uint32_t i = 0; i < r; ++i) A[(size_t)i * r + i] += 1.0;
    return A;
}

// X = A^{-1} B for A r x r and B r x nrhs (column-major), Gaussian
// elimination with partial pivoting
std::vector<double> solve_dense(std::vector<double> A, uint32_t r, std::vector<double> B, uint32_t nrhs) {
    for (uint32_t k = 0; k < r; ++k) {
        uint32_t p = k;
        for (uint32_t i = k + 1; i < r; ++i)
            if (std::fabs(A[(size_t)k * r + i]) > std::fabs(A[(size_t)k * r + p])) p = i;
        PP_CHECK(A[(size_t)k * r + p] != 0.0, PP_ERR_BREAKDOWN, "singular low-rank core I - R");
        if (p != k) {
            for (uint32_t j = 0; j < r; ++j) std::swap(A[(size_t)j * r + k], A[(size_t)j * r + p]);
            for (uint32_t j = 0; j < nrhs; ++j) std::swap(B[(size_t)j * r + k], B[(size_t)j * r + p]);
        }
        for (uint32_t i = k + 1; i < r; ++i) {
            const double f = A[(size_t)k * r + i] / A[(size_t)k * r + k];
            for (uint32_t j = k; j < r; ++j) A[(size_t)j * r + i] -= f * A[(size_t)j * r + k];
            for (uint32_t j = 0; j < nrhs; ++j) B[(size_t)j * r + i] -= f * B[(size_t)j * r + k];
        }
    }
    for (uint32_t j = 0; j < nrhs; ++j)
        for (int i = (int)r - 1; i >= 0; --i) {
            double v = B[(size_t)j * r + i];
            for (uint32_t c = i + 1; c < r; ++c) v -= A[(size_t)c * r + i] * B[(size_t)j * r + c];
            B[(size_t)j * r + i] = v / A[(size_t)i * r + i];
        }
    return B;
}

// Low-rank term of the nonsymmetric case: Arnoldi on G = C^{-1} F B^{-1} E
// (C^{-1} by the recursion below l; CGS2 orthogonalisation), eigenvalues of
// the Hessenberg matrix by double-shift QR, eigenvectors of the largest real
// eigenvalues in [0, 1) by inverse iteration; W = V Y, H = lambda / (1 - lambda)
// (GMSLRDPrecondition.cpp:980-1036 takes leading Schur vectors of H with a
// diagonal core; here the core R = Q^T H Q of the selected invariant subspace
// is kept whole, so the correction is exact on that subspace).
// schur_reference: the reference's selection instead (GMSLRDPrecondition.cpp:
// 1008-1032): the real Schur form H_m = Z T Z^T, the first min(m, rank)
// diagonal entries of T in the order the Schur form holds them (no sorting)
// and their Schur vectors, W = V Z(:, 1:r) and a diagonal core
// H = lambda / (1 - lambda), largest index first as the reference stores it.
void lowrank_arnoldi(pp_ctx *ctx, MlState &S, int l, uint32_t rank, uint32_t steps, bool schur_reference) {
    MlLevel &L = S.lev[l];
    const uint32_t n = L.nrest;
    const uint32_t k = std::min<uint32_t>(std::min<uint32_t>(steps, n), 64);
    if (rank == 0 || n == 0 || k == 0) return;
    double *V = dalloc<double>((size_t)n * (k + 1));
    double *w = dalloc<double>(n), *tb = dalloc<double>(std::max(1u, L.nb));
    double *buf = gmres_coeff_buffer(ctx);
    try {
        std::vector<double> v0(n, 1.0 / std::sqrt((double)n));
        PP_HIP(hipMemcpy(V, v0.data(), n * 8, hipMemcpyHostToDevice));
        const uint32_t ld = k + 1;
        std::vector<double> Hh((size_t)ld * k, 0.0), h1(k + 1), h2(k + 1);
        uint32_t m = 0;
        for (uint32_t j = 0; j < k; ++j) {
            launch_spmv(ctx, L.E, PP_F64, 1.0, V + (size_t)j * n, 0.0, tb, -1, nullptr, false);
            precond_apply(ctx, L.Bpc, PP_F64, tb, tb);
            launch_spmv(ctx, L.F, PP_F64, 1.0, tb, 0.0, w, -1, nullptr, false);
            rec_solve(ctx, S, l, w);
            for (int pass = 0; pass < 2; ++pass) {
                launch_gemv_t(ctx, n, j + 1, V, n, w, nullptr);
                PP_HIP(hipMemcpyAsync(buf + 320, buf, (j + 1) * 8, hipMemcpyDeviceToDevice, ctx->stream));
                launch_gemv_sub(ctx, n, j + 1, V, n, buf + 320, w);
                read_slots(ctx, -1, j + 1, pass == 0 ? h1.data() : h2.data());
            }
            for (uint32_t i = 0; i <= j; ++i) Hh[(size_t)j * ld + i] = h1[i] + h2[i];
            const double hn = std::sqrt(dot_dev(ctx, n, w, w));
            m = j + 1;
            Hh[(size_t)j * ld + j + 1] = hn;
            if (!(hn > 1e-12) || j + 1 == k) break;
            launch_axpby(ctx, PP_F64, n, 1.0 / hn, w, 0.0, w, V + (size_t)(j + 1) * n);
        }
        std::vector<double> Hm((size_t)m * m);
        for (uint32_t c = 0; c < m; ++c)
            for (uint32_t r = 0; r < m; ++r) Hm[(size_t)c * m + r] = Hh[(size_t)c * ld + r];
        if (schur_reference) {
            std::vector<double> T = Hm, Zs, wr, wi;
            PP_CHECK(hessenberg_schur(T, (int)m, Zs, wr, wi), PP_ERR_BREAKDOWN,
                     "Hessenberg Schur iteration did not converge");
            const uint32_t r = std::min<uint32_t>(m, rank);
            L.theta.assign(wr.begin(), wr.end());
            std::vector<double> Zsel((size_t)m * r);
            std::vector<double> hv(r);
            for (uint32_t c = 0; c < r; ++c) {
                const uint32_t i = r - 1 - c;  // i = r-1 .. 0 into columns 0 .. r-1
                hv[c] = wr[i] / (1.0 - wr[i]);
                std::copy(Zs.begin() + (size_t)i * m, Zs.begin() + (size_t)(i + 1) * m, Zsel.begin() + (size_t)c * m);
            }
            L.rank = r;
            if (r) {
                double *Qd = dalloc<double>(Zsel.size());
                PP_HIP(hipMemcpy(Qd, Zsel.data(), Zsel.size() * 8, hipMemcpyHostToDevice));
                L.W = dalloc<double>((size_t)n * r);
                launch_gemm_tall_f64(ctx, n, m, r, V, Qd, L.W);
                PP_HIP(hipStreamSynchronize(ctx->stream));
                dfree(Qd);
                L.Hhost = hv;
                L.Hd = dalloc<double>(r);
                PP_HIP(hipMemcpy(L.Hd, hv.data(), r * 8, hipMemcpyHostToDevice));
                L.save = dalloc<double>(n);
                L.h = dalloc<double>(64 * 1024 + (size_t)r * kLrBlocks);
            }
            dfree(V); dfree(w); dfree(tb);
            return;
        }
        std::vector<double> re, im;
        PP_CHECK(hessenberg_eigenvalues(Hm, (int)m, re, im), PP_ERR_BREAKDOWN,
                 "Hessenberg eigenvalue iteration did not converge");
        std::vector<uint32_t> idx;
        for (uint32_t i = 0; i < m; ++i)
            if (std::fabs(im[i]) <= 1e-10 * std::max(1.0, std::fabs(re[i])) && re[i] < 1.0 - 1e-12)
                idx.push_back(i);
        std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return re[a] > re[b]; });
        if (idx.size() > rank) idx.resize(rank);
        L.theta.assign(re.begin(), re.end());
        // orthonormal basis Q of the eigenvectors' span (modified Gram-Schmidt,
        // dependent vectors dropped): an invariant subspace of H_m
        std::vector<double> Q;
        for (uint32_t c = 0; c < idx.size(); ++c) {
            std::vector<double> y = inverse_iteration(Hm, (int)m, re[idx[c]]);
            const uint32_t q = (uint32_t)(Q.size() / m);
            for (int pass = 0; pass < 2; ++pass)
                for (uint32_t j = 0; j < q; ++j) {
                    double d = 0.0;
                    for (uint32_t i = 0; i < m; ++i) d += Q[(size_t)j * m + i] * y[i];
                    for (uint32_t i = 0; i < m; ++i) y[i] -= d * Q[(size_t)j * m + i];
                }
            double nr = 0.0;
            for (double v : y) nr += v * v;
            nr = std::sqrt(nr);
            if (!(nr > 1e-8)) continue;
            for (double v : y) Q.push_back(v / nr);
        }
        const uint32_t r = (uint32_t)(Q.size() / m);
        L.rank = r;
        if (r) {
            // R = Q^T H_m Q, core K = (I - R)^{-1} - I = (I - R)^{-1} R
            std::vector<double> HQ((size_t)m * r, 0.0), R((size_t)r * r, 0.0);
            for (uint32_t c = 0; c < r; ++c)
                for (uint32_t j = 0; j < m; ++j)
                    for (uint32_t i = 0; i < m; ++i) HQ[(size_t)c * m + i] += Hm[(size_t)j * m + i] * Q[(size_t)c * m + j];
            for (uint32_t c = 0; c < r; ++c)
                for (uint32_t i = 0; i < r; ++i) {
                    double d = 0.0;
                    for (uint32_t j = 0; j < m; ++j) d += Q[(size_t)i * m + j] * HQ[(size_t)c * m + j];
                    R[(size_t)c * r + i] = d;
                }
            std::vector<double> K = solve_dense(identity_minus(R, r), r, R, r);
            L.Hhost.resize(r);
            for (uint32_t c = 0; c < r; ++c) L.Hhost[c] = R[(size_t)c * r + c] / (1.0 - R[(size_t)c * r + c]);
            double *Qd = dalloc<double>(Q.size());
            PP_HIP(hipMemcpy(Qd, Q.data(), Q.size() * 8, hipMemcpyHostToDevice));
            L.W = dalloc<double>((size_t)n * r);
            launch_gemm_tall_f64(ctx, n, m, r, V, Qd, L.W);
            PP_HIP(hipStreamSynchronize(ctx->stream));
            dfree(Qd);
            std::vector<double> ones(r, 1.0);
            L.Hd = dalloc<double>(r);
            PP_HIP(hipMemcpy(L.Hd, ones.data(), r * 8, hipMemcpyHostToDevice));
            L.Md = dalloc<double>((size_t)r * r);
            PP_HIP(hipMemcpy(L.Md, K.data(), (size_t)r * r * 8, hipMemcpyHostToDevice));
            L.save = dalloc<double>(n);
            L.h = dalloc<double>(64 * 1024 + (size_t)r * kLrBlocks);
        }
    } catch (...) {
        dfree(V); dfree(w); dfree(tb);
        throw;
    }
    dfree(V); dfree(w); dfree(tb);
}

}  // namespace

void launch_gemm_tall_f64(pp_ctx *ctx, uint32_t n, uint32_t k, uint32_t r, const double *V, const double *Y,
                          double *W) {
    const uint64_t waves = (uint64_t)((n + 15) / 16) * ((r + 15) / 16);
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    if (blocks == 0) return;
    k_gemm_tall_f64<<<blocks, 256, 0, ctx->stream>>>(n, k, r, V, Y, W);
    PP_HIP(hipGetLastError());
}

void ml_free(pp_precond *M) {
    auto *S = static_cast<MlState *>(M->multilevel);
    if (!S) return;
    for (auto &L : S->lev) {
        pp_precond_destroy(L.Bpc);
        pp_csr_destroy(L.B);
        pp_csr_destroy(L.E);
        pp_csr_destroy(L.F);
        dfree(L.W); dfree(L.Hd); dfree(L.Md); dfree(L.save); dfree(L.t); dfree(L.h);
        dfree(L.W32); dfree(L.save32); dfree(L.t32);
    }
    pp_precond_destroy(S->last);
    pp_csr_destroy(S->Clast);
    dfree(S->dperm);
    dfree(S->xp);
    dfree(S->xp32);
    delete S;
    M->multilevel = nullptr;
}

// setup(): partition, factor every level's B blocks and the last level, then
// build the low-rank corrections bottom-up (each uses the recursion below it).
void ml_setup(pp_ctx *ctx, pp_precond *M) {
    const bool spd = M->opts.kind == PP_PC_AMSED;
    HostCsr A;
    A.n = M->A->nrow;
    A.rp = M->A->h_rp;
    A.ci = M->A->h_ci;
    A.v = M->A->h_v;
    auto *S = new MlState;
    M->multilevel = S;
    try {
        S->symmetric = spd;
        S->n = A.n;
        const uint32_t nlev = M->opts.levels ? M->opts.levels : 3;
        const uint32_t ndom = M->opts.ndom ? M->opts.ndom : 16;
        S->hid = build_hid(A, ndom, nlev);
        const Hid &H = S->hid;
        // the block factors choose their own schedule (AUTO: the subdomain
        // blocks of B run as independent one-workgroup chunks when they fit)
        const int trsv = M->opts.trsv;
        // factor drop tolerance: the reference's are 1e-3 (AMSED) and 1e-2
        // (GMSLR) (AMSEDPrecondition.h:39-51, GMSLRPrecondition.h:25-37); 0 gives
        // exact block factors
        const double tau = M->opts.tau;
        const uint32_t lfil = M->opts.lfil ? M->opts.lfil : 20;
        M->mnnz = 0;
        // last level
        uint32_t nc = 0;
        HostCsr Cl = extract_block(A, H, H.lev_ptr[nlev - 1], A.n, H.lev_ptr[nlev - 1], A.n, nc);
        S->Clast = make_csr(ctx, Cl, nc);
        S->last = make_factor(ctx, S->Clast, spd ? PP_PC_ICT : PP_PC_ILUT, tau, lfil, trsv);
        M->mnnz += S->last->mnnz;
        S->lev.resize(nlev - 1);
        for (uint32_t l = 0; l + 1 < nlev; ++l) {
            MlLevel &L = S->lev[l];
            L.off = H.lev_ptr[l];
            L.nb = H.lev_ptr[l + 1] - H.lev_ptr[l];
            L.nrest = A.n - H.lev_ptr[l + 1];
            uint32_t c = 0;
            HostCsr B = extract_block(A, H, H.lev_ptr[l], H.lev_ptr[l + 1], H.lev_ptr[l], H.lev_ptr[l + 1], c);
            HostCsr E = extract_block(A, H, H.lev_ptr[l], H.lev_ptr[l + 1], H.lev_ptr[l + 1], A.n, c);
            HostCsr F = spd ? transpose_rect(E, L.nrest)
                            : extract_block(A, H, H.lev_ptr[l + 1], A.n, H.lev_ptr[l], H.lev_ptr[l + 1], c);
            L.B = make_csr(ctx, B, L.nb);
            L.E = make_csr(ctx, E, L.nrest);
            L.F = make_csr(ctx, F, L.nb);
            L.Bpc = make_factor(ctx, L.B, spd ? PP_PC_ILDLT : PP_PC_ILUT, tau, lfil, trsv);
            L.t = dalloc<double>(std::max(1u, L.nb));
            M->mnnz += L.Bpc->mnnz;
        }
        S->dperm = dalloc<uint32_t>(A.n);
        PP_HIP(hipMemcpy(S->dperm, H.perm.data(), A.n * sizeof(uint32_t), hipMemcpyHostToDevice));
        S->xp = dalloc<double>(A.n);
        const uint32_t rank = M->opts.rank;
        const uint32_t steps =
            M->opts.lanczos_steps ? M->opts.lanczos_steps : std::max<uint32_t>(2 * rank, rank + 16);
        S->lanczos_steps = steps;
        const int form = M->opts.lowrank == PP_LOWRANK_AUTO ? PP_LOWRANK_REFERENCE : M->opts.lowrank;
        PP_CHECK(form == PP_LOWRANK_REFERENCE || form == PP_LOWRANK_DEFLATION || form == PP_LOWRANK_SUBSPACE,
                 PP_ERR_INVALID, "unknown low-rank form");
        PP_CHECK(spd || form != PP_LOWRANK_DEFLATION, PP_ERR_UNSUPPORTED,
                 "the eigenvalue-deflation low-rank form is AMSED's (symmetric) only");
        S->lowrank = form;
        for (int l = (int)nlev - 2; l >= 0; --l) {
            if (spd) {
                if (form == PP_LOWRANK_SUBSPACE) lowrank_lanczos(ctx, *S, l, rank, steps);
                else
                    lowrank_lanczos_restarted(ctx, *S, l, rank, form, M->opts.eps,
                                              M->opts.lanczos_steps ? M->opts.lanczos_steps : kAmsedMsteps);
            } else {
                lowrank_arnoldi(ctx, *S, l, rank, steps, form == PP_LOWRANK_REFERENCE);
            }
            MlLevel &L = S->lev[l];
            M->mnnz += 2 * L.rank * L.nrest + L.rank;  // AMSEDPrecondition.cpp:1269 convention
        }
        M->low64.nlev = nlev;
        M->up64.nlev = nlev;
    } catch (...) {
        ml_free(M);
        throw;
    }
}

// FP32 copies of every operand of the apply (before any graph capture:
// allocations and conversions happen here, not inside an apply)
void ml_need_f32(pp_ctx *ctx, pp_precond *M) {
    auto *S = static_cast<MlState *>(M->multilevel);
    PP_CHECK(S, PP_ERR_NOT_READY, "multilevel preconditioner not set up");
    if (S->have32) return;
    precond_need_f32(S->last);
    for (MlLevel &L : S->lev) {
        precond_need_f32(L.Bpc);
        csr_need_f32(ctx, L.E);
        csr_need_f32(ctx, L.F);
        L.t32 = dalloc<float>(std::max(1u, L.nb));
        if (L.rank) {
            L.save32 = dalloc<float>(L.nrest);
            L.W32 = dalloc<float>((size_t)L.nrest * L.rank);
            const size_t cnt = (size_t)L.nrest * L.rank;
            k_to_f32<<<(unsigned)std::min<size_t>((cnt + 255) / 256, 4096), 256, 0, ctx->stream>>>(cnt, L.W, L.W32);
            PP_HIP(hipGetLastError());
        }
    }
    S->xp32 = dalloc<float>(S->n);
    PP_HIP(hipStreamSynchronize(ctx->stream));
    S->have32 = true;
}

void ml_apply(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y) {
    auto *S = static_cast<MlState *>(M->multilevel);
    PP_CHECK(S, PP_ERR_NOT_READY, "multilevel preconditioner not set up");
    const unsigned g = std::max(1u, std::min<uint32_t>((S->n + 255) / 256, 4096));
    if (dtype == PP_F64) {
        k_gather<double><<<g, 256, 0, ctx->stream>>>(S->n, S->dperm, (const double *)x, S->xp);
        rec_solve<double>(ctx, *S, -1, S->xp);
        k_scatter<double><<<g, 256, 0, ctx->stream>>>(S->n, S->dperm, S->xp, (double *)y);
    } else {
        PP_CHECK(S->have32, PP_ERR_NOT_READY, "FP32 multilevel apply before its FP32 copies exist");
        k_gather<float><<<g, 256, 0, ctx->stream>>>(S->n, S->dperm, (const float *)x, S->xp32);
        rec_solve<float>(ctx, *S, -1, S->xp32);
        k_scatter<float><<<g, 256, 0, ctx->stream>>>(S->n, S->dperm, S->xp32, (float *)y);
    }
    PP_HIP(hipGetLastError());
}

const pp_precond *ml_block_factor(const pp_precond *M) {
    const auto *S = M ? (const MlState *)M->multilevel : nullptr;
    return S && !S->lev.empty() ? S->lev[0].Bpc : nullptr;
}

}  // namespace pp

extern "C" int pp_ml_info(const pp_precond *M, uint32_t *nlev, uint32_t *level_sizes, uint32_t *ranks,
                          uint32_t cap) {
    if (!M || !M->multilevel) return PP_ERR_NOT_READY;
    auto *S = static_cast<const pp::MlState *>(M->multilevel);
    if (nlev) *nlev = S->hid.nlev;
    for (uint32_t l = 0; l < S->hid.nlev && l < cap; ++l) {
        if (level_sizes) level_sizes[l] = S->hid.lev_ptr[l + 1] - S->hid.lev_ptr[l];
        if (ranks) ranks[l] = l + 1 < S->hid.nlev ? S->lev[l].rank : 0;
    }
    return PP_OK;
}

extern "C" int pp_ml_permutation(const pp_precond *M, uint32_t *perm) {
    if (!M || !M->multilevel || !perm) return PP_ERR_NOT_READY;
    auto *S = static_cast<const pp::MlState *>(M->multilevel);
    std::copy(S->hid.perm.begin(), S->hid.perm.end(), perm);
    return PP_OK;
}

// out (device, n x r col-major) = V (n x k) * Y (k x r) on the matrix cores; all device pointers.
extern "C" int pp_gemm_tall_f64(pp_ctx *ctx, uint32_t n, uint32_t k, uint32_t r, const double *V, const double *Y,
                                double *W) {
    if (!ctx || !V || !Y || !W) return PP_ERR_INVALID;
    try {
        pp::launch_gemm_tall_f64(ctx, n, k, r, V, Y, W);
        PP_HIP(hipStreamSynchronize(ctx->stream));
    } catch (const pp::Error &e) {
        ctx->err = e.what();
        return e.status;
    }
    return PP_OK;
}
