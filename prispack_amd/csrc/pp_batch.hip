// pp_batch.hip -- batched persistent IC(0)-preconditioned CG over many
// independent SPD systems of uneven size.  ONE launch solves the whole batch:
// workgroup s owns system s for the entire solve (SpMV, both triangular
// sweeps level by level with workgroup barriers, workgroup reductions, the
// FP64 refinement loop), so no grid-wide synchronisation is ever needed and
// ~256 systems fill the 256 CUs.  The launch is captured in a hipGraph and
// replayed.  The workgroup size (the row tile each SpMV / sweep step covers)
// is chosen per batch by a small timing autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>

#include "pp_internal.h"

namespace pp {

struct DevBatchSweep {
    uint32_t *perm = nullptr;     // global row processed at global position k
    uint32_t *rp = nullptr;       // [ntot+1] positions (level order within each system)
    uint32_t *ci = nullptr;       // global dependency rows
    double *val64 = nullptr;
    float *val32 = nullptr;
    double *diag64 = nullptr;     // by global row
    float *diag32 = nullptr;
    uint32_t *lev = nullptr;      // concatenated level pointers (global positions)
    uint32_t *lev_off = nullptr;  // [nsys+1] offsets into lev
    void free_all() {
        dfree(perm); dfree(rp); dfree(ci); dfree(val64); dfree(val32);
        dfree(diag64); dfree(diag32); dfree(lev); dfree(lev_off);
        perm = rp = ci = lev = lev_off = nullptr;
        val64 = diag64 = nullptr;
        val32 = diag32 = nullptr;
    }
};

// Everything the persistent kernel reads, by value.
struct BatchArgs {
    uint32_t nsys;
    const uint32_t *roff;  // [nsys+1]
    const uint32_t *rp, *ci;
    const double *v64;
    const float *v32;
    DevBatchSweep lo, up;
    const double *b;
    double *x;
    double *r64;
    float *r32, *e32, *q32, *w32, *d32, *p32;  // rhs, solution, residual, M^-1 r, direction, A d
    double *rr64, *w64, *d64, *p64;             // FP64 mode: residual, M^-1 r, direction, A d
    double tol, inner_tol;
    uint32_t maxit, max_outer;
    // reduced form of every system's two-colour IC(0) (mixed inner solves),
    // or red_n0 == nullptr: per system its n0 first-colour rows come first;
    // kab: KA rows (first colour) then KB rows (second colour), kfe: KE rows
    // then KF rows, all in batch row and column numbering; s1, dt over the
    // second colour's positions
    const uint32_t *red_n0;
    const uint32_t *kab_rp, *kab_ci, *kfe_rp, *kfe_ci;
    const float *kab_v, *kfe_v, *red_s1, *red_dt;
    int32_t *iters;       // [nsys] total inner iterations
    int32_t *outer;       // [nsys]
    int32_t *conv;        // [nsys]
    double *relres;       // [nsys] final ||b - A x|| / ||b||
};

namespace {

template <int BS>
__device__ inline double block_reduce(double v, double *sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) t += sh[w];
    return t;  // every thread holds the total (fixed order)
}

// Rows handled per thread and pass: their index, entry and vector loads are
// issued together (a system's workgroup is alone on its CU, so its
// latency is hidden by loads in flight, not by other waves).
constexpr int kRowsInFlight = 4;

// y = A x on the rows [r0, r1) with values V (FP64 or FP32); each row's sum
// in storage order
template <class T, class V>
__device__ inline void sys_spmv(const uint32_t *rp, const uint32_t *ci, const V *val, const T *x, T *y,
                                uint32_t r0, uint32_t r1) {
    constexpr int U = kRowsInFlight;
    const uint32_t step = blockDim.x;
    for (uint32_t i0 = r0 + threadIdx.x; i0 < r1; i0 += U * step) {
        uint32_t k[U], e[U];
        T sum[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * step;
            k[u] = i < r1 ? rp[i] : 0;
            e[u] = i < r1 ? rp[i + 1] : 0;
            sum[u] = 0;
        }
        for (bool more = true; more;) {
            more = false;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k[u] < e[u]) {
                    sum[u] += (T)val[k[u]] * x[ci[k[u]]];
                    ++k[u];
                    more = true;
                }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t i = i0 + u * step;
            if (i < r1) y[i] = sum[u];
        }
    }
}

// Level-by-level sweep of one system; y may alias b (rows of a level only read
// rows of earlier levels, finished before the barrier, and each row reads its
// own b before writing its y).  mode 0: divide by diag (IC(0) both sweeps).
template <class T>
__device__ inline void sys_sweep(const uint32_t *perm, const uint32_t *rp, const uint32_t *ci, const T *val,
                                 const T *diag, const uint32_t *lev, uint32_t l0, uint32_t l1, const T *b, T *y) {
    constexpr int U = kRowsInFlight;
    const uint32_t step = blockDim.x;
    for (uint32_t l = l0; l < l1; ++l) {
        const uint32_t k1 = lev[l + 1];
        for (uint32_t k0 = lev[l] + threadIdx.x; k0 < k1; k0 += U * step) {
            uint32_t row[U], e[U], f[U];
            T sum[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t k = k0 + u * step;
                const bool ok = k < k1;
                row[u] = ok ? perm[k] : 0;
                e[u] = ok ? rp[k] : 0;
                f[u] = ok ? rp[k + 1] : 0;
                sum[u] = ok ? b[row[u]] : (T)0;
            }
            for (bool more = true; more;) {
                more = false;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (e[u] < f[u]) {
                        sum[u] -= val[e[u]] * y[ci[e[u]]];
                        ++e[u];
                        more = true;
                    }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k0 + u * step < k1) y[row[u]] = sum[u] / diag[row[u]];
        }
        __syncthreads();
    }
}

template <class T>
struct Pick;
template <>
struct Pick<float> {
    __device__ static const float *v(const BatchArgs &a) { return a.v32; }
    __device__ static const float *lv(const DevBatchSweep &s) { return s.val32; }
    __device__ static const float *dg(const DevBatchSweep &s) { return s.diag32; }
};
template <>
struct Pick<double> {
    __device__ static const double *v(const BatchArgs &a) { return a.v64; }
    __device__ static const double *lv(const DevBatchSweep &s) { return s.val64; }
    __device__ static const double *dg(const DevBatchSweep &s) { return s.diag64; }
};

// FlexibleCG (source/Solver/CG.cpp:33-83) on one system in precision T,
// x0 = 0 in x, stopping on the recursive residual.  Returns iterations in the
// reference convention (m + 1).
template <class T, int BS>
__device__ int sys_pcg(const BatchArgs &a, uint32_t s, uint32_t r0, uint32_t r1, const T *b, T *x, T *r, T *w,
                       T *d, T *v, double tol, uint32_t maxit, double *sh, int *converged) {
    const T *A = Pick<T>::v(a);
    // system s owns level pointers lev[off[s] .. off[s+1]-1] (nlev + 1 of them)
    const uint32_t L0 = a.lo.lev_off[s], L1 = a.lo.lev_off[s + 1] - 1;
    const uint32_t U0 = a.up.lev_off[s], U1 = a.up.lev_off[s + 1] - 1;
    auto apply = [&](const T *in, T *out) {
        sys_sweep<T>(a.lo.perm, a.lo.rp, a.lo.ci, Pick<T>::lv(a.lo), Pick<T>::dg(a.lo), a.lo.lev, L0, L1, in, out);
        sys_sweep<T>(a.up.perm, a.up.rp, a.up.ci, Pick<T>::lv(a.up), Pick<T>::dg(a.up), a.up.lev, U0, U1, out, out);
    };
    // r = b (x0 = 0), w = M^{-1} r, d = w, v = A d
    double bb = 0.0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) {
        r[i] = b[i];
        x[i] = 0;
        bb += (double)b[i] * (double)b[i];
    }
    const double normb = sqrt(block_reduce<BS>(bb, sh));
    __syncthreads();
    apply(r, w);
    double dr = 0.0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) {
        d[i] = w[i];
        dr += (double)(w[i] * r[i]);
    }
    dr = block_reduce<BS>(dr, sh);
    __syncthreads();
    sys_spmv<T, T>(a.rp, a.ci, A, d, v, r0, r1);
    __syncthreads();
    double dv = 0.0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) dv += (double)(d[i] * v[i]);
    dv = block_reduce<BS>(dv, sh);
    uint32_t m = maxit;
    *converged = 0;
    if (normb == 0.0) {
        *converged = 1;
        return 1;
    }
    for (uint32_t it = 0;; ++it) {
        // x += alpha d ; r -= alpha v ; (r, r)
        const T alpha = (T)(dr / dv);
        double rr = 0.0;
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) {
            x[i] += alpha * d[i];
            const T rn = r[i] - alpha * v[i];
            r[i] = rn;
            rr += (double)(rn * rn);
        }
        rr = block_reduce<BS>(rr, sh);
        const uint32_t check = it + 1;
        if (check >= maxit) break;
        const double norm = sqrt(rr) / normb;
        if (!(norm == norm)) break;
        if (norm < tol) {
            *converged = 1;
            m = check;
            break;
        }
        __syncthreads();
        apply(r, w);
        double wv = 0.0;
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) wv += (double)(w[i] * v[i]);
        wv = block_reduce<BS>(wv, sh);
        const T beta = (T)(wv / dv);
        dr = 0.0;
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) {
            const T dn = w[i] - beta * d[i];
            d[i] = dn;
            dr += (double)(dn * r[i]);
        }
        dr = block_reduce<BS>(dr, sh);
        __syncthreads();
        sys_spmv<T, T>(a.rp, a.ci, A, d, v, r0, r1);
        __syncthreads();
        dv = 0.0;
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) dv += (double)(d[i] * v[i]);
        dv = block_reduce<BS>(dv, sh);
        if (dv == 0.0 || !(dv == dv)) break;
    }
    __syncthreads();
    return (int)m + 1;
}

// three sums at once (fixed order); sh holds 3 * BS / 64 doubles
template <int BS>
__device__ inline void block_reduce3(double &a, double &b, double &c, double *sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        c += __shfl_down(c, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        sh[wid] = a;
        sh[BS / 64 + wid] = b;
        sh[2 * BS / 64 + wid] = c;
    }
    __syncthreads();
    a = b = c = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) {
        a += sh[w];
        b += sh[BS / 64 + w];
        c += sh[2 * BS / 64 + w];
    }
}

// CG on the reduced system of system s's two-colour IC(0) (the same
// iteration as the single-system reduced CG: single-reduction recurrence,
// u = r): solves A e = rin approximately to the relative reduction tol of
// the inner system's true residual sqrt(sum Dt r^2) / ||rin||.  z holds
// y0 = KA r on the first colour and r on the second; x accumulates in e's
// second-colour part.  Returns the number of residual checks.
template <int BS>
__device__ int sys_red_cg(const BatchArgs &a, uint32_t s, uint32_t r0, uint32_t r1, const float *rin, float *e,
                          float *z, float *w, float *p, float *sv, double tol, uint32_t maxit, double *sh,
                          double *sh3, int *converged) {
    constexpr int U = kRowsInFlight;
    const uint32_t m0 = r0 + a.red_n0[s];
    // r = KF rin (second colour); x = p = s = 0; ||rin||
    sys_spmv<float, float>(a.kfe_rp, a.kfe_ci, a.kfe_v, rin, z, m0, r1);
    double bb = 0.0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) bb += (double)rin[i] * (double)rin[i];
    for (uint32_t i = m0 + threadIdx.x; i < r1; i += BS) {
        e[i] = 0.f;
        p[i] = 0.f;
        sv[i] = 0.f;
    }
    const double normb = sqrt(block_reduce<BS>(bb, sh));
    __syncthreads();
    *converged = 0;
    uint32_t m = maxit;
    double gold = 0.0, aold = 1.0;
    bool first = true;
    for (uint32_t c = 1;; ++c) {
        sys_spmv<float, float>(a.kab_rp, a.kab_ci, a.kab_v, z, z, r0, m0);  // y0 = KA r
        __syncthreads();
        // w = KB [y0; r] with (r, w), (r, r) and sum Dt r^2
        double delta = 0.0, gamma = 0.0, rho = 0.0;
        for (uint32_t i0 = m0 + threadIdx.x; i0 < r1; i0 += U * BS) {
            uint32_t k[U], f[U];
            float sum[U], ri[U], di[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * BS;
                const bool ok = i < r1;
                k[u] = ok ? a.kab_rp[i] : 0;
                f[u] = ok ? a.kab_rp[i + 1] : 0;
                ri[u] = ok ? z[i] : 0.f;
                di[u] = ok ? a.red_dt[i] : 0.f;
                sum[u] = 0.f;
            }
            for (bool more = true; more;) {
                more = false;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (k[u] < f[u]) {
                        sum[u] += a.kab_v[k[u]] * z[a.kab_ci[k[u]]];
                        ++k[u];
                        more = true;
                    }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * BS;
                if (i < r1) {
                    w[i] = sum[u];
                    delta += (double)(ri[u] * sum[u]);
                    gamma += (double)(ri[u] * ri[u]);
                    rho += (double)di[u] * (double)(ri[u] * ri[u]);
                }
            }
        }
        block_reduce3<BS>(delta, gamma, rho, sh3);
        if (c >= maxit) break;
        const double norm = normb > 0.0 ? sqrt(rho) / normb : 0.0;
        if (!(norm == norm) || isinf(norm)) break;
        if (norm < tol) {
            *converged = 1;
            m = c;
            break;
        }
        const double beta = first ? 0.0 : gamma / gold;
        const double den = first ? delta : delta - beta * gamma / aold;
        if (den == 0.0 || !(den == den)) break;
        const double alpha = gamma / den;
        gold = gamma;
        aold = alpha;
        first = false;
        const float fa = (float)alpha, fb = (float)beta;
        for (uint32_t i0 = m0 + threadIdx.x; i0 < r1; i0 += U * BS) {
            float zr[U], wr[U], pr[U], sr[U], er[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * BS;
                if (i < r1) {
                    zr[u] = z[i];
                    wr[u] = w[i];
                    pr[u] = p[i];
                    sr[u] = sv[i];
                    er[u] = e[i];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * BS;
                if (i < r1) {
                    const float pn = zr[u] + fb * pr[u];
                    const float sn = wr[u] + fb * sr[u];
                    p[i] = pn;
                    sv[i] = sn;
                    e[i] = er[u] + fa * pn;
                    z[i] = zr[u] - fa * sn;
                }
            }
        }
        __syncthreads();
    }
    // e = [KE [rin_0; e_1]; e_1], e_1 = s1 x
    __syncthreads();
    for (uint32_t i = m0 + threadIdx.x; i < r1; i += BS) {
        const float ei = a.red_s1[i] * e[i];
        e[i] = ei;
        z[i] = ei;
    }
    for (uint32_t i = r0 + threadIdx.x; i < m0; i += BS) z[i] = rin[i];
    __syncthreads();
    sys_spmv<float, float>(a.kfe_rp, a.kfe_ci, a.kfe_v, z, e, r0, m0);
    __syncthreads();
    return (int)m;
}

// One workgroup per system.  MIXED: FP64 residual r = b - A x, FP32 inner
// FlexibleCG on A e = r / ||r|| (or the reduced CG where every system has
// the two-colour structure), x += ||r|| e, until ||b - A x|| / ||b|| < tol.
template <int BS, bool MIXED>
__global__ __launch_bounds__(BS) void k_batch_pcg(BatchArgs a) {
    __shared__ double sh[BS / 64];
    __shared__ double sh3[3 * (BS / 64)];
    const uint32_t s = blockIdx.x;
    if (s >= a.nsys) return;
    const uint32_t r0 = a.roff[s], r1 = a.roff[s + 1];
    int conv = 0, total = 0, outer = 0;
    double rel = 1.0;
    if (!MIXED) {
        total = sys_pcg<double, BS>(a, s, r0, r1, a.b, a.x, a.rr64, a.w64, a.d64, a.p64, a.tol, a.maxit, sh, &conv);
        outer = 1;
        __syncthreads();
        sys_spmv<double, double>(a.rp, a.ci, a.v64, a.x, a.r64, r0, r1);
        __syncthreads();
        double rr = 0.0, bb = 0.0;
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) {
            const double t = a.b[i] - a.r64[i];
            rr += t * t;
            bb += a.b[i] * a.b[i];
        }
        rr = block_reduce<BS>(rr, sh);
        bb = block_reduce<BS>(bb, sh);
        rel = bb > 0 ? sqrt(rr) / sqrt(bb) : 0.0;
    } else {
        double bb = 0.0;
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) {
            bb += a.b[i] * a.b[i];
            a.x[i] = 0.0;
        }
        const double normb = sqrt(block_reduce<BS>(bb, sh));
        for (;;) {
            __syncthreads();
            sys_spmv<double, double>(a.rp, a.ci, a.v64, a.x, a.r64, r0, r1);
            __syncthreads();
            double rr = 0.0;
            for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) {
                const double t = a.b[i] - a.r64[i];
                a.r64[i] = t;
                rr += t * t;
            }
            const double rn = sqrt(block_reduce<BS>(rr, sh));
            rel = normb > 0 ? rn / normb : 0.0;
            if (rel < a.tol || !(rel == rel)) {
                conv = rel < a.tol;
                break;
            }
            if ((uint32_t)outer >= a.max_outer || (uint32_t)total >= a.maxit) break;
            const float sc = (float)(1.0 / rn);
            for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) a.r32[i] = (float)a.r64[i] * sc;
            __syncthreads();
            const double want = fmax(a.inner_tol, 0.5 * a.tol * normb / rn);
            int ic = 0;
            if (a.red_n0)
                total += sys_red_cg<BS>(a, s, r0, r1, a.r32, a.e32, a.q32, a.w32, a.d32, a.p32, want,
                                        a.maxit - (uint32_t)total, sh, sh3, &ic);
            else
                total += sys_pcg<float, BS>(a, s, r0, r1, a.r32, a.e32, a.q32, a.w32, a.d32, a.p32, want,
                                            a.maxit - (uint32_t)total, sh, &ic);
            ++outer;
            for (uint32_t i = r0 + threadIdx.x; i < r1; i += BS) a.x[i] += rn * (double)a.e32[i];
        }
    }
    if (threadIdx.x == 0) {
        a.iters[s] = total;
        a.outer[s] = outer;
        a.conv[s] = conv;
        a.relres[s] = rel;
    }
}

}  // namespace
}  // namespace pp

// ============================================================================ host side
struct pp_batch {
    pp_ctx *ctx = nullptr;
    uint32_t nsys = 0, ntot = 0, nnz = 0;
    std::vector<uint32_t> h_roff, h_rp, h_ci;
    std::vector<double> h_v;
    uint32_t *roff = nullptr, *rp = nullptr, *ci = nullptr;
    double *v64 = nullptr;
    float *v32 = nullptr;
    pp::DevBatchSweep lo, up;
    double *r64 = nullptr, *rr64 = nullptr, *w64 = nullptr, *d64 = nullptr, *p64 = nullptr;
    float *r32 = nullptr, *e32 = nullptr, *q32 = nullptr, *w32 = nullptr, *d32 = nullptr, *p32 = nullptr;
    int32_t *iters = nullptr, *outer = nullptr, *conv = nullptr;
    double *relres = nullptr;
    int block = 256;
    bool symmetric = true;
    uint32_t max_levels = 0;
    // PP_ORDER_MULTICOLOR: every system is solved as P_s A_s P_s^T (rows
    // grouped by colour, so its IC(0) sweeps have one level per colour);
    // the kernel works on the permuted copy below, b and x are permuted once
    // per solve.  Position k of the permuted batch holds original row gperm[k];
    // permuted entry e holds original entry vmap[e].
    int ordering = PP_ORDER_NATURAL;
    std::vector<uint32_t> h_gperm, h_prp, h_pci, vmap;
    uint32_t *gperm = nullptr, *prp = nullptr, *pci = nullptr;
    double *pv64 = nullptr, *bp = nullptr, *xp = nullptr;
    float *pv32 = nullptr;
    // the reduced form of every system's two-colour IC(0) (multicolour order
    // only, and only when every system has the structure; see BatchArgs)
    uint32_t *red_n0 = nullptr, *kab_rp = nullptr, *kab_ci = nullptr, *kfe_rp = nullptr, *kfe_ci = nullptr;
    float *kab_v = nullptr, *kfe_v = nullptr, *red_s1 = nullptr, *red_dt = nullptr;
};

namespace pp {
namespace {

struct HostBatchSweep {
    std::vector<uint32_t> perm, rp{0}, ci, lev, lev_off{0};
    std::vector<double> val, diag;
};

// Append the gather-form sweep of one system (local CSR `R`, rows offset by
// goff): rows grouped by dependency level, dependencies in storage order.
void append_sweep(HostBatchSweep &H, const HostCsr &R, bool lower, bool diag_first_in_row, uint32_t goff,
                  uint32_t &nlev_out) {
    const uint32_t n = R.n;
    std::vector<uint32_t> lev(n, 0), rp(n + 1, 0), ci;
    std::vector<double> val, dg(n, 1.0);
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t k = R.rp[i]; k < R.rp[i + 1]; ++k) {
            if (R.ci[k] == i) {
                dg[i] = R.v[k];
                continue;
            }
            ci.push_back(R.ci[k]);
            val.push_back(R.v[k]);
        }
        rp[i + 1] = (uint32_t)ci.size();
    }
    (void)diag_first_in_row;
    uint32_t maxl = 0;
    auto visit = [&](uint32_t i) {
        uint32_t l = 0;
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) l = std::max(l, lev[ci[k]] + 1);
        lev[i] = l;
        maxl = std::max(maxl, l);
    };
    if (lower)
        for (uint32_t i = 0; i < n; ++i) visit(i);
    else
        for (uint32_t i = n; i-- > 0;) visit(i);
    const uint32_t nlev = n ? maxl + 1 : 0;
    std::vector<uint32_t> lp(nlev + 1, 0);
    for (uint32_t i = 0; i < n; ++i) lp[lev[i] + 1]++;
    for (uint32_t l = 0; l < nlev; ++l) lp[l + 1] += lp[l];
    std::vector<uint32_t> fill(lp.begin(), lp.end() - (nlev ? 1 : 0)), perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[fill[lev[i]]++] = i;
    const uint32_t base = (uint32_t)H.perm.size();
    for (uint32_t q = 0; q < n; ++q) {
        const uint32_t i = perm[q];
        H.perm.push_back(goff + i);
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
            H.ci.push_back(goff + ci[k]);
            H.val.push_back(val[k]);
        }
        H.rp.push_back((uint32_t)H.ci.size());
    }
    for (uint32_t l = 0; l <= nlev; ++l) H.lev.push_back(base + lp[l]);
    H.lev_off.push_back((uint32_t)H.lev.size());
    H.diag.insert(H.diag.end(), dg.begin(), dg.end());
    nlev_out = nlev;
}

// Concatenate a part built by append_sweep into an empty HostBatchSweep
// onto H (positions, entry offsets and level pointers shifted).
void append_parts(HostBatchSweep &H, const HostBatchSweep &P) {
    const uint32_t bp = (uint32_t)H.perm.size(), be = (uint32_t)H.ci.size(), bl = (uint32_t)H.lev.size();
    H.perm.insert(H.perm.end(), P.perm.begin(), P.perm.end());
    H.ci.insert(H.ci.end(), P.ci.begin(), P.ci.end());
    H.val.insert(H.val.end(), P.val.begin(), P.val.end());
    H.diag.insert(H.diag.end(), P.diag.begin(), P.diag.end());
    for (size_t q = 1; q < P.rp.size(); ++q) H.rp.push_back(be + P.rp[q]);
    for (uint32_t l : P.lev) H.lev.push_back(bp + l);
    for (size_t q = 1; q < P.lev_off.size(); ++q) H.lev_off.push_back(bl + P.lev_off[q]);
}

template <class T>
T *upload(const std::vector<T> &h) {
    T *d = dalloc<T>(h.size());
    if (!h.empty()) PP_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

std::vector<float> to_f32(const std::vector<double> &v) { return std::vector<float>(v.begin(), v.end()); }

void upload_batch_sweep(const HostBatchSweep &H, DevBatchSweep &D) {
    D.free_all();
    D.perm = upload(H.perm);
    D.rp = upload(H.rp);
    D.ci = upload(H.ci);
    D.val64 = upload(H.val);
    D.val32 = upload(to_f32(H.val));
    D.diag64 = upload(H.diag);
    D.diag32 = upload(to_f32(H.diag));
    D.lev = upload(H.lev);
    D.lev_off = upload(H.lev_off);
}

// The permuted pattern of every system (multicolour order within each
// diagonal block) and the entry map; values follow through vmap.
void batch_build_order(pp_batch *B) {
    B->h_gperm.assign(B->ntot, 0);
    B->h_prp.assign(B->ntot + 1, 0);
    B->h_pci.assign(B->nnz, 0);
    B->vmap.assign(B->nnz, 0);
    // a system's permuted entries occupy the same range as its original ones
    parallel_for(B->nsys, 1, [&](size_t s0, size_t s1) {
        std::vector<std::pair<uint32_t, uint32_t>> row;  // (new column, original entry)
        for (size_t s = s0; s < s1; ++s) {
            const uint32_t r0 = B->h_roff[s], r1 = B->h_roff[s + 1], n = r1 - r0;
            HostCsr A;
            A.n = n;
            A.rp.assign(n + 1, 0);
            for (uint32_t i = 0; i < n; ++i) A.rp[i + 1] = B->h_rp[r0 + i + 1] - B->h_rp[r0];
            A.ci.resize(A.rp[n]);
            A.v.assign(A.rp[n], 0.0);
            for (uint32_t k = 0; k < A.rp[n]; ++k) A.ci[k] = B->h_ci[B->h_rp[r0] + k] - r0;
            uint32_t ncol = 0;
            const std::vector<uint32_t> order = multicolor_order(A, ncol);
            std::vector<uint32_t> inv(n);
            for (uint32_t k = 0; k < n; ++k) inv[order[k]] = k;
            uint32_t q = B->h_rp[r0];
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t i = order[k];
                B->h_gperm[r0 + k] = r0 + i;
                row.clear();
                for (uint32_t e = B->h_rp[r0 + i]; e < B->h_rp[r0 + i + 1]; ++e)
                    row.emplace_back(inv[B->h_ci[e] - r0], e);
                std::sort(row.begin(), row.end());
                for (const auto &pe : row) {
                    B->h_pci[q] = r0 + pe.first;
                    B->vmap[q++] = pe.second;
                }
                B->h_prp[r0 + k + 1] = q;
            }
        }
    });
}

// values of the permuted copy from the original values
void batch_upload_permuted_values(pp_batch *B) {
    std::vector<double> pv(B->nnz);
    parallel_for(B->nnz, 1u << 16, [&](size_t e0, size_t e1) {
        for (size_t e = e0; e < e1; ++e) pv[e] = B->h_v[B->vmap[e]];
    });
    if (!B->pv64) B->pv64 = dalloc<double>(B->nnz);
    if (!B->pv32) B->pv32 = dalloc<float>(B->nnz);
    PP_HIP(hipMemcpy(B->pv64, pv.data(), B->nnz * sizeof(double), hipMemcpyHostToDevice));
    const std::vector<float> p32 = to_f32(pv);
    PP_HIP(hipMemcpy(B->pv32, p32.data(), B->nnz * sizeof(float), hipMemcpyHostToDevice));
}

void batch_free_reduced(pp_batch *B) {
    dfree(B->red_n0); dfree(B->kab_rp); dfree(B->kab_ci); dfree(B->kfe_rp); dfree(B->kfe_ci);
    dfree(B->kab_v); dfree(B->kfe_v); dfree(B->red_s1); dfree(B->red_dt);
    B->red_n0 = B->kab_rp = B->kab_ci = B->kfe_rp = B->kfe_ci = nullptr;
    B->kab_v = B->kfe_v = B->red_s1 = B->red_dt = nullptr;
}

// The systems' reduced operators in batch numbering (FP32 values: they run
// only in the mixed solves' inner iterations).
void batch_upload_reduced(pp_batch *B, const std::vector<ReducedHost> &reds) {
    const uint32_t ns = B->nsys, nt = B->ntot;
    std::vector<uint32_t> n0(ns), ab_rp(nt + 1, 0), fe_rp(nt + 1, 0);
    for (uint32_t s = 0; s < ns; ++s) {
        const ReducedHost &H = reds[s];
        const uint32_t r0 = B->h_roff[s];
        n0[s] = H.n0;
        for (uint32_t i = 0; i < H.n0; ++i) {
            ab_rp[r0 + i + 1] = ab_rp[r0 + i] + (H.ka_rp[i + 1] - H.ka_rp[i]);
            fe_rp[r0 + i + 1] = fe_rp[r0 + i] + (H.ke_rp[i + 1] - H.ke_rp[i]);
        }
        for (uint32_t q = 0; q < H.n1; ++q) {
            const uint32_t i = r0 + H.n0 + q;
            ab_rp[i + 1] = ab_rp[i] + (H.kb_rp[q + 1] - H.kb_rp[q]);
            fe_rp[i + 1] = fe_rp[i] + (H.kf_rp[q + 1] - H.kf_rp[q]);
        }
    }
    std::vector<uint32_t> ab_ci(ab_rp[nt]), fe_ci(fe_rp[nt]);
    std::vector<float> ab_v(ab_rp[nt]), fe_v(fe_rp[nt]), s1(nt, 0.f), dt(nt, 0.f);
    parallel_for(ns, 1, [&](size_t s0, size_t s1e) {
        for (size_t s = s0; s < s1e; ++s) {
            const ReducedHost &H = reds[s];
            const uint32_t r0 = B->h_roff[s];
            auto put = [&](const std::vector<uint32_t> &rp, const std::vector<uint32_t> &ci,
                           const std::vector<double> &v, uint32_t rows, uint32_t row0, const std::vector<uint32_t> &grp,
                           std::vector<uint32_t> &gci, std::vector<float> &gv) {
                for (uint32_t i = 0; i < rows; ++i)
                    for (uint32_t k = rp[i], q = grp[row0 + i]; k < rp[i + 1]; ++k, ++q) {
                        gci[q] = r0 + ci[k];
                        gv[q] = (float)v[k];
                    }
            };
            put(H.ka_rp, H.ka_ci, H.ka_v, H.n0, r0, ab_rp, ab_ci, ab_v);
            put(H.kb_rp, H.kb_ci, H.kb_v, H.n1, r0 + H.n0, ab_rp, ab_ci, ab_v);
            put(H.ke_rp, H.ke_ci, H.ke_v, H.n0, r0, fe_rp, fe_ci, fe_v);
            put(H.kf_rp, H.kf_ci, H.kf_v, H.n1, r0 + H.n0, fe_rp, fe_ci, fe_v);
            for (uint32_t q = 0; q < H.n1; ++q) {
                s1[r0 + H.n0 + q] = (float)H.s1[q];
                dt[r0 + H.n0 + q] = (float)H.dt[q];
            }
        }
    });
    B->red_n0 = upload(n0);
    B->kab_rp = upload(ab_rp);
    B->kab_ci = upload(ab_ci);
    B->kab_v = upload(ab_v);
    B->kfe_rp = upload(fe_rp);
    B->kfe_ci = upload(fe_ci);
    B->kfe_v = upload(fe_v);
    B->red_s1 = upload(s1);
    B->red_dt = upload(dt);
}

// IC(0) of every diagonal block (of the permuted copy under multicolour
// ordering), as two gather-form sweeps.
void batch_factor(pp_batch *B) {
    const bool perm = B->ordering == PP_ORDER_MULTICOLOR;
    const std::vector<uint32_t> &Rp = perm ? B->h_prp : B->h_rp;
    const std::vector<uint32_t> &Ci = perm ? B->h_pci : B->h_ci;
    std::vector<HostBatchSweep> los(B->nsys), ups(B->nsys);
    std::vector<uint32_t> nlev(B->nsys, 0), fail_row(B->nsys, 0);
    std::vector<int> status(B->nsys, PP_OK);
    std::vector<ReducedHost> reds(perm ? B->nsys : 0);
    std::atomic<bool> red_ok{perm};
    // systems factor independently; their sweeps are concatenated in order
    parallel_for(B->nsys, 1, [&](size_t s0, size_t s1) {
        for (size_t s = s0; s < s1; ++s) {
            const uint32_t r0 = B->h_roff[s], r1 = B->h_roff[s + 1], n = r1 - r0;
            HostCsr A;
            A.n = n;
            A.rp.assign(n + 1, 0);
            for (uint32_t i = 0; i < n; ++i) A.rp[i + 1] = Rp[r0 + i + 1] - Rp[r0];
            A.ci.resize(A.rp[n]);
            A.v.resize(A.rp[n]);
            for (uint32_t k = 0; k < A.rp[n]; ++k) {
                const uint32_t e = Rp[r0] + k;
                A.ci[k] = Ci[e] - r0;
                A.v[k] = B->h_v[perm ? B->vmap[e] : e];
            }
            HostCsr lt;
            status[s] = factor_ic0(A, lt, fail_row[s]);
            if (status[s] != PP_OK) continue;
            if (red_ok && !reduced_host(n, A.rp, A.ci, A.v, reds[s])) red_ok = false;
            const HostCsr L = transpose(lt);
            uint32_t nl = 0, nu = 0;
            append_sweep(los[s], L, true, false, r0, nl);
            append_sweep(ups[s], lt, false, true, r0, nu);
            nlev[s] = std::max(nl, nu);
        }
    });
    for (uint32_t s = 0; s < B->nsys; ++s)  // the first failing system, as a serial loop reports it
        if (status[s] != PP_OK)
            throw Error(status[s], "batch system " + std::to_string(s) + ": IC(0) pivot not positive at row " +
                                       std::to_string(fail_row[s]));
    HostBatchSweep lo, up;
    B->max_levels = 0;
    for (uint32_t s = 0; s < B->nsys; ++s) {
        append_parts(lo, los[s]);
        append_parts(up, ups[s]);
        B->max_levels = std::max(B->max_levels, nlev[s]);
    }
    upload_batch_sweep(lo, B->lo);
    upload_batch_sweep(up, B->up);
    batch_free_reduced(B);
    if (red_ok) batch_upload_reduced(B, reds);
}

BatchArgs make_args(pp_batch *B, const pp_solve_opts &o, const double *b, double *x) {
    BatchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nsys = B->nsys;
    a.roff = B->roff;
    const bool perm = B->ordering == PP_ORDER_MULTICOLOR;
    a.rp = perm ? B->prp : B->rp;
    a.ci = perm ? B->pci : B->ci;
    a.v64 = perm ? B->pv64 : B->v64;
    a.v32 = perm ? B->pv32 : B->v32;
    a.lo = B->lo;
    a.up = B->up;
    a.b = b;
    a.x = x;
    a.r64 = B->r64;
    a.r32 = B->r32; a.e32 = B->e32; a.q32 = B->q32; a.w32 = B->w32; a.d32 = B->d32; a.p32 = B->p32;
    a.rr64 = B->rr64; a.w64 = B->w64; a.d64 = B->d64; a.p64 = B->p64;
    a.tol = o.tol;
    a.inner_tol = o.inner_tol > 0 ? o.inner_tol : 1e-6;
    a.maxit = o.maxit;
    a.max_outer = o.max_outer ? o.max_outer : 20;
    a.red_n0 = perm ? B->red_n0 : nullptr;
    a.kab_rp = B->kab_rp;
    a.kab_ci = B->kab_ci;
    a.kab_v = B->kab_v;
    a.kfe_rp = B->kfe_rp;
    a.kfe_ci = B->kfe_ci;
    a.kfe_v = B->kfe_v;
    a.red_s1 = B->red_s1;
    a.red_dt = B->red_dt;
    a.iters = B->iters;
    a.outer = B->outer;
    a.conv = B->conv;
    a.relres = B->relres;
    return a;
}

void launch_batch(hipStream_t st, int block, bool mixed, const BatchArgs &a) {
    const unsigned g = a.nsys;
#define PP_BATCH_CASE(BS)                                                          \
    if (block == BS) {                                                             \
        if (mixed) k_batch_pcg<BS, true><<<g, BS, 0, st>>>(a);                      \
        else k_batch_pcg<BS, false><<<g, BS, 0, st>>>(a);                           \
        PP_HIP(hipGetLastError());                                                 \
        return;                                                                    \
    }
    PP_BATCH_CASE(256)
    PP_BATCH_CASE(512)
    PP_BATCH_CASE(1024)
#undef PP_BATCH_CASE
    throw Error(PP_ERR_INVALID, "batch workgroup size must be 256, 512 or 1024");
}

// The single persistent launch, captured into a graph and replayed.
void run_batch(pp_ctx *ctx, pp_batch *B, bool mixed, const BatchArgs &a) {
    Graph G;
    PP_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    try {
        launch_batch(ctx->stream, B->block, mixed, a);
    } catch (...) {
        hipGraph_t tmp;
        (void)hipStreamEndCapture(ctx->stream, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
    }
    PP_HIP(hipStreamEndCapture(ctx->stream, &G.g));
    PP_HIP(hipGraphInstantiate(&G.exec, G.g, nullptr, nullptr, 0));
    PP_HIP(hipGraphLaunch(G.exec, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
}

void batch_free(pp_batch *B) {
    dfree(B->roff); dfree(B->rp); dfree(B->ci); dfree(B->v64); dfree(B->v32);
    B->lo.free_all();
    B->up.free_all();
    dfree(B->r64); dfree(B->rr64); dfree(B->w64); dfree(B->d64); dfree(B->p64);
    dfree(B->r32); dfree(B->e32); dfree(B->q32); dfree(B->w32); dfree(B->d32); dfree(B->p32);
    dfree(B->iters); dfree(B->outer); dfree(B->conv); dfree(B->relres);
    dfree(B->gperm); dfree(B->prp); dfree(B->pci); dfree(B->pv64); dfree(B->pv32); dfree(B->bp); dfree(B->xp);
    batch_free_reduced(B);
}

}  // namespace
}  // namespace pp

using namespace pp;

template <class F>
static int guarded_batch(pp_ctx *ctx, F &&f) {
    try {
        PP_HIP(hipSetDevice(ctx->device));
        f();
        return PP_OK;
    } catch (const pp::Error &e) {
        ctx->err = e.what();
        return e.status;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return PP_ERR_INVALID;
    }
}

extern "C" {

int pp_batch_create(pp_ctx *ctx, uint32_t nsys, const uint32_t *row_off, const uint32_t *rowptr,
                    const uint32_t *colidx, const double *values, pp_batch **out) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_batch(ctx, [&] {
        PP_CHECK(nsys > 0 && row_off && rowptr && colidx && values && out, PP_ERR_INVALID, "NULL argument");
        PP_CHECK(row_off[0] == 0, PP_ERR_INVALID, "row_off[0] must be 0");
        std::unique_ptr<pp_batch> B(new pp_batch);
        B->ctx = ctx;
        B->nsys = nsys;
        B->ntot = row_off[nsys];
        B->nnz = rowptr[B->ntot];
        B->h_roff.assign(row_off, row_off + nsys + 1);
        B->h_rp.assign(rowptr, rowptr + B->ntot + 1);
        B->h_ci.assign(colidx, colidx + B->nnz);
        B->h_v.assign(values, values + B->nnz);
        // block-diagonal and symmetric (IC(0)-CG needs SPD blocks)
        for (uint32_t s = 0; s < nsys; ++s)
            PP_CHECK(row_off[s] <= row_off[s + 1], PP_ERR_INVALID, "row_off not monotone");
        parallel_for(nsys, 1, [&](size_t s0, size_t s1) {
            for (size_t s = s0; s < s1; ++s)
                for (uint32_t i = row_off[s]; i < row_off[s + 1]; ++i)
                    for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
                        PP_CHECK(colidx[k] >= row_off[s] && colidx[k] < row_off[s + 1], PP_ERR_INVALID,
                                 "entry outside its system's diagonal block");
        });
        B->roff = upload(B->h_roff);
        B->rp = upload(B->h_rp);
        B->ci = upload(B->h_ci);
        B->v64 = upload(B->h_v);
        B->v32 = upload(to_f32(B->h_v));
        batch_factor(B.get());
        const size_t n = B->ntot;
        B->r64 = dalloc<double>(n); B->rr64 = dalloc<double>(n); B->w64 = dalloc<double>(n);
        B->d64 = dalloc<double>(n); B->p64 = dalloc<double>(n);
        B->r32 = dalloc<float>(n); B->e32 = dalloc<float>(n); B->q32 = dalloc<float>(n);
        B->w32 = dalloc<float>(n); B->d32 = dalloc<float>(n); B->p32 = dalloc<float>(n);
        B->iters = dalloc<int>(nsys); B->outer = dalloc<int>(nsys); B->conv = dalloc<int>(nsys);
        B->relres = dalloc<double>(nsys);
        *out = B.release();
    });
}

int pp_batch_set_values(pp_ctx *ctx, pp_batch *B, const double *values, int on_device) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_batch(ctx, [&] {
        PP_CHECK(B && values, PP_ERR_INVALID, "NULL argument");
        if (on_device)
            PP_HIP(hipMemcpy(B->h_v.data(), values, B->nnz * sizeof(double), hipMemcpyDeviceToHost));
        else
            std::memcpy(B->h_v.data(), values, B->nnz * sizeof(double));
        PP_HIP(hipMemcpy(B->v64, B->h_v.data(), B->nnz * sizeof(double), hipMemcpyHostToDevice));
        std::vector<float> v32 = to_f32(B->h_v);
        PP_HIP(hipMemcpy(B->v32, v32.data(), B->nnz * sizeof(float), hipMemcpyHostToDevice));
        if (B->ordering == PP_ORDER_MULTICOLOR) batch_upload_permuted_values(B);
        batch_factor(B);
    });
}

int pp_batch_set_ordering(pp_ctx *ctx, pp_batch *B, int ordering) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_batch(ctx, [&] {
        PP_CHECK(B, PP_ERR_INVALID, "NULL batch");
        PP_CHECK(ordering == PP_ORDER_NATURAL || ordering == PP_ORDER_MULTICOLOR, PP_ERR_INVALID, "unknown ordering");
        PP_HIP(hipStreamSynchronize(ctx->stream));
        B->ordering = ordering;
        if (ordering == PP_ORDER_MULTICOLOR && !B->gperm) {
            batch_build_order(B);
            B->gperm = upload(B->h_gperm);
            B->prp = upload(B->h_prp);
            B->pci = upload(B->h_pci);
            B->bp = dalloc<double>(B->ntot);
            B->xp = dalloc<double>(B->ntot);
            batch_upload_permuted_values(B);
        }
        batch_factor(B);
    });
}

static int batch_solve_impl(pp_ctx *ctx, pp_batch *B, const pp_solve_opts *opts, const double *b, double *x,
                            int32_t *iters_out, pp_solve_report *rep) {
    return guarded_batch(ctx, [&] {
        PP_CHECK(B && opts && b && x, PP_ERR_INVALID, "NULL argument");
        PP_CHECK(opts->tol > 0 && opts->maxit >= 1, PP_ERR_INVALID, "bad tolerance / maxit");
        const bool mixed = opts->precision == PP_MIXED;
        const bool perm = B->ordering == PP_ORDER_MULTICOLOR;
        BatchArgs a = make_args(B, *opts, perm ? B->bp : b, perm ? B->xp : x);
        PP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        if (perm) launch_permute(ctx, B->ntot, B->gperm, PP_F64, b, B->bp, true);
        run_batch(ctx, B, mixed, a);
        if (perm) {
            launch_permute(ctx, B->ntot, B->gperm, PP_F64, B->xp, x, false);
            PP_HIP(hipStreamSynchronize(ctx->stream));
        }
        PP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        PP_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0;
        PP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        std::vector<int32_t> it(B->nsys), ou(B->nsys), cv(B->nsys);
        std::vector<double> rr(B->nsys);
        PP_HIP(hipMemcpy(it.data(), B->iters, B->nsys * sizeof(int32_t), hipMemcpyDeviceToHost));
        PP_HIP(hipMemcpy(ou.data(), B->outer, B->nsys * sizeof(int32_t), hipMemcpyDeviceToHost));
        PP_HIP(hipMemcpy(cv.data(), B->conv, B->nsys * sizeof(int32_t), hipMemcpyDeviceToHost));
        PP_HIP(hipMemcpy(rr.data(), B->relres, B->nsys * sizeof(double), hipMemcpyDeviceToHost));
        if (iters_out) std::copy(it.begin(), it.end(), iters_out);
        if (rep) {
            std::memset(rep, 0, sizeof(*rep));
            rep->iters = *std::max_element(it.begin(), it.end());
            rep->inner_iters = 0;
            for (int32_t v : it) rep->inner_iters += v;
            rep->outer_iters = *std::max_element(ou.begin(), ou.end());
            rep->converged = std::all_of(cv.begin(), cv.end(), [](int32_t c) { return c != 0; });
            rep->true_relres = *std::max_element(rr.begin(), rr.end());
            rep->final_error = rep->true_relres;
            rep->solve_ms = ms;
            rep->status = PP_OK;
            rep->cg_variant = mixed && a.red_n0 ? PP_CG_REDUCED : PP_CG_REFERENCE;
        }
    });
}

int pp_batch_solve(pp_ctx *ctx, pp_batch *B, const pp_solve_opts *opts, const double *b, double *x,
                   int32_t *iters_out, pp_solve_report *rep) {
    if (!ctx) return PP_ERR_INVALID;
    return batch_solve_impl(ctx, B, opts, b, x, iters_out, rep);
}

// IC(0)-CG needs SPD blocks, so A^T = A and the adjoint system is the same one.
int pp_batch_solve_t(pp_ctx *ctx, pp_batch *B, const pp_solve_opts *opts, const double *g, double *lambda,
                     int32_t *iters_out, pp_solve_report *rep) {
    if (!ctx) return PP_ERR_INVALID;
    return batch_solve_impl(ctx, B, opts, g, lambda, iters_out, rep);
}

// Times a bounded number of mixed-precision iterations for each workgroup
// size and keeps the fastest.
int pp_batch_tune(pp_ctx *ctx, pp_batch *B, int *chosen) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_batch(ctx, [&] {
        PP_CHECK(B, PP_ERR_INVALID, "NULL batch");
        pp_solve_opts o;
        pp_solve_opts_default(&o, PP_SOLVER_CG);
        o.precision = PP_MIXED;
        o.maxit = 12;
        o.tol = 1e-30;
        double *b = dalloc<double>(B->ntot);
        double *x = dalloc<double>(B->ntot);
        std::vector<double> ones(B->ntot, 1.0);
        PP_HIP(hipMemcpy(b, ones.data(), B->ntot * sizeof(double), hipMemcpyHostToDevice));
        BatchArgs a = make_args(B, o, b, x);
        float best = 1e30f;
        int best_bs = B->block;
        for (int bs : {256, 512, 1024}) {
            launch_batch(ctx->stream, bs, true, a);  // warm-up
            PP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
            for (int rep = 0; rep < 3; ++rep) launch_batch(ctx->stream, bs, true, a);
            PP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
            PP_HIP(hipEventSynchronize(ctx->ev1));
            float ms = 0;
            PP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            if (ms < best) {
                best = ms;
                best_bs = bs;
            }
        }
        dfree(b);
        dfree(x);
        B->block = best_bs;
        if (chosen) *chosen = best_bs;
    });
}

void pp_batch_destroy(pp_batch *B) {
    if (!B) return;
    (void)hipSetDevice(B->ctx->device);
    (void)hipStreamSynchronize(B->ctx->stream);
    batch_free(B);
    delete B;
}

}  // extern "C"

namespace pp {
namespace {
__global__ void k_batch_outer(uint32_t n, const uint32_t *rp, const uint32_t *ci, double scale, const double *lam,
                              const double *x, double *g) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double li = scale * lam[i];
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) g[k] = li * x[ci[k]];
    }
}
}  // namespace
}  // namespace pp

extern "C" int pp_batch_outer_on_pattern(pp_ctx *ctx, const pp_batch *B, double scale, const double *lambda,
                                         const double *x, double *grad) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_batch(ctx, [&] {
        PP_CHECK(B && lambda && x && grad, PP_ERR_INVALID, "NULL argument");
        const unsigned g = std::min<uint32_t>((B->ntot + 255) / 256, 4096);
        pp::k_batch_outer<<<std::max(1u, g), 256, 0, ctx->stream>>>(B->ntot, B->rp, B->ci, scale, lambda, x, grad);
        PP_HIP(hipGetLastError());
        PP_HIP(hipStreamSynchronize(ctx->stream));
    });
}
