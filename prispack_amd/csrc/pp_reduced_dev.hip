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
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <tuple>
#include <vector>

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

// n0: the first row holding an entry left of its diagonal
__global__ void k_rd_split(uint32_t n, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                           uint32_t *n0) {
    uint32_t m = UINT32_MAX;
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] < i) {
                m = min(m, i);
                break;
            }
    for (int o = 32; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0 && m != UINT32_MAX) atomicMin(n0, m);
}

// the two-colour structure: one diagonal entry per row, every other entry
// in the other colour
__global__ void k_rd_check(uint32_t n, uint32_t n0, const uint32_t *__restrict__ rp,
                           const uint32_t *__restrict__ ci, int *fail) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        uint32_t diag = 0;
        bool bad = false;
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const uint32_t j = ci[k];
            if (j == i) ++diag;
            else bad |= (i < n0) == (j < n0);
        }
        if (bad || diag != 1) atomicOr(fail, 1);
    }
}

// The operators' patterns (every row of the matrix holds one diagonal):
// KA row i = row i's off-diagonal columns, KE row i = i then those, KB and
// KF row q = row n0 + q's off-diagonal columns then n0 + q; their row
// pointers follow from the matrix's (KA: rp[i] - i, KE: rp[i], KB and KF:
// rp[n0 + q] - rp[n0]).
__global__ void k_rd_patterns(uint32_t n, uint32_t n0, const uint32_t *__restrict__ rp,
                              const uint32_t *__restrict__ ci, uint32_t *ka_rp, uint32_t *ka_ci, uint32_t *ke_rp,
                              uint32_t *ke_ci, uint32_t *kb_rp, uint32_t *kb_ci, uint32_t *kf_rp, uint32_t *kf_ci) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i <= n; i += gridDim.x * kB) {
        if (i <= n0) {
            ka_rp[i] = rp[i] - i;
            ke_rp[i] = rp[i];
        }
        if (i >= n0) kb_rp[i - n0] = kf_rp[i - n0] = rp[i] - rp[n0];
        if (i == n) continue;
        if (i < n0) {
            uint32_t pa = rp[i] - i, pe = rp[i];
            ke_ci[pe++] = i;
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const uint32_t j = ci[k];
                if (j == i) continue;
                ka_ci[pa++] = j;
                ke_ci[pe++] = j;
            }
        } else {
            uint32_t pb = rp[i] - rp[n0];
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const uint32_t j = ci[k];
                if (j == i) continue;
                kb_ci[pb] = j;
                kf_ci[pb++] = j;
            }
            kb_ci[pb] = i;
            kf_ci[pb] = i;
        }
    }
}

// the largest absolute row sum (each row summed in order; NaN rows skipped,
// as std::max skips them); non-negative doubles order as their bit patterns
__global__ void k_rd_anorm(uint32_t n, const uint32_t *__restrict__ rp, const double *__restrict__ v,
                           unsigned long long *out) {
    unsigned long long m = 0;  // the bits of +0.0
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        double t = 0.0;
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) t += fabs(v[k]);
        if (t == t) m = max(m, (unsigned long long)__double_as_longlong(t));
    }
    for (int o = 32; o > 0; o >>= 1) {  // one atomic per wavefront
        const unsigned long long h = (unsigned long long)(uint32_t)__shfl_xor((int)(uint32_t)(m >> 32), o) << 32;
        m = max(m, h | (uint32_t)__shfl_xor((int)(uint32_t)m, o));
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// the highest pattern number in pid[0, n) (one atomic per wavefront: the
// rows mostly share one pattern, so per-row atomics would queue on one word)
__global__ void k_rd_max_pid(uint32_t n, const uint8_t *__restrict__ pid, uint32_t *out) {
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) m = max(m, (uint32_t)pid[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// first[p] = the first row of pid[0, n) with pattern p: the lanes of a
// wavefront sharing a pattern are represented by their lowest lane (its row
// is their smallest), one atomic per pattern and wavefront
__global__ void k_rd_first_seen(uint32_t n, const uint8_t *__restrict__ pid, uint32_t *first) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t base = blockIdx.x * kB + (threadIdx.x & ~63u); base < n; base += gridDim.x * kB) {
        const uint32_t i = base + lane;
        const bool valid = i < n;
        const int p = valid ? (int)pid[i] : -1;
        unsigned long long pending = __ballot(valid);
        while (pending) {
            const int leader = __ffsll((long long)pending) - 1;
            const int lp = __shfl(p, leader);
            const unsigned long long same = __ballot(valid && p == lp) & pending;
            if ((int)lane == leader) atomicMin(first + lp, i);
            pending &= ~same;
        }
    }
}

__global__ void k_rd_remap(uint32_t n, const uint8_t *__restrict__ pid, const uint8_t *__restrict__ map,
                           uint8_t *out) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) out[i] = map[pid[i]];
}

}  // namespace

int reduced_build_device(pp_ctx *ctx, pp_precond *M, pp_csr *As, ReducedForm *R) {
    const uint32_t n = As->nrow;
    if (n != As->ncol || n < 2) return 0;
    uint32_t *u = dalloc<uint32_t>(256 + 2);  // n0, the highest pattern, first rows of the patterns
    int *fail = nullptr;
    auto cleanup = [&] {
        dfree(u);
        dfree(fail);
    };
    try {
        fail = dalloc<int>(1);
        PP_HIP(hipMemsetAsync(u, 0xff, sizeof(uint32_t), ctx->stream));
        PP_HIP(hipMemsetAsync(fail, 0, sizeof(int), ctx->stream));
        k_rd_split<<<grid_of(n), kB, 0, ctx->stream>>>(n, As->rp, As->ci, u);
        PP_HIP(hipGetLastError());
        uint32_t n0 = 0;
        PP_HIP(hipMemcpyAsync(&n0, u, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        if (n0 == 0 || n0 >= n) {
            cleanup();
            return 0;
        }
        k_rd_check<<<grid_of(n), kB, 0, ctx->stream>>>(n, n0, As->rp, As->ci, fail);
        PP_HIP(hipGetLastError());
        int bad = 0;
        PP_HIP(hipMemcpyAsync(&bad, fail, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        // the factor's first level must be exactly the first colour
        if (bad || M->low64.nlev != 2 || M->up64.nlev != 2 || !M->lvl0_contiguous || M->lvl0_lo != 0 ||
            M->lvl0_hi != n0) {
            cleanup();
            return 0;
        }
        const uint32_t n1 = n - n0;
        // the row pointers on the host (row blocks, widths)
        std::vector<uint32_t> hrp;
        if (As->h_rp.size() != (size_t)n + 1) {
            hrp.resize((size_t)n + 1);
            PP_HIP(hipMemcpy(hrp.data(), As->rp, ((size_t)n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
        }
        const std::vector<uint32_t> &rp = hrp.empty() ? As->h_rp : hrp;
        // KA and KB take the sliced-column form of the matrix's own
        // two-colour copy (the same offsets row for row, its pattern numbers
        // renumbered for the second colour) when the widths agree; else the
        // host construction runs
        const EllOp &S = As->ell;
        std::atomic<uint32_t> wmax[2];
        wmax[0] = 0;
        wmax[1] = 0;
        parallel_for(n, 1u << 16, [&](size_t i0, size_t i1) {
            uint32_t m[2] = {0, 0};
            for (size_t i = i0; i < i1; ++i) {
                uint32_t &w = m[i < n0 ? 0 : 1];
                w = std::max(w, rp[i + 1] - rp[i] - 1);
            }
            for (int c = 0; c < 2; ++c) {
                uint32_t cur = wmax[c].load();
                while (m[c] > cur && !wmax[c].compare_exchange_weak(cur, m[c])) {
                }
            }
        });
        const uint32_t wa = wmax[0].load(), wb = wmax[1].load();
        const bool fits = S.ok && S.pid && S.csplit == n0 && S.dsplit == n0 && S.cbase == (int64_t)n0 &&
                          S.cbase1 == -(int64_t)n0 && S.dbase == 0 && wa == S.width && wb == S.width &&
                          (double)wa * n0 <= 1.25 * (double)(rp[n0] - (uint64_t)n0) &&
                          (double)wb * n1 <= 1.25 * (double)(rp[n] - (uint64_t)rp[n0] - n1);
        if (!fits) {
            cleanup();
            return -1;
        }
        R->n0 = n0;
        R->n1 = n1;
        // the operators' arrays, filled before their CSR objects take them
        // over (those compute their row blocks' column forms from them)
        std::vector<uint32_t> h_ka(n0 + 1), h_ke(rp.begin(), rp.begin() + n0 + 1), h_kb(n1 + 1), h_kf(n1 + 1);
        parallel_for((size_t)n + 2, 1u << 16, [&](size_t t0, size_t t1) {
            for (size_t t = t0; t < t1; ++t) {
                if (t <= n0) h_ka[t] = rp[t] - (uint32_t)t;
                else h_kb[t - n0 - 1] = h_kf[t - n0 - 1] = rp[t - 1] - rp[n0];
            }
        });
        struct Raw {
            uint32_t *rp = nullptr, *ci = nullptr;
            double *v = nullptr;
        } ra, re, rb, rf;
        auto raw_free = [&] {
            for (Raw *r : {&ra, &re, &rb, &rf}) {
                dfree(r->rp);
                dfree(r->ci);
                dfree(r->v);
                *r = Raw();
            }
        };
        try {
            for (auto [r, rows, nnz] : {std::make_tuple(&ra, n0, h_ka[n0]), std::make_tuple(&re, n0, h_ke[n0]),
                                        std::make_tuple(&rb, n1, h_kb[n1]), std::make_tuple(&rf, n1, h_kf[n1])}) {
                r->rp = dalloc<uint32_t>((size_t)rows + 1);
                r->ci = dalloc<uint32_t>(std::max<uint32_t>(nnz, 1));
                r->v = dalloc<double>(std::max<uint32_t>(nnz, 1));
            }
            R->s1_64 = dalloc<double>(n1);
            R->dt_64 = dalloc<double>(n1);
            if (!R->dwork) R->dwork = dalloc<double>(n);
            if (!R->dfail) R->dfail = dalloc<int>(1);
        } catch (...) {
            raw_free();
            throw;
        }
        k_rd_patterns<<<grid_of(n + 1), kB, 0, ctx->stream>>>(n, n0, As->rp, As->ci, ra.rp, ra.ci, re.rp, re.ci,
                                                              rb.rp, rb.ci, rf.rp, rf.ci);
        PP_HIP(hipMemsetAsync(R->dfail, 0, sizeof(int), ctx->stream));
        PP_HIP(hipMemsetAsync(u, 0, sizeof(unsigned long long), ctx->stream));
        double *d = R->dwork;
        k_rr_diag<<<grid_of(n), kB, 0, ctx->stream>>>(n, n0, As->rp, As->ci, As->v64, d, R->dfail);
        k_rr_second<<<grid_of(n1), kB, 0, ctx->stream>>>(n0, n1, As->rp, As->ci, As->v64, d, rb.rp, rb.v, rf.rp,
                                                         rf.v, R->dt_64, R->s1_64, R->dfail);
        k_rr_first<<<grid_of(n0), kB, 0, ctx->stream>>>(n0, As->rp, As->ci, As->v64, d, R->dt_64, ra.rp, ra.v, re.rp,
                                                        re.v);
        k_rd_anorm<<<grid_of(n), kB, 0, ctx->stream>>>(n, As->rp, As->v64, (unsigned long long *)u);
        const hipError_t launched = hipGetLastError();
        if (launched != hipSuccess) raw_free();
        PP_HIP(launched);
        // each CSR object owns its arrays from here on (and frees them if it throws)
        auto take = [&](pp_csr *&K, uint32_t rows, Raw &r, std::vector<uint32_t> &&h) {
            Raw t = r;
            r = Raw();
            K = csr_from_device(ctx, rows, n, t.rp, t.ci, t.v, std::move(h));
        };
        try {
            take(R->KA, n0, ra, std::move(h_ka));
            take(R->KE, n0, re, std::move(h_ke));
            take(R->KB, n1, rb, std::move(h_kb));
            take(R->KF, n1, rf, std::move(h_kf));
        } catch (...) {
            raw_free();
            throw;
        }
        unsigned long long an = 0;
        PP_HIP(hipMemcpyAsync(&bad, R->dfail, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipMemcpyAsync(&an, u, sizeof(an), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        if (bad) {  // a pivot that is not positive: no reduced form
            cleanup();
            return 0;
        }
        std::memcpy(&R->anorm, &an, sizeof(double));
        // KA: rows [0, n0) of S, its pattern numbers unchanged (numbered by
        // first appearance, the first colour's patterns are S's first ones)
        EllOp &ka = R->ka, &kb = R->kb;
        ka.n = n0;
        ka.width = S.width;
        ka.cbase = n0;
        ka.dbase = 0;
        PP_HIP(hipMemsetAsync(u + 1, 0, sizeof(uint32_t), ctx->stream));
        k_rd_max_pid<<<grid_of(n0), kB, 0, ctx->stream>>>(n0, S.pid, u + 1);
        PP_HIP(hipMemsetAsync(u + 2, 0xff, 256 * sizeof(uint32_t), ctx->stream));
        k_rd_first_seen<<<grid_of(n1), kB, 0, ctx->stream>>>(n1, S.pid + n0, u + 2);
        PP_HIP(hipGetLastError());
        std::vector<uint32_t> hu(258);
        std::vector<int16_t> ptab((size_t)S.npat * 8);
        PP_HIP(hipMemcpyAsync(hu.data(), u, hu.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipMemcpyAsync(ptab.data(), S.ptab, ptab.size() * sizeof(int16_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        ka.npat = hu[1] + 1;
        ka.pid = dalloc<uint8_t>(n0 + 1);
        ka.ptab = dalloc<int16_t>((size_t)ka.npat * 8);
        PP_HIP(hipMemcpyAsync(ka.pid, S.pid, n0, hipMemcpyDeviceToDevice, ctx->stream));
        PP_HIP(hipMemcpyAsync(ka.ptab, S.ptab, (size_t)ka.npat * 8 * sizeof(int16_t), hipMemcpyDeviceToDevice,
                              ctx->stream));
        ka.v64 = dalloc<double>((size_t)ka.width * n0);
        ka.ok = true;
        // KB: rows [n0, n) of S, patterns renumbered by first appearance there
        std::vector<std::pair<uint32_t, uint32_t>> seen;
        for (uint32_t p = 0; p < S.npat; ++p)
            if (hu[2 + p] != UINT32_MAX) seen.emplace_back(hu[2 + p], p);
        std::sort(seen.begin(), seen.end());
        std::vector<uint8_t> map(256, 0);
        std::vector<int16_t> tb;
        for (size_t r = 0; r < seen.size(); ++r) {
            map[seen[r].second] = (uint8_t)r;
            tb.insert(tb.end(), ptab.begin() + (size_t)seen[r].second * 8, ptab.begin() + (size_t)seen[r].second * 8 + 8);
        }
        kb.n = n1;
        kb.width = S.width;
        kb.cbase = 0;
        kb.dbase = n0;
        kb.npat = (uint32_t)seen.size();
        kb.pid = dalloc<uint8_t>(n1 + 1);
        kb.ptab = dalloc<int16_t>(tb.size());
        uint8_t *dmap = (uint8_t *)(u + 2);  // the first rows are read: the table's room holds the map
        PP_HIP(hipMemcpyAsync(dmap, map.data(), 256, hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipMemcpyAsync(kb.ptab, tb.data(), tb.size() * sizeof(int16_t), hipMemcpyHostToDevice, ctx->stream));
        k_rd_remap<<<grid_of(n1), kB, 0, ctx->stream>>>(n1, S.pid + n0, dmap, kb.pid);
        PP_HIP(hipGetLastError());
        kb.v64 = dalloc<double>((size_t)kb.width * n1);
        kb.dg64 = dalloc<double>(n1);
        kb.ok = true;
        ell_fill_values(ctx, ka, PP_F64, R->KA->rp, R->KA->v64);
        ell_fill_values(ctx, kb, PP_F64, R->KB->rp, R->KB->v64);
        PP_HIP(hipStreamSynchronize(ctx->stream));
        R->ok = true;
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
    return 1;
}

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
    // the largest absolute row sum of the new values
    unsigned long long *an = (unsigned long long *)R->dwork;  // d is no longer read
    PP_HIP(hipMemsetAsync(an, 0, sizeof(unsigned long long), ctx->stream));
    k_rd_anorm<<<grid_of(n), kB, 0, ctx->stream>>>(n, As->rp, As->v64, an);
    PP_HIP(hipGetLastError());
    unsigned long long bits = 0;
    PP_HIP(hipMemcpyAsync(&bits, an, sizeof(bits), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(&R->anorm, &bits, sizeof(double));
    return true;
}

}  // namespace pp
