// pp_kernels.hip -- CDNA4 (gfx950) kernels of the Krylov hot path: CSR SpMV,
// BLAS-1 with deterministic reductions, level-scheduled triangular sweeps and
// the PCG update steps.  64-wide wavefronts, 256-thread workgroups.
#include <hip/hip_runtime.h>
#include <type_traits>

#include <algorithm>
#include <atomic>

#include "pp_internal.h"

namespace pp {

// ---------------------------------------------------------------------------
// Deterministic reductions.  A workgroup reduces with a fixed butterfly inside
// each wavefront and a fixed left-to-right pass over the wavefronts; a
// one-workgroup k_finish launch then sums the per-workgroup partials in index
// order.  The result depends only on the grid shape, never on scheduling.
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Valid in thread 0 only.
__device__ inline double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; ++w) t += sh[w];
    }
    __syncthreads();
    return t;
}

// Three block sums in one LDS round (sh holds 3 * waves doubles); the same
// operation order per value as block_sum.  Valid in thread 0.
__device__ inline void block_sum3(double &a, double &b, double &c, double *sh) {
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    if (lane == 0) {
        sh[wid] = a;
        sh[nw + wid] = b;
        sh[2 * nw + wid] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0, z = 0.0;
        for (int w = 0; w < nw; ++w) {
            x += sh[w];
            y += sh[nw + w];
            z += sh[2 * nw + w];
        }
        a = x;
        b = y;
        c = z;
    }
    __syncthreads();
}

// XCD-aware work split.  Workgroups are dispatched round-robin over the
// eight XCDs (workgroup i runs on XCD i % 8), each with its own L2.  Giving
// XCD q the contiguous range [q nb / 8, (q+1) nb / 8) of row blocks keeps the
// x entries that neighbouring row blocks share in one L2, instead of every
// XCD fetching all of x.  The grid is a multiple of eight when this is used.
constexpr uint32_t kXcds = 8;
struct XcdRange {
    uint32_t first, step, end;
};
__device__ inline XcdRange xcd_range(uint32_t b0, uint32_t b1) {
    const uint32_t nb = b1 - b0, G = gridDim.x;
    if (G % kXcds != 0) return {b0 + blockIdx.x, G, b1};
    const uint32_t q = blockIdx.x % kXcds, per = G / kXcds;
    const uint32_t lo = b0 + (uint32_t)((uint64_t)nb * q / kXcds), hi = b0 + (uint32_t)((uint64_t)nb * (q + 1) / kXcds);
    return {lo + blockIdx.x / kXcds, per, hi};
}
static inline unsigned xcd_grid(uint32_t nblocks, unsigned cap) {
    unsigned g = std::min<unsigned>(nblocks, cap);
    if (g >= kXcds) g -= g % kXcds;
    return std::max(1u, g);
}

// Streamed operand load: non-temporal (NT) or ordinary.
template <bool NT, class V>
__device__ inline V ld_stream(const V *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// Publishes this workgroup's partial sum (plain store).  A separate
// one-workgroup k_finish launch adds the partials in index order: the kernel
// boundary makes them visible, so no workgroup pays a device-wide fence or a
// contended atomic (on a multi-XCD part an agent-scope fence writes back and
// invalidates the L2, which costs more than the whole streaming kernel).
__device__ inline void publish_partial(double blockval, double *partials) {
    if (threadIdx.x == 0) partials[blockIdx.x] = blockval;
}

// Products a_k * x[col_k] of the entries [e0, e0 + cnt) into prod[0..cnt),
// unit stride over the entries (coalesced).  Columns come from 16-bit
// offsets on one of the block's two bases (blo != kNoBase) or from ci.
// BATCH: a thread issues all its value / column loads, then all its x
// gathers, before its first product -- up to kSpmvNnzBlock / kBlock
// independent loads in flight per thread instead of a load-use chain per
// entry, which leaves HBM latency exposed.
template <class T, bool NT, bool BATCH>
__device__ inline void stage_products(T *prod, const T *__restrict__ val, const uint32_t *__restrict__ ci,
                                      const uint16_t *__restrict__ ci16, uint32_t blo, uint32_t bhi, uint32_t e0,
                                      uint32_t cnt, const T *__restrict__ x) {
    if constexpr (BATCH) {
        constexpr int PER = kSpmvNnzBlock / kBlock;
        uint32_t c[PER];
        T v[PER], xv[PER];
        if (blo != kNoBase) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const uint32_t k = threadIdx.x + j * kBlock;
                c[j] = 0;
                v[j] = 0;
                if (k < cnt) {
                    c[j] = ld_stream<NT>(ci16 + e0 + k);
                    v[j] = ld_stream<NT>(val + e0 + k);
                }
            }
#pragma unroll
            for (int j = 0; j < PER; ++j) c[j] = ((c[j] & kHiBase) ? bhi : blo) + (c[j] & kOffMask);
        } else {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const uint32_t k = threadIdx.x + j * kBlock;
                c[j] = 0;
                v[j] = 0;
                if (k < cnt) {
                    c[j] = ld_stream<NT>(ci + e0 + k);
                    v[j] = ld_stream<NT>(val + e0 + k);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            xv[j] = 0;
            if (threadIdx.x + j * kBlock < cnt) xv[j] = x[c[j]];
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t k = threadIdx.x + j * kBlock;
            if (k < cnt) prod[k] = v[j] * xv[j];
        }
    } else if (blo != kNoBase) {
        for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) {
            const uint32_t o = ld_stream<NT>(ci16 + e0 + k);
            prod[k] = ld_stream<NT>(val + e0 + k) * x[((o & kHiBase) ? bhi : blo) + (o & kOffMask)];
        }
    } else {
        for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) prod[k] = ld_stream<NT>(val + e0 + k) * x[ld_stream<NT>(ci + e0 + k)];
    }
}

static inline unsigned grid_for(pp_ctx *ctx, uint64_t n, int per_thread = 1) {
    uint64_t g = (n + (uint64_t)kBlock * per_thread - 1) / ((uint64_t)kBlock * per_thread);
    uint64_t cap = (uint64_t)ctx->cu_count * 8;
    if (cap > (uint64_t)kMaxPartials) cap = kMaxPartials;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

// ---------------------------------------------------------------------------
// CSR SpMV, "CSR-stream": each workgroup owns a row block whose nonzeros fit
// in LDS.  Phase 1 streams values and column indices with unit stride
// (coalesced) and stores the products a_k * x[col_k]; phase 2 gives every row
// one thread that sums its products in storage order starting from zero --
// the exact operation sequence of CSRMatrix::MatPVecWithRawPtr
// (CSRMatrix.h:360-377), so the FP64 result is bit-identical to it.  A row
// longer than the LDS stage gets a whole workgroup and a tree sum.
// Optional fused dot: partial sum of w_row * y_row.
// ---------------------------------------------------------------------------
template <class T, bool DOT, bool NT, bool BATCH>
__global__ __launch_bounds__(256) void k_spmv(const BlockDesc *__restrict__ desc,
                                              const uint32_t *__restrict__ rp,
                                              const uint32_t *__restrict__ ci,
                                              const uint16_t *__restrict__ ci16,
                                              const T *__restrict__ val, const T *__restrict__ x,
                                              T *__restrict__ y, uint32_t nb, T alpha, T beta, int beta_zero,
                                              const T *__restrict__ dotw, const T *__restrict__ aux,
                                              const T *__restrict__ wgt, double *partials, const int *stop) {
    __shared__ T prod[kSpmvNnzBlock];
    __shared__ double sh[3 * (kBlock / 64)];
    if (stop && *stop) return;
    // DOT: (dotw, y); with aux also (aux, dotw) and (aux, aux) into the next
    // two partial rows (the three reductions of the single-reduction PCG);
    // wgt weights the last one: sum wgt_i aux_i^2
    double dacc = 0.0, dacc2 = 0.0, dacc3 = 0.0;
    const XcdRange xr = xcd_range(0, nb);
    for (uint32_t b = xr.first; b < xr.end; b += xr.step) {
        const BlockDesc d0 = desc[b], d1 = desc[b + 1];  // one scalar load each, no chain
        const uint32_t r0 = d0.r0, r1 = d1.r0;
        const uint32_t k0 = d0.k0, k1 = d1.k0;
        const uint32_t cnt = k1 - k0;
        if (cnt <= kSpmvNnzBlock) {
            // the row operands of phase 2 are loaded first, so their latency
            // overlaps the product stage
            const bool mine = threadIdx.x < r1 - r0;
            const uint32_t row = r0 + threadIdx.x;
            uint32_t pa = 0, pe = 0;
            T yold = 0, dw = 0, ax = 0, wg = 1;
            if (mine) {
                pa = rp[row];
                pe = rp[row + 1];
                if (!beta_zero) yold = y[row];
                if (DOT) {
                    dw = dotw[row];
                    if (aux) ax = aux[row];
                    if (wgt) wg = wgt[row];
                }
            }
            // products a_k * x[col_k] into LDS (uniform choice of offsets)
            stage_products<T, NT, BATCH>(prod, val, ci, ci16, ci16 ? d0.blo : kNoBase, d0.bhi, k0, cnt, x);
            __syncthreads();
            if (mine) {
                uint32_t a = pa - k0;
                const uint32_t e = pe - k0;
                T s = 0;
                for (; a < e; ++a) s += prod[a];
                T out = alpha * s;
                if (!beta_zero) out += beta * yold;
                if constexpr (NT) __builtin_nontemporal_store(out, y + row);  // streamed out, not re-read here
                else y[row] = out;
                if (DOT) {
                    const double u = (double)dw;
                    dacc += u * (double)out;
                    if (aux) {
                        const double rr = (double)ax;
                        dacc2 += rr * u;
                        dacc3 += rr * rr * (double)wg;
                    }
                }
            }
            __syncthreads();
        } else {
            double s = 0.0;
            for (uint32_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) s += (double)(val[k] * x[ci[k]]);
            s = block_sum(s, sh);
            if (threadIdx.x == 0) {
                T out = alpha * (T)s;
                if (!beta_zero) out += beta * y[r0];
                y[r0] = out;
                if (DOT) {
                    const double u = (double)dotw[r0];
                    dacc += u * (double)out;
                    if (aux) {
                        const double rr = (double)aux[r0];
                        dacc2 += rr * u;
                        dacc3 += rr * rr * (wgt ? (double)wgt[r0] : 1.0);
                    }
                }
            }
            __syncthreads();
        }
    }
    if (DOT) {
        if (aux) {
            block_sum3(dacc, dacc2, dacc3, sh);
            publish_partial(dacc, partials);
            publish_partial(dacc2, partials + kMaxPartials);
            publish_partial(dacc3, partials + 2 * kMaxPartials);
        } else {
            publish_partial(block_sum(dacc, sh), partials);
        }
    }
}

// Entry-slot count of the sliced-column kernels.  W > 0: exactly W slots,
// known at compile time, so every slot's loads are issued unconditionally
// and back to back; W < 0: the runtime width, at most -W slots, each behind
// a test (measured 22% slower on the 128^3 7-point SpMV, 43.1 against
// 35.3 us).  ell_width_dispatch picks the exact form for the widths of the
// 7- and 5-point stencils (natural and two-colour orders).
// The offsets of a row from its pattern (EllOp::pid / ptab): one 16-byte
// table entry, eight 16-bit slots.
template <int MAXW>
__device__ inline void ell_unpack(const int4 t, int16_t *o) {
    const uint32_t w[4] = {(uint32_t)t.x, (uint32_t)t.y, (uint32_t)t.z, (uint32_t)t.w};
#pragma unroll
    for (int j = 0; j < MAXW; ++j) o[j] = (int16_t)((w[j / 2] >> (16 * (j & 1))) & 0xffffu);
}
template <int MAXW>
__device__ inline void ell_pattern(const int16_t *__restrict__ ptab, uint32_t p, int16_t *o) {
    ell_unpack<MAXW>(*reinterpret_cast<const int4 *>(ptab + 8 * (size_t)p), o);
}

template <int W>
struct EllWidth {
    static constexpr int kMax = W > 0 ? W : -W;
    __device__ static inline bool has(int j, uint32_t width) { return W > 0 || j < (int)width; }
};
// The same for matrices whose rows use offset patterns (widths 4 to 7 only).
#define PP_ELL_PAT_WIDTHS(w, CALL)             \
    switch (w) {                               \
        case 4: CALL(4); break;                \
        case 5: CALL(5); break;                \
        case 6: CALL(6); break;                \
        default: CALL(7); break;               \
    }
#define PP_ELL_WIDTHS(w, CALL)                 \
    switch (w) {                               \
        case 4: CALL(4); break;                \
        case 5: CALL(5); break;                \
        case 6: CALL(6); break;                \
        case 7: CALL(7); break;                \
        default:                               \
            if ((w) <= 8) CALL(-8);            \
            else CALL(-16);                    \
    }

// ---------------------------------------------------------------------------
// Sliced-column SpMV (EllOp): one row per thread, entry j of all rows
// contiguous, so every value and offset load is coalesced and no LDS stage
// or barrier is needed.  Each row sums its products in storage order
// starting from zero, the diagonal term last, each product rounded before
// its sum, then y = alpha * sum (+ beta * y): the operation sequence of
// k_spmv on the CSR form, so the results are the same bit for bit.  x and y
// may be one vector when the rows write where no row reads (the reduced
// CG's KA: first colour from second).  DOT: the same sums as k_spmv.
// ---------------------------------------------------------------------------
template <class T, bool DOT, int W, bool PAT>
__global__ __launch_bounds__(256) void k_ell(uint32_t n, uint32_t width, const int16_t *__restrict__ off,
                                             const uint8_t *__restrict__ pid, const int16_t *__restrict__ ptab,
                                             const T *__restrict__ val, int64_t cbase, const T *__restrict__ dg,
                                             int64_t dbase, uint32_t csplit, int64_t cbase1, uint32_t dsplit,
                                             const T *x, T *y, T alpha, T beta, int beta_zero,
                                             const T *__restrict__ dotw, const T *__restrict__ aux,
                                             const T *__restrict__ wgt, double *partials, const int *stop) {
    __shared__ double sh[3 * (kBlock / 64)];
    if (stop && *stop) return;
    double dacc = 0.0, dacc2 = 0.0, dacc3 = 0.0;
    const uint32_t nb = (n + kBlock - 1) / kBlock;
    const XcdRange xr = xcd_range(0, nb);
    for (uint32_t b = xr.first; b < xr.end; b += xr.step) {
        const uint32_t i = b * kBlock + threadIdx.x;
        if (i < n) {
            constexpr int MAXW = EllWidth<W>::kMax;
            int16_t o[MAXW];
            T v[MAXW], xv[MAXW];
            if constexpr (PAT) ell_pattern<MAXW>(ptab, pid[i], o);
#pragma unroll
            for (int j = 0; j < MAXW; ++j)
                if (EllWidth<W>::has(j, width)) {
                    const size_t e = (size_t)j * n + i;
                    if constexpr (!PAT) o[j] = off[e];
                    v[j] = val[e];
                }
            T dgi = 0, xd = 0, yold = 0, u = 0, ax = 0, wg = 1;
            if (dg) {
                dgi = dg[i];
                xd = x[dbase + i];
            }
            if (!beta_zero) yold = y[i];
            if (DOT) {
                u = dotw[i];
                if (aux) ax = aux[i];
                if (wgt) wg = wgt[i];
            }
            const int64_t cb = i < csplit ? cbase : cbase1;
            const bool dfirst = i < dsplit;
#pragma unroll
            for (int j = 0; j < MAXW; ++j)
                if (EllWidth<W>::has(j, width)) xv[j] = x[(int64_t)i + cb + o[j]];
            T sum = 0;
            {
#pragma clang fp contract(off)
                if (dg && dfirst) sum += dgi * xd;
#pragma unroll
                for (int j = 0; j < MAXW; ++j)
                    if (EllWidth<W>::has(j, width)) sum += v[j] * xv[j];
                if (dg && !dfirst) sum += dgi * xd;
            }
            T out = alpha * sum;
            if (!beta_zero) out += beta * yold;
            y[i] = out;
            if (DOT) {
                dacc += (double)u * (double)out;
                if (aux) {
                    dacc2 += (double)ax * (double)u;
                    dacc3 += (double)ax * (double)ax * (double)wg;
                }
            }
        }
    }
    if (DOT) {
        if (aux) {
            block_sum3(dacc, dacc2, dacc3, sh);
            publish_partial(dacc, partials);
            publish_partial(dacc2, partials + kMaxPartials);
            publish_partial(dacc3, partials + 2 * kMaxPartials);
        } else {
            publish_partial(block_sum(dacc, sh), partials);
        }
    }
}

// The same product with two consecutive rows per thread (no sums): each
// thread loads the pair's offsets as one 32-bit word and the pair's values
// as one 2-element vector per entry slot, and stores the two results as one
// vector, so a wavefront moves twice the bytes per load instruction.  Each
// row's operation sequence is the one above, so y is the same bit for bit.
// Needs an even row count and y aligned to two elements (checked by the
// launcher).
template <class T, int W, bool PAT>
__global__ __launch_bounds__(256) void k_ell2(uint32_t n, uint32_t width, const int16_t *__restrict__ off,
                                              const uint8_t *__restrict__ pid, const int16_t *__restrict__ ptab,
                                              const T *__restrict__ val, int64_t cbase, const T *__restrict__ dg,
                                              int64_t dbase, uint32_t csplit, int64_t cbase1, uint32_t dsplit,
                                              const T *x, T *y, T alpha, T beta, int beta_zero, const int *stop) {
    using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
    if (stop && *stop) return;
    const uint32_t npair = n / 2;
    const uint32_t nb = (npair + kBlock - 1) / kBlock;
    const XcdRange xr = xcd_range(0, nb);
    for (uint32_t b = xr.first; b < xr.end; b += xr.step) {
        const uint32_t t = b * kBlock + threadIdx.x;
        if (t >= npair) continue;
        const uint32_t i = 2 * t;
        constexpr int MAXW = EllWidth<W>::kMax;
        int16_t o0[MAXW], o1[MAXW];
        T v0[MAXW], v1[MAXW], x0[MAXW], x1[MAXW];
        if constexpr (PAT) {
            const uint32_t pp = *reinterpret_cast<const uint16_t *>(pid + i);
            ell_pattern<MAXW>(ptab, pp & 0xffu, o0);
            ell_pattern<MAXW>(ptab, pp >> 8, o1);
        }
#pragma unroll
        for (int j = 0; j < MAXW; ++j)
            if (EllWidth<W>::has(j, width)) {
                const size_t e = (size_t)j * n + i;
                if constexpr (!PAT) {
                    const uint32_t ow = *reinterpret_cast<const uint32_t *>(off + e);
                    o0[j] = (int16_t)(ow & 0xffffu);
                    o1[j] = (int16_t)(ow >> 16);
                }
                const T2 vv = *reinterpret_cast<const T2 *>(val + e);
                v0[j] = vv.x;
                v1[j] = vv.y;
            }
        T dg0 = 0, dg1 = 0, xd0 = 0, xd1 = 0, yo0 = 0, yo1 = 0;
        if (dg) {
            dg0 = dg[i];
            dg1 = dg[i + 1];
            xd0 = x[dbase + i];
            xd1 = x[dbase + i + 1];
        }
        if (!beta_zero) {
            const T2 yy = *reinterpret_cast<const T2 *>(y + i);
            yo0 = yy.x;
            yo1 = yy.y;
        }
        const int64_t cb0 = i < csplit ? cbase : cbase1, cb1 = i + 1 < csplit ? cbase : cbase1;
        const bool df0 = i < dsplit, df1 = i + 1 < dsplit;
#pragma unroll
        for (int j = 0; j < MAXW; ++j)
            if (EllWidth<W>::has(j, width)) {
                x0[j] = x[(int64_t)i + cb0 + o0[j]];
                x1[j] = x[(int64_t)i + 1 + cb1 + o1[j]];
            }
        T s0 = 0, s1 = 0;
        {
#pragma clang fp contract(off)
            if (dg) {  // a diagonal that comes first in its row
                if (df0) s0 += dg0 * xd0;
                if (df1) s1 += dg1 * xd1;
            }
#pragma unroll
            for (int j = 0; j < MAXW; ++j)
                if (EllWidth<W>::has(j, width)) {
                    s0 += v0[j] * x0[j];
                    s1 += v1[j] * x1[j];
                }
            if (dg) {
                if (!df0) s0 += dg0 * xd0;
                if (!df1) s1 += dg1 * xd1;
            }
        }
        T r0 = alpha * s0, r1 = alpha * s1;
        if (!beta_zero) {
            r0 += beta * yo0;
            r1 += beta * yo1;
        }
        T2 out;
        out.x = r0;
        out.y = r1;
        *reinterpret_cast<T2 *>(y + i) = out;
    }
}

// With the sums, at most 1024 workgroups, each over several 256-row
// slices: fewer partials for the consumer to add, measured 2% faster per
// reduced iteration than one slice per workgroup at 1M rows (512: slower).
unsigned ell_grid(const EllOp &E, bool dot) {
    constexpr unsigned kEllDotGroups = 1024;
    const uint32_t nb = (E.n + kBlock - 1) / kBlock;
    return nb ? xcd_grid(nb, dot ? kEllDotGroups : (1u << 20)) : 0;
}

// grid: 0 for ell_grid's, else that many workgroups (a caller whose
// consumers expect a given number of partials)
void launch_ell(pp_ctx *ctx, const EllOp &E, int dtype, double alpha, const void *x, double beta, void *y,
                int dot_slot, const void *dotw, const void *aux, const void *wgt, unsigned grid) {
    const bool dot = dot_slot >= 0;
    // two rows per thread where the pair loads and the pair store are aligned
    const size_t esz = dtype == PP_F64 ? 8 : 4;
    if (!dot && grid == 0 && E.n >= 2 && E.n % 2 == 0 && ((uintptr_t)y % (2 * esz)) == 0) {
        const uint32_t nb2 = (E.n / 2 + kBlock - 1) / kBlock;
        const unsigned g2 = xcd_grid(nb2, 1u << 20);
        const int bz2 = beta == 0.0;
#define PP_ELL2(T, W, P, VALS, DG)                                                                             \
    k_ell2<T, W, P><<<g2, kBlock, 0, ctx->stream>>>(E.n, E.width, E.off, E.pid, E.ptab, VALS, E.cbase, DG, E.dbase, \
                                                     E.csplit, E.cbase1, E.dsplit, (const T *)x, (T *)y, (T)alpha, \
                                                     (T)beta, bz2, ctx->stop_word)
#define PP_ELL2_64(W) PP_ELL2(double, W, false, E.v64, E.dg64)
#define PP_ELL2_32(W) PP_ELL2(float, W, false, E.v32, E.dg32)
#define PP_ELL2_64P(W) PP_ELL2(double, W, true, E.v64, E.dg64)
#define PP_ELL2_32P(W) PP_ELL2(float, W, true, E.v32, E.dg32)
        if (E.pid) {
            if (dtype == PP_F64) {
                PP_ELL_PAT_WIDTHS(E.width, PP_ELL2_64P)
            } else {
                PP_ELL_PAT_WIDTHS(E.width, PP_ELL2_32P)
            }
        } else if (dtype == PP_F64) {
            PP_ELL_WIDTHS(E.width, PP_ELL2_64)
        } else {
            PP_ELL_WIDTHS(E.width, PP_ELL2_32)
        }
#undef PP_ELL2_64
#undef PP_ELL2_32
#undef PP_ELL2_64P
#undef PP_ELL2_32P
#undef PP_ELL2
        PP_HIP(hipGetLastError());
        return;
    }
    if (grid == 0) grid = ell_grid(E, dot);
    if (grid == 0) return;
    double *part = dot ? ctx->red.partials + (size_t)dot_slot * kMaxPartials : nullptr;
    const int bz = beta == 0.0;
#define PP_ELL(T, DOTV, W, P, VALS, DG)                                                                       \
    k_ell<T, DOTV, W, P><<<grid, kBlock, 0, ctx->stream>>>(E.n, E.width, E.off, E.pid, E.ptab, VALS, E.cbase, DG, \
                                                          E.dbase, E.csplit, E.cbase1, E.dsplit, (const T *)x,  \
                                                          (T *)y, (T)alpha, (T)beta, bz, (const T *)dotw,       \
                                                          (const T *)aux, (const T *)wgt, part, ctx->stop_word)
#define PP_ELL_64D(W) PP_ELL(double, true, W, false, E.v64, E.dg64)
#define PP_ELL_64N(W) PP_ELL(double, false, W, false, E.v64, E.dg64)
#define PP_ELL_32D(W) PP_ELL(float, true, W, false, E.v32, E.dg32)
#define PP_ELL_32N(W) PP_ELL(float, false, W, false, E.v32, E.dg32)
#define PP_ELL_64DP(W) PP_ELL(double, true, W, true, E.v64, E.dg64)
#define PP_ELL_64NP(W) PP_ELL(double, false, W, true, E.v64, E.dg64)
#define PP_ELL_32DP(W) PP_ELL(float, true, W, true, E.v32, E.dg32)
#define PP_ELL_32NP(W) PP_ELL(float, false, W, true, E.v32, E.dg32)
    if (E.pid) {
        if (dtype == PP_F64) {
            if (dot) {
                PP_ELL_PAT_WIDTHS(E.width, PP_ELL_64DP)
            } else {
                PP_ELL_PAT_WIDTHS(E.width, PP_ELL_64NP)
            }
        } else {
            if (dot) {
                PP_ELL_PAT_WIDTHS(E.width, PP_ELL_32DP)
            } else {
                PP_ELL_PAT_WIDTHS(E.width, PP_ELL_32NP)
            }
        }
    } else if (dtype == PP_F64) {
        if (dot) {
            PP_ELL_WIDTHS(E.width, PP_ELL_64D)
        } else {
            PP_ELL_WIDTHS(E.width, PP_ELL_64N)
        }
    } else {
        if (dot) {
            PP_ELL_WIDTHS(E.width, PP_ELL_32D)
        } else {
            PP_ELL_WIDTHS(E.width, PP_ELL_32N)
        }
    }
#undef PP_ELL_64D
#undef PP_ELL_64N
#undef PP_ELL_32D
#undef PP_ELL_32N
#undef PP_ELL_64DP
#undef PP_ELL_64NP
#undef PP_ELL_32DP
#undef PP_ELL_32NP
#undef PP_ELL
    PP_HIP(hipGetLastError());
}

template <class T>
__global__ void k_ell_fill(uint32_t n, uint32_t width, const uint32_t *__restrict__ rp, const T *__restrict__ v,
                           T *__restrict__ out, uint32_t tail) {
    const size_t total = (size_t)width * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const uint32_t j = (uint32_t)(e / n), i = (uint32_t)(e - (size_t)j * n);
        const uint32_t k = rp[i] + j;
        out[e] = k < rp[i + 1] - tail ? v[k] : (T)0;
    }
}

// The same for a form with a separate diagonal: rows below dsplit hold it
// first (the slots are their other entries), the rest last; dg refilled too
template <class T>
__global__ void k_ell_fill_dg(uint32_t n, uint32_t width, uint32_t dsplit, const uint32_t *__restrict__ rp,
                              const T *__restrict__ v, T *__restrict__ out, T *__restrict__ dg) {
    const size_t total = (size_t)width * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const uint32_t j = (uint32_t)(e / n), i = (uint32_t)(e - (size_t)j * n);
        const bool first = i < dsplit;
        const uint32_t k = rp[i] + (first ? 1 : 0) + j, k1 = rp[i + 1] - (first ? 0 : 1);
        out[e] = k < k1 ? v[k] : (T)0;
        if (j == 0) dg[i] = v[first ? rp[i] : rp[i + 1] - 1];
    }
}

// out[j * n + i] from the CSR values v of rows i (row pointers rp), on the
// null stream, waited for (setup only)
template <class T>
void ell_fill_rows(uint32_t n, uint32_t width, const uint32_t *rp, const T *v, T *out) {
    const uint64_t total = (uint64_t)width * n;
    const unsigned g = (unsigned)std::min<uint64_t>((total + kBlock * 4 - 1) / (kBlock * 4), 8192);
    k_ell_fill<T><<<std::max(1u, g), kBlock, 0, 0>>>(n, width, rp, v, out, 0u);
    PP_HIP(hipGetLastError());
    PP_HIP(hipStreamSynchronize(0));
}
template void ell_fill_rows<double>(uint32_t, uint32_t, const uint32_t *, const double *, double *);
template void ell_fill_rows<float>(uint32_t, uint32_t, const uint32_t *, const float *, float *);

void ell_fill_values(pp_ctx *ctx, const EllOp &E, int dtype, const uint32_t *rp, const void *v) {
    const unsigned g = grid_for(ctx, (uint64_t)E.width * E.n, 4);
    if (E.dg64) {  // a separate diagonal: out of the slots, into dg
        if (dtype == PP_F64)
            k_ell_fill_dg<double><<<g, kBlock, 0, ctx->stream>>>(E.n, E.width, E.dsplit, rp, (const double *)v, E.v64,
                                                                 E.dg64);
        else
            k_ell_fill_dg<float><<<g, kBlock, 0, ctx->stream>>>(E.n, E.width, E.dsplit, rp, (const float *)v, E.v32,
                                                                E.dg32);
    } else if (dtype == PP_F64) {
        k_ell_fill<double><<<g, kBlock, 0, ctx->stream>>>(E.n, E.width, rp, (const double *)v, E.v64, 0u);
    } else {
        k_ell_fill<float><<<g, kBlock, 0, ctx->stream>>>(E.n, E.width, rp, (const float *)v, E.v32, 0u);
    }
    PP_HIP(hipGetLastError());
}

void build_row_blocks(const std::vector<uint32_t> &rp, std::vector<uint32_t> &blocks) {
    blocks.clear();
    const uint32_t n = (uint32_t)rp.size() - 1;
    // rows of at most kSpmvNnzBlock / kSpmvRowsBlock entries: every block
    // takes kSpmvRowsBlock rows (the rule below, without its serial pass)
    std::atomic<uint32_t> longest{0};
    parallel_for(n, 1u << 16, [&](size_t i0, size_t i1) {
        uint32_t m = 0;
        for (size_t i = i0; i < i1; ++i) m = std::max(m, rp[i + 1] - rp[i]);
        uint32_t cur = longest.load();
        while (m > cur && !longest.compare_exchange_weak(cur, m)) {
        }
    });
    if (longest.load() <= kSpmvNnzBlock / kSpmvRowsBlock) {
        blocks.reserve(n / kSpmvRowsBlock + 2);
        for (uint32_t r = 0; r < n; r += kSpmvRowsBlock) blocks.push_back(r);
        blocks.push_back(n);
        return;
    }
    uint32_t r = 0;
    blocks.push_back(0);
    while (r < n) {
        uint32_t start = r;
        uint32_t nnz0 = rp[r];
        if (rp[r + 1] - nnz0 > kSpmvNnzBlock) {
            r = r + 1;  // a long row alone
        } else {
            while (r < n && r - start < kSpmvRowsBlock && rp[r + 1] - nnz0 <= kSpmvNnzBlock) ++r;
        }
        blocks.push_back(r);
    }
}

// result = sum of partials[0..np) in index order (one workgroup); op
// FIN_PCG_CHECK also runs the PCG convergence test on it.
enum FinishOp : int { FIN_SUM = 0, FIN_PCG_CHECK = 1 };
__device__ void pcg_check(double rr, const double *scal, int *ctl, double *hist, uint32_t hist_cap);
// ring != NULL (the last check of an iteration body): afterwards the control
// words are copied into the next mapped pinned slot for the host (plain
// vector stores over the fabric), also when the body was skipped.  The
// slot's C_PUB word, written last behind a system-scope fence, is the
// body's sequence number: the host polls it instead of waiting on an event.
__device__ inline void publish_ctl(int *ctl, int *ring) {
    const int seq = ctl[C_PUB] + 1;
    ctl[C_PUB] = seq;
    int *slot = ring + ((seq - 1) % pp_ctx::kRing) * C_COUNT;
    for (int i = 0; i < C_COUNT; ++i)
        if (i != C_PUB) slot[i] = ctl[i];
    __threadfence_system();
    slot[C_PUB] = seq;
}

// One workgroup of kFinishThreads: each thread loads its (at most four)
// partials before adding them, so the loads overlap instead of forming a
// chain of memory latencies.
constexpr int kFinishThreads = 1024;
__device__ inline double sum_partials(const double *__restrict__ partials, unsigned np) {
    double v[kMaxPartials / kFinishThreads];
#pragma unroll
    for (int j = 0; j < kMaxPartials / kFinishThreads; ++j) {
        const unsigned i = threadIdx.x + j * kFinishThreads;
        v[j] = i < np ? partials[i] : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < kMaxPartials / kFinishThreads; ++j) s += v[j];
    return s;
}

__global__ __launch_bounds__(1024) void k_finish(const double *__restrict__ partials, unsigned np, double *result,
                                                 const int *stop, int op, const double *scal, int *ctl, double *hist,
                                                 uint32_t hist_cap, int *ring) {
    __shared__ double sh[kFinishThreads / 64];
    if (stop && *stop) {
        if (ring && threadIdx.x == 0) publish_ctl(ctl, ring);
        return;
    }
    double s = block_sum(sum_partials(partials, np), sh);
    if (threadIdx.x == 0) {
        *result = s;
        if (op == FIN_PCG_CHECK) pcg_check(s, scal, ctl, hist, hist_cap);
        if (ring) publish_ctl(ctl, ring);
    }
}

static void launch_finish(pp_ctx *ctx, int slot, unsigned np, const int *stop, int op = FIN_SUM,
                          double *hist = nullptr, uint32_t hist_cap = 0, int *ring = nullptr) {
    k_finish<<<1, kFinishThreads, 0, ctx->stream>>>(ctx->red.partials + (size_t)slot * kMaxPartials, np,
                                                    ctx->scal + slot, stop, op, ctx->scal, ctx->ctl, hist, hist_cap,
                                                    ring);
}

void launch_spmv(pp_ctx *ctx, const pp_csr *A, int dtype, double alpha, const void *x, double beta,
                 void *y, int dot_slot, const void *dot_with, bool, const void *aux, const void *wgt) {
    // one row block per workgroup; a launch with inner products keeps to
    // kMaxPartials workgroups (one partial each), looping over blocks
    const bool dotl = dot_slot >= 0;
    unsigned grid = A->nblocks ? xcd_grid(A->nblocks, dotl ? (unsigned)kMaxPartials : (1u << 20)) : 0;
    if (grid == 0) return;
    double *part = dot_slot >= 0 ? ctx->red.partials + (size_t)dot_slot * kMaxPartials : nullptr;
    const int bz = beta == 0.0;
    const bool dot = dot_slot >= 0;
    const int var = dtype == PP_F64 ? A->spmv_variant : A->spmv_variant32;
    if (var == kVarEll) {  // with sums: as many workgroups (partials) as the CSR-stream launch
        launch_ell(ctx, A->ell, dtype, alpha, x, beta, y, dot_slot, dot_with, aux, wgt, dot ? grid : 0);
        if (dot && !aux) launch_finish(ctx, dot_slot, grid, ctx->stop_word);  // with aux the caller finishes
        PP_HIP(hipGetLastError());
        return;
    }
    const uint16_t *c16 = (var & kVarIdx16) ? A->ci16 : nullptr;
#define PP_SPMV(T, DOTV, NTV, BV, VALS, DW, PART)                                                                 \
    k_spmv<T, DOTV, NTV, BV><<<grid, kBlock, 0, ctx->stream>>>(A->blk_desc, A->rp, A->ci, c16, VALS, (const T *)x, \
                                                              (T *)y, A->nblocks, (T)alpha, (T)beta, bz, DW,     \
                                                              (const T *)aux, (const T *)wgt, PART,          \
                                                              ctx->stop_word)
#define PP_SPMV_VAR(T, DOTV, VALS, DW, PART)                                           \
    switch (var & (kVarNT | kVarBatch)) {                                               \
        case 0: PP_SPMV(T, DOTV, false, false, VALS, DW, PART); break;                  \
        case kVarNT: PP_SPMV(T, DOTV, true, false, VALS, DW, PART); break;              \
        case kVarBatch: PP_SPMV(T, DOTV, false, true, VALS, DW, PART); break;           \
        default: PP_SPMV(T, DOTV, true, true, VALS, DW, PART); break;                   \
    }
    if (dtype == PP_F64) {
        const double *dw = (const double *)dot_with;
        if (dot) PP_SPMV_VAR(double, true, A->v64, dw, part)
        else PP_SPMV_VAR(double, false, A->v64, nullptr, nullptr)
    } else {
        const float *dw = (const float *)dot_with;
        if (dot) PP_SPMV_VAR(float, true, A->v32, dw, part)
        else PP_SPMV_VAR(float, false, A->v32, nullptr, nullptr)
    }
#undef PP_SPMV_VAR
#undef PP_SPMV
    if (dot && !aux) launch_finish(ctx, dot_slot, grid, ctx->stop_word);  // with aux the caller finishes
    PP_HIP(hipGetLastError());
}

// Reads a buffer larger than the last-level cache (256 MB on MI355X) so that
// what a timed launch finds cached is what a solve leaves there: nothing of
// its own operands.  The store never happens (the buffer holds zeros); it
// only keeps the loads alive.
__global__ __launch_bounds__(256) void k_flush(const uint4 *__restrict__ p, size_t n, uint32_t *sink) {
    uint32_t acc = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint4 v = p[i];  // ordinary loads: they must take the cache lines
        acc |= v.x | v.y | v.z | v.w;
    }
    if (acc == 0xFFFFFFFFu) *sink = acc;
}

float time_cold_ms(pp_ctx *ctx, const std::function<void()> &fn, int reps, bool cold) {
    constexpr size_t kFlushBytes = 384ull << 20;
    if (cold && !ctx->flush_buf) {
        ctx->flush_buf = dalloc<uint4>(kFlushBytes / sizeof(uint4) + 1);
        launch_fill_zero(ctx, ctx->flush_buf, kFlushBytes + sizeof(uint4));
    }
    uint4 *buf = (uint4 *)ctx->flush_buf;
    float best = 0.f;
    fn();  // warm-up (code, TLB)
    for (int r = 0; r < reps; ++r) {
        if (cold)
            k_flush<<<ctx->cu_count * 8, 256, 0, ctx->stream>>>(buf, kFlushBytes / sizeof(uint4),
                                                                 (uint32_t *)(buf + kFlushBytes / sizeof(uint4)));
        PP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        fn();
        PP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        PP_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        PP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        if (r == 0 || ms < best) best = ms;
    }
    return best;
}

// ---------------------------------------------------------------------------
// BLAS-1
// ---------------------------------------------------------------------------
__global__ void k_fill_zero(uint32_t *p, size_t nwords) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

void launch_fill_zero(pp_ctx *ctx, void *p, size_t bytes) {
    size_t words = bytes / 4;
    if (words) k_fill_zero<<<grid_for(ctx, words), kBlock, 0, ctx->stream>>>((uint32_t *)p, words);
    if (bytes % 4) PP_HIP(hipMemsetAsync((char *)p + words * 4, 0, bytes % 4, ctx->stream));
    PP_HIP(hipGetLastError());
}

// Streaming copy: 16-byte non-temporal loads and stores, four in flight per
// thread.  The reference point of the HBM roofline (what a kernel that only
// moves bytes reaches on this device).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_stream(size_t n16, const u32x4 *__restrict__ s, u32x4 *__restrict__ d) {
    const size_t step = (size_t)gridDim.x * kBlock * 4;
    for (size_t i = (size_t)blockIdx.x * kBlock * 4 + threadIdx.x; i < n16; i += step) {
        u32x4 a[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + (size_t)k * kBlock < n16) a[k] = __builtin_nontemporal_load(s + i + (size_t)k * kBlock);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + (size_t)k * kBlock < n16) __builtin_nontemporal_store(a[k], d + i + (size_t)k * kBlock);
    }
}

void launch_copy_stream(pp_ctx *ctx, const void *src, void *dst, size_t bytes, unsigned groups) {
    const size_t n16 = bytes / 16;
    if (groups == 0) groups = (unsigned)ctx->cu_count * 64;
    if (n16) k_copy_stream<<<groups, kBlock, 0, ctx->stream>>>(n16, (const u32x4 *)src, (u32x4 *)dst);
    if (bytes % 16)
        PP_HIP(hipMemcpyAsync((char *)dst + n16 * 16, (const char *)src + n16 * 16, bytes % 16,
                              hipMemcpyDeviceToDevice, ctx->stream));
    PP_HIP(hipGetLastError());
}

template <class S, class D>
__global__ void k_convert(uint32_t n, const S *s, D *d, const int *stop) {
    if (stop && *stop) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = (D)s[i];
}

void launch_convert(pp_ctx *ctx, uint32_t n, int sdt, const void *s, int ddt, void *d) {
    unsigned g = grid_for(ctx, n);
    if (sdt == PP_F64 && ddt == PP_F32)
        k_convert<double, float><<<g, kBlock, 0, ctx->stream>>>(n, (const double *)s, (float *)d, ctx->stop_word);
    else if (sdt == PP_F32 && ddt == PP_F64)
        k_convert<float, double><<<g, kBlock, 0, ctx->stream>>>(n, (const float *)s, (double *)d, ctx->stop_word);
    else if (sdt == PP_F64)
        k_convert<double, double><<<g, kBlock, 0, ctx->stream>>>(n, (const double *)s, (double *)d, ctx->stop_word);
    else
        k_convert<float, float><<<g, kBlock, 0, ctx->stream>>>(n, (const float *)s, (float *)d, ctx->stop_word);
    PP_HIP(hipGetLastError());
}

// z = a x + b y, evaluated as (a*x) + (b*y) with both products rounded, the
// order of HostVector::add (BaseVector.h:279-301).
template <class T>
__global__ void k_axpby(uint32_t n, T a, const T *x, T b, const T *y, T *z) {
#pragma clang fp contract(off)
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T t1 = a * x[i];
        T t2 = b * y[i];
        z[i] = t1 + t2;
    }
}

void launch_axpby(pp_ctx *ctx, int dtype, uint32_t n, double a, const void *x, double b, const void *y,
                  void *z) {
    unsigned g = grid_for(ctx, n);
    if (dtype == PP_F64)
        k_axpby<double><<<g, kBlock, 0, ctx->stream>>>(n, a, (const double *)x, b, (const double *)y, (double *)z);
    else
        k_axpby<float><<<g, kBlock, 0, ctx->stream>>>(n, (float)a, (const float *)x, (float)b, (const float *)y,
                                                      (float *)z);
    PP_HIP(hipGetLastError());
}

// dot with the accumulator of the working precision inside a thread, FP64
// across threads/workgroups.
template <class T>
__global__ __launch_bounds__(256) void k_dot(uint32_t n, const T *x, const T *y, double *partials,
                                             const int *stop) {
    __shared__ double sh[4];
    if (stop && *stop) return;
    T acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += x[i] * y[i];
    publish_partial(block_sum((double)acc, sh), partials);
}

void launch_dot(pp_ctx *ctx, int dtype, uint32_t n, const void *x, const void *y, int slot) {
    unsigned g = grid_for(ctx, n, 4);
    double *part = ctx->red.partials + (size_t)slot * kMaxPartials;
    if (dtype == PP_F64)
        k_dot<double><<<g, kBlock, 0, ctx->stream>>>(n, (const double *)x, (const double *)y, part, ctx->stop_word);
    else
        k_dot<float><<<g, kBlock, 0, ctx->stream>>>(n, (const float *)x, (const float *)y, part, ctx->stop_word);
    launch_finish(ctx, slot, g, ctx->stop_word);
    PP_HIP(hipGetLastError());
}

__global__ void k_outer_on_pattern(uint32_t n, const uint32_t *rp, const uint32_t *ci, double scale,
                                   const double *lam, const double *x, double *g) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double li = scale * lam[i];
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) g[k] = li * x[ci[k]];
    }
}

void launch_outer_on_pattern(pp_ctx *ctx, const pp_csr *A, double scale, const double *lam, const double *x,
                             double *g) {
    k_outer_on_pattern<<<grid_for(ctx, A->nrow), kBlock, 0, ctx->stream>>>(A->nrow, A->rp, A->ci, scale, lam, x, g);
    PP_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Level-scheduled triangular sweep, one launch per level.  Rows of a level
// are stored contiguously (level order); each thread finishes one row by
// subtracting its dependencies in storage order with no fused multiply-add,
// which is the operation sequence of the reference's serial sweeps
// (IncompleteCholesky.cpp:250-295, IncompleteLU.cpp:269-297,
// IncompleteLDLT.cpp:189-233) -- bit-identical results in FP64.
// mode: 0 divide by diag, 1 unit, 2 multiply by diag, 3 divide first.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_sweep_level(const uint32_t *__restrict__ perm,
                                                     const uint32_t *__restrict__ rp,
                                                     const uint32_t *__restrict__ ci,
                                                     const T *__restrict__ val, const T *__restrict__ diag,
                                                     int mode, const T *b, T *x, uint32_t k0, uint32_t k1,
                                                     long long off, const int *stop, int mode2,
                                                     const T *__restrict__ diag2) {
#pragma clang fp contract(off)
    const uint32_t k = k0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= k1) return;
    if (stop && *stop) return;
    // perm == nullptr: the level's rows are contiguous, row = position + off
    const uint32_t i = perm ? perm[k] : (uint32_t)((long long)k + off);
    T s = b[i];
    if (mode == 3) s = s / diag[i];
    for (uint32_t e = rp[k]; e < rp[k + 1]; ++e) {
        T p = val[e] * x[ci[e]];
        s = s - p;
    }
    if (mode == 0) s = s / diag[i];
    else if (mode == 2) s = s * diag[i];
    // fused first level of the next sweep (rows without dependencies)
    if (mode2 == 0 || mode2 == 3) s = s / diag2[i];
    else if (mode2 == 2) s = s * diag2[i];
    x[i] = s;
}

// One wide level, LDS-staged like the SpMV: each workgroup takes a row block,
// loads its dependencies' values, columns and x gathers coalesced into LDS
// products, then one thread per row subtracts them in storage order.  Same
// operations in the same order as k_sweep_level (bit-identical results).
// Rows of one level never depend on each other, so x may alias b.
template <class T, bool NT, bool BATCH>
__global__ __launch_bounds__(256) void k_sweep_stream(const uint32_t *__restrict__ perm,
                                                      const uint32_t *__restrict__ rp,
                                                      const uint32_t *__restrict__ ci,
                                                      const T *__restrict__ val, const T *__restrict__ diag,
                                                      int mode, const T *b, T *x,
                                                      const BlockDesc *__restrict__ desc,
                                                      const uint16_t *__restrict__ ci16, uint32_t blk0,
                                                      uint32_t blk1, long long off, const int *stop, int mode2,
                                                      const T *__restrict__ diag2) {
#pragma clang fp contract(off)
    __shared__ T prod[kSpmvNnzBlock];
    if (stop && *stop) return;
    const XcdRange xr = xcd_range(blk0, blk1);
    for (uint32_t bi = xr.first; bi < xr.end; bi += xr.step) {
        const BlockDesc d0 = desc[bi], d1 = desc[bi + 1];
        const uint32_t r0 = d0.r0, r1 = d1.r0;
        const uint32_t e0 = d0.k0, cnt = d1.k0 - e0;
        const bool staged = cnt <= kSpmvNnzBlock;
        // row operands first (their latency overlaps the product stage)
        static_assert(kSpmvRowsBlock <= (uint32_t)kBlock, "one row per thread in phase 2");
        const bool mine = r0 + threadIdx.x < r1;
        uint32_t pi = 0, pa = 0, pe = 0;
        T pb = 0, pd = 0, pd2 = 0;
        if (mine) {
            const uint32_t kk = r0 + threadIdx.x;
            pi = perm ? perm[kk] : (uint32_t)((long long)kk + off);
            pa = rp[kk];
            pe = rp[kk + 1];
            pb = b[pi];
            if (mode != 1) pd = diag[pi];
            if (mode2 == 0 || mode2 == 2 || mode2 == 3) pd2 = diag2[pi];
        }
        if (staged) stage_products<T, NT, BATCH>(prod, val, ci, ci16, ci16 ? d0.blo : kNoBase, d0.bhi, e0, cnt, x);
        __syncthreads();
        if (mine) {  // row blocks hold at most kBlock rows: one row per thread
            T s = pb;
            if (mode == 3) s = s / pd;
            if (staged) {
                for (uint32_t e = pa; e < pe; ++e) s = s - prod[e - e0];
            } else {  // a single row longer than the staging buffer
                for (uint32_t e = pa; e < pe; ++e) {
                    T p = val[e] * x[ci[e]];
                    s = s - p;
                }
            }
            if (mode == 0) s = s / pd;
            else if (mode == 2) s = s * pd;
            if (mode2 == 0 || mode2 == 3) s = s / pd2;
            else if (mode2 == 2) s = s * pd2;
            x[pi] = s;
        }
        __syncthreads();
    }
}

// One contiguous level in sliced-column form (DevSweep::EllLevel): row
// row0 + t per thread, entry j of all rows contiguous (coalesced loads, no
// LDS stage).  The same operation sequence as k_sweep_stream (b, the
// pre-division, the products subtracted in storage order each rounded, the
// row's diagonal mode, the next sweep's first-level mode), so the results
// are the same bit for bit; padding entries subtract a zero.
template <class T, int W>
__global__ __launch_bounds__(256) void k_sweep_ell(uint32_t nrows, uint32_t width, const int16_t *__restrict__ off,
                                                   const T *__restrict__ val, long long row0, long long cbase,
                                                   const T *__restrict__ diag, int mode, const T *b, T *x,
                                                   const int *stop, int mode2, const T *__restrict__ diag2) {
#pragma clang fp contract(off)
    if (stop && *stop) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    const long long i = row0 + t;
    constexpr int MAXW = EllWidth<W>::kMax;
    int16_t o[MAXW];
    T v[MAXW], xv[MAXW];
#pragma unroll
    for (int j = 0; j < MAXW; ++j)
        if (EllWidth<W>::has(j, width)) {
            const size_t e = (size_t)j * nrows + t;
            o[j] = off[e];
            v[j] = val[e];
        }
    const T pb = b[i];
    const T pd = mode != 1 ? diag[i] : (T)0;
    const T pd2 = (mode2 == 0 || mode2 == 2 || mode2 == 3) ? diag2[i] : (T)0;
#pragma unroll
    for (int j = 0; j < MAXW; ++j)
        if (EllWidth<W>::has(j, width)) xv[j] = x[i + cbase + o[j]];
    T s = pb;
    if (mode == 3) s = s / pd;
#pragma unroll
    for (int j = 0; j < MAXW; ++j)
        if (EllWidth<W>::has(j, width)) {
            const T p = v[j] * xv[j];
            s = s - p;
        }
    if (mode == 0) s = s / pd;
    else if (mode == 2) s = s * pd;
    if (mode2 == 0 || mode2 == 3) s = s / pd2;
    else if (mode2 == 2) s = s * pd2;
    x[i] = s;
}

// Levels [l0, l1) of one sweep; mode2 >= 0 appends the next sweep's
// dependency-free first level (diag2) to every row of these levels.
template <class T>
void launch_levels(pp_ctx *ctx, const DevSweep<T> &s, int mode, const T *b, T *x, uint32_t l0, uint32_t l1,
                   int mode2, const T *diag2) {
    // early exit only inside a solve's iteration body (StopScope): a sweep
    // launched elsewhere (a preconditioner setup, a plain apply) always runs,
    // whatever stop word an earlier solve left behind
    const int *stop = ctx->stop_word;
    const bool ell = s.variant == kVarEll;
    const int var = ell ? s.ell_fallback : s.variant;  // the staged kernel's loads on the other levels
    for (uint32_t l = l0; l < l1; ++l) {
        uint32_t k0 = s.level_ptr[l], k1 = s.level_ptr[l + 1];
        const bool contiguous = l < s.level_off.size() && s.level_off[l] != DevSweep<T>::kNoOffset;
        if (ell && l < s.ell.size() && s.ell[l].ok) {
            const auto &E = s.ell[l];
            const unsigned g = (E.nrows + kBlock - 1) / kBlock;
#define PP_SWEEP_ELL(W)                                                                                     \
    k_sweep_ell<T, W><<<g, kBlock, 0, ctx->stream>>>(E.nrows, E.width, E.off, E.val, E.row0, E.cbase, s.diag, \
                                                     mode, b, x, stop, mode2, diag2)
            PP_ELL_WIDTHS(E.width, PP_SWEEP_ELL)
#undef PP_SWEEP_ELL
            continue;
        }
        if (s.blocks && 2 * l + 1 < s.level_blk.size() && s.level_blk[2 * l + 1] > s.level_blk[2 * l]) {
            const uint32_t b0 = s.level_blk[2 * l], b1 = s.level_blk[2 * l + 1];
            const unsigned g = xcd_grid(b1 - b0, 8192u);
            const uint16_t *c16 = (var & kVarIdx16) ? s.ci16 : nullptr;
            const long long off = contiguous ? s.level_off[l] : 0;
            const uint32_t *perm = contiguous ? nullptr : s.perm;
#define PP_SWEEP(NTV, BV)                                                                                         \
    k_sweep_stream<T, NTV, BV><<<g, kBlock, 0, ctx->stream>>>(perm, s.rp, s.ci, s.val, s.diag, mode, b, x, s.desc, \
                                                             c16, b0, b1, off, stop, mode2, diag2)
            switch (var & (kVarNT | kVarBatch)) {
                case 0: PP_SWEEP(false, false); break;
                case kVarNT: PP_SWEEP(true, false); break;
                case kVarBatch: PP_SWEEP(false, true); break;
                default: PP_SWEEP(true, true); break;
            }
#undef PP_SWEEP
            continue;
        }
        unsigned g = (k1 - k0 + kBlock - 1) / kBlock;
        k_sweep_level<T><<<g, kBlock, 0, ctx->stream>>>(contiguous ? nullptr : s.perm, s.rp, s.ci, s.val, s.diag,
                                                        mode, b, x, k0, k1, contiguous ? s.level_off[l] : 0, stop,
                                                        mode2, diag2);
    }
    PP_HIP(hipGetLastError());
}

template <class T>
void launch_sweep(pp_ctx *ctx, const DevSweep<T> &s, int mode, const T *b, T *x) {
    launch_levels(ctx, s, mode, b, x, 0, s.nlev, -1, (const T *)nullptr);
}

// y = U^{-1} L^{-1} x: the forward sweep reads its right-hand side from x
// and writes y (x may alias y), the backward sweep runs in place on y.
// fuse: the last level of `lo` holds exactly the rows of the first level of
// `up` (which have no dependencies), so one kernel finishes the forward
// sweep and starts the backward one.
// first_level > 0: the forward sweep's levels below it are already in y.
template <class T>
void launch_sweep_pair(pp_ctx *ctx, const DevSweep<T> &lo, int mlo, const DevSweep<T> &up, int mup, bool fuse,
                       const T *x, T *y, uint32_t first_level) {
    if (!fuse || lo.nlev == 0 || up.nlev == 0 || first_level >= lo.nlev) {
        launch_levels(ctx, lo, mlo, x, y, std::min(first_level, lo.nlev), lo.nlev, -1, (const T *)nullptr);
        launch_sweep(ctx, up, mup, y, y);
        return;
    }
    launch_levels(ctx, lo, mlo, x, y, first_level, lo.nlev - 1, -1, (const T *)nullptr);
    launch_levels(ctx, lo, mlo, x, y, lo.nlev - 1, lo.nlev, mup, up.diag);
    launch_levels(ctx, up, mup, y, y, 1, up.nlev, -1, (const T *)nullptr);
}
template void launch_sweep_pair<double>(pp_ctx *, const DevSweep<double> &, int, const DevSweep<double> &, int,
                                        bool, const double *, double *, uint32_t);
template void launch_sweep_pair<float>(pp_ctx *, const DevSweep<float> &, int, const DevSweep<float> &, int, bool,
                                       const float *, float *, uint32_t);

// Both sweeps of an apply, one workgroup per chunk (PP_TRSV_BLOCK).  The
// chunk's part of the vector lives in LDS; level after level the threads take
// the level's rows and a workgroup barrier separates levels.  Per row: the
// operations of k_sweep_level in the same order (bit-identical results).
// A thread's share of the next level, loaded while the current level
// computes: its first entry of the level (for the products) and its first
// row (for the sums).  Loaded into registers that the caller alternates, so
// the barrier between levels never waits for these loads.
template <class T>
struct ChunkAhead {
    uint32_t c, i, e0, e1;
    T v, d;
    bool eok, rok;
};

template <class T>
__device__ inline void chunk_ahead(ChunkAhead<T> &q, uint32_t L, uint32_t Lend, uint32_t E, uint32_t Eend,
                                   const DevChunkSweep<T> &S) {
    const uint32_t e = E + threadIdx.x, k = L + threadIdx.x;
    q.eok = e < Eend;
    q.rok = k < Lend;
    // unconditional loads (index 0 stands in when out of range; every array
    // holds at least one element): without branches around them the
    // compiler can wait for exactly these loads where they are used
    const uint32_t es = q.eok ? e : 0, ks = q.rok ? k : 0;
    q.c = S.ci[es];
    q.v = S.val[es];
    q.i = S.row[ks];
    q.e0 = S.rp[ks];
    q.e1 = S.rp[ks + 1];
    q.d = S.pdiag ? S.pdiag[ks] : T(1);
}

// s - p[a] - p[a+1] - ... - p[b-1] (p in LDS), each subtraction rounded,
// in index order.  The reads go out eight at a time before the first of
// their subtractions, so a row waits on LDS once per eight entries instead
// of once per entry (the row sums are the chunked sweep's critical path: a
// few rows per level, a dozen or more entries per row).  Reads past b
// repeat p[b-1] and subtract +0 instead, which leaves s unchanged (x - (+0)
// is x for every x, -0 and NaN included).
template <class T>
__device__ inline T sub_run(T s, const T *p, uint32_t a, uint32_t b) {
#pragma clang fp contract(off)
    for (uint32_t e = a; e < b; e += 8) {
        T q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = p[min(e + u, b - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) s = s - (e + u < b ? q[u] : T(0));
    }
    return s;
}

template <class T>
__device__ inline void chunk_row_finish(T s, uint32_t i, T d, int mode, T *xs) {
#pragma clang fp contract(off)
    if (mode == 0) s = s / d;
    else if (mode == 2) s = s * d;
    xs[i] = s;
}

// One level of a chunked sweep (entries [E0, E1) of rows [L0, L1)), then the
// level boundaries advance.  Staged (the level's entries fit the product
// buffer): every thread forms products of the level's entries into LDS,
// barrier, then each row's thread subtracts its products in storage order.
// Otherwise each row's thread forms and subtracts its own products.  Per
// row the operations of k_sweep_level in the same order (bit-identical).
template <class T>
__device__ inline void chunk_level(ChunkAhead<T> &cur, ChunkAhead<T> &nxt, const uint32_t *__restrict__ lp,
                                   const uint32_t *__restrict__ le, uint32_t nlev, uint32_t l, uint32_t &L0,
                                   uint32_t &L1, uint32_t &L2, uint32_t &E0, uint32_t &E1, uint32_t &E2,
                                   const DevChunkSweep<T> &S, int mode, T *xs, T *prod, uint32_t cap) {
#pragma clang fp contract(off)
    const uint32_t L3 = l + 3 <= nlev ? lp[l + 3] : L2, E3 = l + 3 <= nlev ? le[l + 3] : E2;
    nxt.eok = nxt.rok = false;
    if (l + 1 < nlev) chunk_ahead<T>(nxt, L1, L2, E1, E2, S);
    if (E1 - E0 <= cap) {
        if (cur.eok) {
            T p = cur.v * xs[cur.c];
            prod[threadIdx.x] = p;
        }
        for (uint32_t e = E0 + threadIdx.x + blockDim.x; e < E1; e += blockDim.x) {
            T p = S.val[e] * xs[S.ci[e]];
            prod[e - E0] = p;
        }
        __syncthreads();
        if (cur.rok) {
            T s = xs[cur.i];
            if (mode == 3) s = s / cur.d;
            s = sub_run(s, prod, cur.e0 - E0, cur.e1 - E0);
            chunk_row_finish(s, cur.i, cur.d, mode, xs);
        }
        for (uint32_t k = L0 + threadIdx.x + blockDim.x; k < L1; k += blockDim.x) {
            const uint32_t i = S.row[k];
            const T d = S.pdiag ? S.pdiag[k] : T(1);
            T s = xs[i];
            if (mode == 3) s = s / d;
            s = sub_run(s, prod, S.rp[k] - E0, S.rp[k + 1] - E0);
            chunk_row_finish(s, i, d, mode, xs);
        }
    } else {
        for (uint32_t k = L0 + threadIdx.x; k < L1; k += blockDim.x) {
            const uint32_t i = S.row[k];
            const T d = S.pdiag ? S.pdiag[k] : T(1);
            T s = xs[i];
            if (mode == 3) s = s / d;
            for (uint32_t e = S.rp[k], e1 = S.rp[k + 1]; e < e1; ++e) {
                T p = S.val[e] * xs[S.ci[e]];
                s = s - p;
            }
            chunk_row_finish(s, i, d, mode, xs);
        }
    }
    __syncthreads();
    L0 = L1;
    L1 = L2;
    L2 = L3;
    E0 = E1;
    E1 = E2;
    E2 = E3;
}

// Levels of one sweep over a chunk, in place on xs (the chunk's right-hand
// side on entry).
template <class T>
__device__ inline void chunk_sweep(const uint32_t *__restrict__ lp, const uint32_t *__restrict__ le, uint32_t nlev,
                                   const DevChunkSweep<T> &S, int mode, T *xs, T *prod, uint32_t cap) {
    if (nlev == 0) return;
    uint32_t L0 = lp[0], L1 = lp[1], L2 = nlev > 1 ? lp[2] : L1;
    uint32_t E0 = le[0], E1 = le[1], E2 = nlev > 1 ? le[2] : E1;
    ChunkAhead<T> ra, rb;
    chunk_ahead<T>(ra, L0, L1, E0, E1, S);
    for (uint32_t l = 0; l < nlev; l += 2) {
        chunk_level<T>(ra, rb, lp, le, nlev, l, L0, L1, L2, E0, E1, E2, S, mode, xs, prod, cap);
        if (l + 1 < nlev)
            chunk_level<T>(rb, ra, lp, le, nlev, l + 1, L0, L1, L2, E0, E1, E2, S, mode, xs, prod, cap);
    }
}

// Both sweeps of an apply, one workgroup per chunk (PP_TRSV_BLOCK); LDS holds
// the chunk's part of the vector (max_rows entries) and the product buffer
// (cap entries).
template <class T>
__global__ __launch_bounds__(1024) void k_sweep_chunks(const uint32_t *__restrict__ start, DevChunkSweep<T> lo,
                                                       int lo_mode, DevChunkSweep<T> up, int up_mode, const T *x,
                                                       T *y, uint32_t max_rows, uint32_t lvl_words, uint32_t cap,
                                                       const int *stop) {
    extern __shared__ unsigned char lds_raw[];
    T *xs = reinterpret_cast<T *>(lds_raw);
    // the chunk's level boundaries of both sweeps (lp, le) staged in LDS, so
    // a level reads them without waiting on global loads; then the products
    uint32_t *lvl = reinterpret_cast<uint32_t *>(xs + max_rows);
    T *prod = reinterpret_cast<T *>(lvl + lvl_words);
    if (stop && *stop) return;
    const uint32_t c = blockIdx.x, r0 = start[c], r1 = start[c + 1];
    const uint32_t lo0 = lo.lev_off[c], lo1 = lo.lev_off[c + 1], up0 = up.lev_off[c], up1 = up.lev_off[c + 1];
    const uint32_t nlo = lo1 - lo0, nup = up1 - up0;  // boundaries of each sweep (levels + 1)
    uint32_t *llp = lvl, *lle = lvl + nlo, *ulp = lvl + 2 * nlo, *ule = lvl + 2 * nlo + nup;
    for (uint32_t j = threadIdx.x; j < r1 - r0; j += blockDim.x) xs[j] = x[r0 + j];
    for (uint32_t j = threadIdx.x; j < nlo; j += blockDim.x) {
        llp[j] = lo.lp[lo0 + j];
        lle[j] = lo.le[lo0 + j];
    }
    for (uint32_t j = threadIdx.x; j < nup; j += blockDim.x) {
        ulp[j] = up.lp[up0 + j];
        ule[j] = up.le[up0 + j];
    }
    __syncthreads();
    chunk_sweep<T>(llp, lle, nlo - 1, lo, lo_mode, xs, prod, cap);
    chunk_sweep<T>(ulp, ule, nup - 1, up, up_mode, xs, prod, cap);
    for (uint32_t j = threadIdx.x; j < r1 - r0; j += blockDim.x) y[r0 + j] = xs[j];
}

template <class T>
void launch_sweep_pair_chunks(pp_ctx *ctx, const ChunkSet &C, const DevChunkSweep<T> &lo, int mlo,
                              const DevChunkSweep<T> &up, int mup, const T *x, T *y) {
    const uint32_t lvl_words = 2 * (lo.max_levels + 1) + 2 * (up.max_levels + 1);
    const size_t xbytes = (size_t)C.max_rows * sizeof(T) + (size_t)lvl_words * 4;
    PP_CHECK(C.nchunk && xbytes <= ctx->max_lds && lo.row && up.row, PP_ERR_INVALID,
             "chunked sweep: storage missing or chunk exceeds LDS");
    // the product buffer takes what LDS has left, up to the widest level
    const uint32_t want = std::max(lo.max_level_nnz, up.max_level_nnz);
    const uint32_t cap = (uint32_t)std::min<size_t>(want, (ctx->max_lds - xbytes) / sizeof(T));
    const size_t lds = xbytes + (size_t)cap * sizeof(T);
    if (lds > 65536)
        PP_HIP(hipFuncSetAttribute((const void *)k_sweep_chunks<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    const unsigned threads = C.max_rows >= 1024 ? 1024u : std::max(64u, (C.max_rows + 63) / 64 * 64);
    k_sweep_chunks<T><<<C.nchunk, threads, lds, ctx->stream>>>(C.dstart, lo, mlo, up, mup, x, y, C.max_rows,
                                                              lvl_words, cap, ctx->stop_word);
    PP_HIP(hipGetLastError());
}
template void launch_sweep_pair_chunks<double>(pp_ctx *, const ChunkSet &, const DevChunkSweep<double> &, int,
                                               const DevChunkSweep<double> &, int, const double *, double *);
template void launch_sweep_pair_chunks<float>(pp_ctx *, const ChunkSet &, const DevChunkSweep<float> &, int,
                                              const DevChunkSweep<float> &, int, const float *, float *);

template <class T>
__global__ void k_level0(uint32_t lo, uint32_t hi, int mode, const T *__restrict__ diag, const T *__restrict__ x,
                         T *__restrict__ y, const int *stop) {
#pragma clang fp contract(off)
    if (stop && *stop) return;
    for (uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += gridDim.x * blockDim.x) {
        T s = x[i];
        if (mode == 0 || mode == 3) s = s / diag[i];
        else if (mode == 2) s = s * diag[i];
        y[i] = s;
    }
}

template <class T>
void launch_level0(pp_ctx *ctx, uint32_t lo, uint32_t hi, int mode, const T *diag, const T *x, T *y) {
    if (hi <= lo) return;
    k_level0<T><<<grid_for(ctx, hi - lo, 4), kBlock, 0, ctx->stream>>>(lo, hi, mode, diag, x, y, ctx->stop_word);
    PP_HIP(hipGetLastError());
}
template void launch_level0<double>(pp_ctx *, uint32_t, uint32_t, int, const double *, const double *, double *);
template void launch_level0<float>(pp_ctx *, uint32_t, uint32_t, int, const float *, const float *, float *);
template void launch_sweep<double>(pp_ctx *, const DevSweep<double> &, int, const double *, double *);
template void launch_sweep<float>(pp_ctx *, const DevSweep<float> &, int, const float *, float *);

// ---------------------------------------------------------------------------
// Sync-free triangular sweep: ONE launch per sweep.  Rows keep the level
// order, so every dependency of a row lives in a lower workgroup index, and
// workgroups are dispatched in index order on each XCD -- the lowest
// unfinished row always has its inputs.  The output vector is pre-filled with
// a signalling-NaN pattern that no arithmetic produces; a row polls each
// dependency's value with an agent-scope atomic load until it is no longer
// that pattern, so one 64-bit (32-bit) word carries both the value and its
// ready flag and no release/acquire fences are needed.  Dependencies are
// consumed in storage order (same arithmetic as the level kernel, so the FP64
// result is bit-identical).  A wall-clock deadline bounds every wait: on
// expiry the sweep stops and raises an error word instead of hanging.
// ---------------------------------------------------------------------------
template <class T>
struct SentinelBits;
template <>
struct SentinelBits<double> {
    using U = unsigned long long;
    static constexpr U value = 0x7FF4B5A5C3D2E1F0ull;  // signalling NaN
};
template <>
struct SentinelBits<float> {
    using U = unsigned int;
    static constexpr U value = 0x7FA5C3E1u;  // signalling NaN
};

// Skipped, like the sweep that follows it, once the stop word is set: an
// iteration issued past convergence leaves its output as it was instead of
// a vector of sentinels (which a later product with a zero coefficient would
// turn into NaN).
template <class T>
__global__ void k_fill_sentinel(uint32_t n, T *y, const int *ctl) {
    using U = typename SentinelBits<T>::U;
    if (ctl[C_STOP]) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        reinterpret_cast<U *>(y)[i] = SentinelBits<T>::value;
}

template <class T>
__global__ __launch_bounds__(256) void k_sweep_syncfree(const uint32_t *__restrict__ perm,
                                                        const uint32_t *__restrict__ rp,
                                                        const uint32_t *__restrict__ ci,
                                                        const T *__restrict__ val, const T *__restrict__ diag,
                                                        int mode, const T *__restrict__ b, T *y, uint32_t n,
                                                        long long deadline, int *ctl) {
#pragma clang fp contract(off)
    using U = typename SentinelBits<T>::U;
    constexpr U kSent = SentinelBits<T>::value;
    if (ctl[C_STOP]) return;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    bool done = k >= n;
    uint32_t i = 0, e = 0, e1 = 0;
    T s = 0;
    if (!done) {
        i = perm[k];
        s = b[i];
        if (mode == 3) s = s / diag[i];
        e = rp[k];
        e1 = rp[k + 1];
    }
    U *ybits = reinterpret_cast<U *>(y);
    const long long t0 = wall_clock64();
    for (unsigned iter = 0;; ++iter) {
        if (!done) {
            while (e < e1) {
                const U bits = __hip_atomic_load(&ybits[ci[e]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (bits == kSent) break;
                const T p = val[e] * __builtin_bit_cast(T, bits);
                s = s - p;
                ++e;
            }
            if (e == e1) {
                if (mode == 0) s = s / diag[i];
                else if (mode == 2) s = s * diag[i];
                __hip_atomic_store(&ybits[i], __builtin_bit_cast(U, s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                done = true;
            }
        }
        if (__all(done)) break;
        if (wall_clock64() - t0 > deadline) {
            if (!done) atomicOr(&ctl[C_SPIN], 1);
            break;
        }
        // another wave already gave up: its dependents would wait in vain
        if ((iter & 63) == 63 && __hip_atomic_load(&ctl[C_SPIN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            break;
        __builtin_amdgcn_s_sleep(1);
    }
}

template <class T>
void launch_sweep_syncfree(pp_ctx *ctx, const DevSweep<T> &s, int mode, const T *b, T *y) {
    if (s.n == 0) return;
    const unsigned g = (s.n + kBlock - 1) / kBlock;
    k_fill_sentinel<T><<<grid_for(ctx, s.n), kBlock, 0, ctx->stream>>>(s.n, y, ctx->ctl);
    k_sweep_syncfree<T><<<g, kBlock, 0, ctx->stream>>>(s.perm, s.rp, s.ci, s.val, s.diag, mode, b, y, s.n,
                                                        ctx->spin_deadline, ctx->ctl);
    PP_HIP(hipGetLastError());
}
template void launch_sweep_syncfree<double>(pp_ctx *, const DevSweep<double> &, int, const double *, double *);
template void launch_sweep_syncfree<float>(pp_ctx *, const DevSweep<float> &, int, const float *, float *);

// ---------------------------------------------------------------------------
// PCG steps (FlexibleCG, source/Solver/CG.cpp:49-79):
//   d = w - beta d with beta = (w, A d_old) / (d_old, A d_old)   [fused (d, r)]
//   v = A d                                                       [fused (d, v)]
//   x += alpha d, r -= alpha v with alpha = (d, r) / (d, v)       [fused (r, r) + check]
// Every kernel returns at once when the stop word is set, so iterations
// launched past convergence cost only their launch.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_pcg_d(uint32_t n, int first, const T *__restrict__ w,
                                               T *__restrict__ d, const T *__restrict__ r,
                                               const double *scal, const int *ctl, double *partials) {
#pragma clang fp contract(off)
    __shared__ double sh[4];
    if (ctl[C_STOP]) return;
    const T beta = first ? (T)0 : (T)(scal[S_WV] / scal[S_DV]);
    T acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T dn;
        if (first) dn = w[i];
        else {
            T t1 = w[i];
            T t2 = (-beta) * d[i];
            dn = t1 + t2;
        }
        d[i] = dn;
        acc += dn * r[i];
    }
    publish_partial(block_sum((double)acc, sh), partials);
}

template <class T>
__global__ __launch_bounds__(256) void k_pcg_wv(uint32_t n, const T *__restrict__ w, const T *__restrict__ v,
                                                const int *ctl, double *partials) {
    __shared__ double sh[4];
    if (ctl[C_STOP]) return;
    T acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += w[i] * v[i];
    publish_partial(block_sum((double)acc, sh), partials);
}

// The relative residual norm of a check: scal[S_TMP0] and its bits in the
// control words (published with them to the host's ring).
__device__ inline void record_norm(double *scal, int *ctl, double norm) {
    scal[S_TMP0] = norm;
    const unsigned long long b = (unsigned long long)__double_as_longlong(norm);
    ctl[C_NORM_LO] = (int)(unsigned)(b & 0xffffffffull);
    ctl[C_NORM_HI] = (int)(unsigned)(b >> 32);
    ctl[C_NORM_SET] = 1;
}

// Convergence check of CG.cpp:57-63 (run by k_finish on (r, r)): the check
// index c counts the reference loop index i.
__device__ void pcg_check(double rr, const double *scal, int *ctl, double *hist, uint32_t hist_cap) {
    const double tol = scal[S_TOL];
    const uint32_t maxit = (uint32_t)ctl[C_MAXIT];
    if (ctl[C_LAST]) {  // the body after the final check has run
        ctl[C_STOP] = 1;
        ctl[C_M] = (int)maxit;
        return;
    }
    const int c = ctl[C_CHECK];
    const double norm = sqrt(rr) / sqrt(scal[S_BB]);
    if (hist && (uint32_t)c - 1u < hist_cap) hist[c - 1] = norm;
    record_norm(const_cast<double *>(scal), ctl, norm);
    if (!(norm == norm) || isinf(norm)) {
        ctl[C_STATUS] = PP_ERR_NOT_FINITE;
        ctl[C_STOP] = 1;
        ctl[C_M] = c;
        return;
    }
    if (norm < tol) {
        ctl[C_CONV] = 1;
        ctl[C_STOP] = 1;
        ctl[C_M] = c;
        return;
    }
    if ((uint32_t)c >= maxit - 1) ctl[C_LAST] = 1;
    ctl[C_CHECK] = c + 1;
}

template <class T>
__global__ __launch_bounds__(256) void k_pcg_xr(uint32_t n, T *__restrict__ x, T *__restrict__ r,
                                                const T *__restrict__ d, const T *__restrict__ v,
                                                const double *scal, int *ctl, double *partials) {
#pragma clang fp contract(off)
    __shared__ double sh[4];
    if (ctl[C_STOP]) return;
    const double den = scal[S_DV];
    if (den == 0.0 || !(den == den)) {  // (d, A d) = 0: breakdown; NaN: non-finite data
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctl[C_STATUS] = den == 0.0 ? PP_ERR_BREAKDOWN : PP_ERR_NOT_FINITE;
            ctl[C_STOP] = 1;
            ctl[C_M] = ctl[C_CHECK];
        }
        return;
    }
    const T alpha = (T)(scal[S_DR] / den);
    T acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T t1 = alpha * d[i];
        T xn = x[i] + t1;
        x[i] = xn;
        T t2 = (-alpha) * v[i];
        T rn = r[i] + t2;
        r[i] = rn;
        acc += rn * rn;
    }
    publish_partial(block_sum((double)acc, sh), partials);
}

template <class T>
void pcg_d_update(pp_ctx *ctx, uint32_t n, bool first, const T *w, T *d, const T *r) {
    unsigned g = grid_for(ctx, n, 4);
    k_pcg_d<T><<<g, kBlock, 0, ctx->stream>>>(n, first ? 1 : 0, w, d, r, ctx->scal, ctx->ctl,
                                              ctx->red.partials + (size_t)S_DR * kMaxPartials);
    launch_finish(ctx, S_DR, g, ctx->ctl + C_STOP);
    PP_HIP(hipGetLastError());
}

template <class T>
void pcg_wv(pp_ctx *ctx, uint32_t n, const T *w, const T *v) {
    unsigned g = grid_for(ctx, n, 4);
    k_pcg_wv<T><<<g, kBlock, 0, ctx->stream>>>(n, w, v, ctx->ctl, ctx->red.partials + (size_t)S_WV * kMaxPartials);
    launch_finish(ctx, S_WV, g, ctx->ctl + C_STOP);
    PP_HIP(hipGetLastError());
}

template <class T>
void pcg_xr_update(pp_ctx *ctx, uint32_t n, T *x, T *r, const T *d, const T *v, int hist_on, bool publish) {
    unsigned g = grid_for(ctx, n, 4);
    k_pcg_xr<T><<<g, kBlock, 0, ctx->stream>>>(n, x, r, d, v, ctx->scal, ctx->ctl,
                                               ctx->red.partials + (size_t)S_RR * kMaxPartials);
    launch_finish(ctx, S_RR, g, ctx->ctl + C_STOP, FIN_PCG_CHECK, hist_on ? ctx->hist : nullptr,
                  hist_on ? ctx->hist_cap : 0, publish ? ctx->ring_dev : nullptr);
    PP_HIP(hipGetLastError());
}

template void pcg_d_update<double>(pp_ctx *, uint32_t, bool, const double *, double *, const double *);
template void pcg_d_update<float>(pp_ctx *, uint32_t, bool, const float *, float *, const float *);
template void pcg_wv<double>(pp_ctx *, uint32_t, const double *, const double *);
template void pcg_wv<float>(pp_ctx *, uint32_t, const float *, const float *);
template void pcg_xr_update<double>(pp_ctx *, uint32_t, double *, double *, const double *, const double *, int,
                                    bool);
template void pcg_xr_update<float>(pp_ctx *, uint32_t, float *, float *, const float *, const float *, int, bool);

// ---------------------------------------------------------------------------
// Single-reduction preconditioned CG (Chronopoulos & Gear, J. Comput. Appl.
// Math. 25 (1989)): with u = M^{-1} r and w = A u, the three inner products
// gamma = (r, u), delta = (w, u), rho = (r, r) come out of the SpMV in one
// reduction; then
//   beta = gamma / gamma_old,  alpha = gamma / (delta - beta gamma / alpha_old)
//   p = u + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s
// (first iteration: beta = 0, alpha = gamma / delta).  Same Krylov iterates
// as CG in exact arithmetic; three kernels per iteration besides the sweeps.
// The convergence test is on rho = (r_i, r_i) before the update.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_cgsr_finish(const double *__restrict__ partials, unsigned np,
                                                      double *scal, int *ctl, double *hist, uint32_t hist_cap,
                                                      int *ring) {
    __shared__ double sh[3 * (kFinishThreads / 64)];
    if (ctl[C_STOP]) {
        if (ring && threadIdx.x == 0) publish_ctl(ctl, ring);
        return;
    }
    double t[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) t[q] = sum_partials(partials + (size_t)q * kMaxPartials, np);
    double v[3] = {t[0], t[1], t[2]};
    block_sum3(v[0], v[1], v[2], sh);
    if (threadIdx.x == 0) {
        const double delta = v[0], gamma = v[1], rho = v[2];
        scal[S_CG_DELTA] = delta;
        scal[S_CG_GAMMA] = gamma;
        scal[S_CG_RHO] = rho;
        pcg_check(rho, scal, ctl, hist, hist_cap);
        if (!ctl[C_STOP]) {
            const bool first = scal[S_CG_FIRST] != 0.0;
            const double beta = first ? 0.0 : gamma / scal[S_CG_GOLD];
            const double den = first ? delta : delta - beta * gamma / scal[S_CG_ALPHA];
            if (den == 0.0 || !(den == den)) {
                ctl[C_STATUS] = den == 0.0 ? PP_ERR_BREAKDOWN : PP_ERR_NOT_FINITE;
                ctl[C_STOP] = 1;
                ctl[C_M] = ctl[C_CHECK];
            } else {
                scal[S_CG_ALPHA] = gamma / den;
                scal[S_CG_BETA] = beta;
                scal[S_CG_GOLD] = gamma;
                scal[S_CG_FIRST] = 0.0;
            }
        }
        if (ring) publish_ctl(ctl, ring);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_cgsr_update(uint32_t n, T *__restrict__ u, const T *__restrict__ w,
                                                     T *__restrict__ p, T *__restrict__ s, T *__restrict__ x,
                                                     T *__restrict__ r, const double *scal, const int *ctl,
                                                     uint32_t l0, uint32_t l1, int mode0,
                                                     const T *__restrict__ diag0) {
#pragma clang fp contract(off)
    if (ctl[C_STOP]) return;
    const T alpha = (T)scal[S_CG_ALPHA], beta = (T)scal[S_CG_BETA];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T pn = u[i] + beta * p[i];
        const T sn = w[i] + beta * s[i];
        p[i] = pn;
        s[i] = sn;
        x[i] = x[i] + alpha * pn;
        const T rn = r[i] - alpha * sn;
        r[i] = rn;
        if (i >= l0 && i < l1) {  // first forward level of the next apply (no dependencies)
            T q = rn;
            if (mode0 == 0 || mode0 == 3) q = q / diag0[i];
            else if (mode0 == 2) q = q * diag0[i];
            u[i] = q;
        }
    }
}

unsigned spmv_grid(const pp_csr *A) { return A->nblocks ? xcd_grid(A->nblocks, kMaxPartials) : 0; }

void cgsr_finish(pp_ctx *ctx, unsigned np, int hist_on, bool publish) {
    k_cgsr_finish<<<1, kFinishThreads, 0, ctx->stream>>>(ctx->red.partials + (size_t)S_CG_DELTA * kMaxPartials, np,
                                                         ctx->scal,
                                                 ctx->ctl, hist_on ? ctx->hist : nullptr,
                                                 hist_on ? ctx->hist_cap : 0, publish ? ctx->ring_dev : nullptr);
    PP_HIP(hipGetLastError());
}

template <class T>
void cgsr_update(pp_ctx *ctx, uint32_t n, T *u, const T *w, T *p, T *s, T *x, T *r, uint32_t lvl0_lo,
                 uint32_t lvl0_hi, int mode0, const T *diag0) {
    k_cgsr_update<T><<<grid_for(ctx, n, 4), kBlock, 0, ctx->stream>>>(n, u, w, p, s, x, r, ctx->scal, ctx->ctl,
                                                                     lvl0_lo, lvl0_hi, mode0, diag0);
    PP_HIP(hipGetLastError());
}
template void cgsr_update<double>(pp_ctx *, uint32_t, double *, const double *, double *, double *, double *,
                                  double *, uint32_t, uint32_t, int, const double *);
template void cgsr_update<float>(pp_ctx *, uint32_t, float *, const float *, float *, float *, float *, float *,
                                 uint32_t, uint32_t, int, const float *);

// Update of the single-reduction CG on the reduced (Schur-complement) system,
// where the preconditioned residual is the residual itself:
// p = r + beta p, s = w + beta s, x += alpha p, r -= alpha s.
template <class T>
__global__ __launch_bounds__(256) void k_red_update(uint32_t n, const T *__restrict__ w, T *__restrict__ p,
                                                    T *__restrict__ s, T *__restrict__ x, T *__restrict__ r,
                                                    const double *scal, const int *ctl) {
#pragma clang fp contract(off)
    if (ctl[C_STOP]) return;
    const T alpha = (T)scal[S_CG_ALPHA], beta = (T)scal[S_CG_BETA];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T ri = r[i];
        const T pn = ri + beta * p[i];
        const T sn = w[i] + beta * s[i];
        p[i] = pn;
        s[i] = sn;
        x[i] = x[i] + alpha * pn;
        r[i] = ri - alpha * sn;
    }
}

template <class T>
void red_update(pp_ctx *ctx, uint32_t n, const T *w, T *p, T *s, T *x, T *r) {
    k_red_update<T><<<grid_for(ctx, n, 4), kBlock, 0, ctx->stream>>>(n, w, p, s, x, r, ctx->scal, ctx->ctl);
    PP_HIP(hipGetLastError());
}
template void red_update<double>(pp_ctx *, uint32_t, const double *, double *, double *, double *, double *);
template void red_update<float>(pp_ctx *, uint32_t, const float *, float *, float *, float *, float *);

// The convergence test and step lengths of one reduced-CG iteration from
// the three inner products' partials (see red_step in pp_internal.h); every
// workgroup of kFinishThreads threads computes the same, workgroup 0 also
// writes the iteration state.  The test mirrors pcg_check and
// k_cgsr_finish: a pending last check stops with m = maxit; a non-finite
// norm or norm < tol stops with m = the check index; a zero or non-finite
// denominator is a breakdown.  t: this thread's sums of the three partial
// arrays (sum_partials); prior_stop: ctl[C_STOP] as read at the kernel's
// start -- an earlier check stopped, so this one only republishes the
// control words.  Returns false when the iteration stops.
__device__ inline bool red_check(const double t[3], int prior_stop, double *scal, int *ctl, int *ring, int parity,
                                 double *sh, double *bc, int *go, double &alpha_out, double &beta_out) {
#pragma clang fp contract(off)
    double v[3] = {t[0], t[1], t[2]};
    block_sum3(v[0], v[1], v[2], sh);
    if (threadIdx.x == 0 && prior_stop) {
        *go = 0;
        if (ring && blockIdx.x == 0) publish_ctl(ctl, ring);
    } else if (threadIdx.x == 0) {
        const double *st = scal + (parity ? S_RED1 : S_RED0);
        double *nx = scal + (parity ? S_RED0 : S_RED1);
        const double delta = v[0], gamma = v[1], rho = v[2];
        const double gold = st[RS_GOLD], aold = st[RS_AOLD];
        const bool first = st[RS_FIRST] != 0.0, last = st[RS_LAST] != 0.0;
        const int c = (int)st[RS_CHECK];
        const int maxit = ctl[C_MAXIT];
        const double norm = sqrt(rho) / sqrt(scal[S_BB]);
        int stop = 0, conv = 0, status = PP_OK, m = 0;
        bool den_fail = false;
        double alpha = 0.0, beta = 0.0;
        if (last) {
            stop = 1;
            m = maxit;
        } else if (!(norm == norm) || isinf(norm)) {
            stop = 1;
            status = PP_ERR_NOT_FINITE;
            m = c;
        } else if (norm < scal[S_TOL]) {
            stop = 1;
            conv = 1;
            m = c;
        } else {
            beta = first ? 0.0 : gamma / gold;
            const double den = first ? delta : delta - beta * gamma / aold;
            if (den == 0.0 || !(den == den)) {
                stop = 1;
                den_fail = true;
                status = den == 0.0 ? PP_ERR_BREAKDOWN : PP_ERR_NOT_FINITE;
                m = c + 1;
            } else {
                alpha = gamma / den;
            }
        }
        bc[0] = alpha;
        bc[1] = beta;
        *go = !stop;
        if (blockIdx.x == 0) {
            scal[S_CG_DELTA] = delta;
            scal[S_CG_GAMMA] = gamma;
            scal[S_CG_RHO] = rho;
            if (!last) record_norm(scal, ctl, norm);
            const bool next_last = !stop && c >= maxit - 1;
            if (stop) {
                // the state a resumed solve continues from
                scal[S_CG_GOLD] = gold;
                scal[S_CG_ALPHA] = aold;
                scal[S_CG_FIRST] = first ? 1.0 : 0.0;
                ctl[C_CONV] = conv;
                if (status != PP_OK) ctl[C_STATUS] = status;
                ctl[C_M] = m;
                if (den_fail) ctl[C_CHECK] = c + 1;  // the check itself had passed
                ctl[C_STOP] = 1;
            } else {
                nx[RS_GOLD] = gamma;
                nx[RS_AOLD] = alpha;
                nx[RS_FIRST] = 0.0;
                nx[RS_LAST] = next_last ? 1.0 : 0.0;
                nx[RS_CHECK] = (double)(c + 1);
                scal[S_CG_ALPHA] = alpha;
                scal[S_CG_BETA] = beta;
                scal[S_CG_GOLD] = gamma;
                scal[S_CG_FIRST] = 0.0;
                ctl[C_CHECK] = c + 1;
                if (next_last) ctl[C_LAST] = 1;
            }
            if (ring) publish_ctl(ctl, ring);
        }
    }
    __syncthreads();
    alpha_out = bc[0];
    beta_out = bc[1];
    return *go != 0;
}

// Fused finish + update of the reduced CG (see red_step in pp_internal.h).
template <class T>
__global__ __launch_bounds__(1024) void k_red_step(const double *__restrict__ partials, unsigned np, double *scal,
                                                   int *ctl, int *ring, int parity, uint32_t n,
                                                   const T *__restrict__ w, T *__restrict__ p, T *__restrict__ s,
                                                   T *__restrict__ x, T *__restrict__ r) {
#pragma clang fp contract(off)
    __shared__ double sh[3 * (kFinishThreads / 64)];
    __shared__ double bc[2];
    __shared__ int go;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (ctl[C_STOP]) {  // an earlier check stopped (or this launch's first workgroup already did)
        if (ring && lead) publish_ctl(ctl, ring);
        return;
    }
    // the first kU elements of this thread are loaded before the reduction
    // (they do not depend on alpha / beta), so their latency overlaps it
    constexpr int kU = 4;
    const uint32_t stride = gridDim.x * blockDim.x, i0 = blockIdx.x * blockDim.x + threadIdx.x;
    T vr[kU], vp[kU], vw[kU], vs[kU], vx[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        const uint32_t i = i0 + u * stride;
        vr[u] = vp[u] = vw[u] = vs[u] = vx[u] = 0;
        if (i < n) {
            vr[u] = r[i];
            vp[u] = p[i];
            vw[u] = w[i];
            vs[u] = s[i];
            vx[u] = x[i];
        }
    }
    double t[3], a64, b64;
#pragma unroll
    for (int q = 0; q < 3; ++q) t[q] = sum_partials(partials + (size_t)q * kMaxPartials, np);
    if (!red_check(t, 0, scal, ctl, ring, parity, sh, bc, &go, a64, b64)) return;
    const T alpha = (T)a64, beta = (T)b64;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        const uint32_t i = i0 + u * stride;
        if (i < n) {
            const T pn = vr[u] + beta * vp[u];
            const T sn = vw[u] + beta * vs[u];
            p[i] = pn;
            s[i] = sn;
            x[i] = vx[u] + alpha * pn;
            r[i] = vr[u] - alpha * sn;
        }
    }
    for (uint32_t i = i0 + kU * stride; i < n; i += stride) {
        const T ri = r[i];
        const T pn = ri + beta * p[i];
        const T sn = w[i] + beta * s[i];
        p[i] = pn;
        s[i] = sn;
        x[i] = x[i] + alpha * pn;
        r[i] = ri - alpha * sn;
    }
}

// The reduced-CG iteration in two launches instead of three (red_step split
// over the two products, sliced-column forms of KA and KB with exact
// widths).  k_red_ka: the check and step lengths of the previous iteration
// (red_check, as k_red_step), then y0 = KA r_new over the first colour,
// with the previous iteration's residual update folded into the gathers:
// r_new[j] = r[j] - alpha (w[j] + beta s[j]), the operations of k_red_step
// in its order, so the products see the bits the separate update stores.
// Nothing of the second colour is written here (other workgroups still
// gather the old r, w, s).  Every load the check and the thread's first
// rows need is issued before the first wait -- the stop flag, the partials,
// then the rows' matrix entries and gathers -- so the check's latency and
// the launch's start overlap the row traffic.  Addresses are 32-bit byte
// offsets from uniform bases (red_fused_ok: every vector under 4 GiB), one
// register per gathered neighbour for r, w and s together.
// k_red_ka's rows loaded ahead of the check per thread: as many as the
// registers of 1024-thread groups hold without spilling -- one for the wide
// FP64 forms, three for FP32 up to six entries a row, two otherwise (four
// FP32 rows of six spill, and measured no faster than three)
template <class T, int W>
constexpr int red_ka_rows() {
    return sizeof(T) == 8 && W >= 6 ? 1 : (sizeof(T) == 4 && W <= 6 ? 3 : 2);
}
template <class T, int W, bool PAT, int kU>
__global__ __launch_bounds__(1024) void k_red_ka(uint32_t n0, const int16_t *__restrict__ off,
                                                 const uint8_t *__restrict__ pid, const int16_t *__restrict__ ptab,
                                                 const T *__restrict__ val, int64_t cbase, const T *z, T *y,
                                                 uint32_t rbase, const T *__restrict__ w, const T *__restrict__ s,
                                                 const double *__restrict__ partials, unsigned np, double *scal,
                                                 int *ctl, int *ring, int parity) {
#pragma clang fp contract(off)
    __shared__ double sh[3 * (kFinishThreads / 64)];
    __shared__ double bc[2];
    __shared__ int go;
    // read by the one thread that tests it, as a vector load issued here (a
    // scalar or wave-uniform read would be waited for at once)
    int prior_stop = 0;
    if (threadIdx.x == 0) prior_stop = __hip_atomic_load(ctl + C_STOP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double t[3];  // at most one partial per thread (np <= kFinishThreads, red_ka)
#pragma unroll
    for (int q = 0; q < 3; ++q) t[q] = threadIdx.x < np ? partials[(size_t)q * kMaxPartials + threadIdx.x] : 0.0;
    // XCD-aware rows: the workgroups of XCD q (workgroup b runs on XCD b % 8)
    // take the contiguous eighth q of the rows, so the gathered neighbours
    // stay in that XCD's L2
    uint32_t hi = n0, stride = gridDim.x * blockDim.x, i0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (gridDim.x % kXcds == 0) {
        const uint32_t q = blockIdx.x % kXcds;
        const uint32_t lo = (uint32_t)((uint64_t)n0 * q / kXcds);
        hi = (uint32_t)((uint64_t)n0 * (q + 1) / kXcds);
        stride = gridDim.x / kXcds * blockDim.x;
        i0 = lo + blockIdx.x / kXcds * blockDim.x + threadIdx.x;
    }
    // the thread's first kU rows (red_ka_rows): matrix entries and every
    // gathered r, w, s loaded before the check (none of them depends on
    // alpha / beta).  The loads go out in three waves, for
    // all kU rows at once: pattern ids (or offsets), then the patterns and
    // the matrix entries, then the gathers -- so a thread waits on two
    // dependent latencies in all, not on two for every row (memory loads
    // complete in issue order).  Rows past the end load row n0 - 1 and
    // store nothing.
    constexpr uint32_t kB = sizeof(T);
    const char *zb = (const char *)z;  // z, w and s at the same byte offset of a second-colour index
    const char *wb = (const char *)w - (size_t)rbase * kB;
    const char *sb = (const char *)s - (size_t)rbase * kB;
    const uint32_t cb = (uint32_t)cbase;
    T v[kU][W], gr[kU][W], gw[kU][W], gs[kU][W];
    auto load_row = [&](uint32_t i, T *vv, T *rr, T *ww, T *ss) {
        int16_t o[W];
        if constexpr (PAT) ell_pattern<W>(ptab, pid[i], o);
        const uint32_t ib = i * kB;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if constexpr (!PAT) o[j] = off[(size_t)j * n0 + i];
            vv[j] = *(const T *)((const char *)val + (size_t)j * n0 * kB + ib);
        }
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const uint32_t gb = (i + cb + (uint32_t)(int32_t)o[j]) * kB;  // in z; the second colour starts at rbase
            rr[j] = *(const T *)(zb + gb);
            ww[j] = *(const T *)(wb + gb);
            ss[j] = *(const T *)(sb + gb);
        }
    };
    uint32_t ri[kU];
    int16_t o[kU][W];
    uint32_t pv[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        ri[u] = min(i0 + u * stride, n0 - 1);
        if constexpr (PAT) pv[u] = pid[ri[u]];
        else
#pragma unroll
            for (int j = 0; j < W; ++j) o[u][j] = off[(size_t)j * n0 + ri[u]];
    }
    __builtin_amdgcn_sched_barrier(0);
    int4 pt[kU];
    if constexpr (PAT)
#pragma unroll
        for (int u = 0; u < kU; ++u) pt[u] = *reinterpret_cast<const int4 *>(ptab + 8 * (size_t)pv[u]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
        for (int j = 0; j < W; ++j) v[u][j] = *(const T *)((const char *)val + (size_t)j * n0 * kB + ri[u] * kB);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        if constexpr (PAT) ell_unpack<W>(pt[u], o[u]);
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const uint32_t gb = (ri[u] + cb + (uint32_t)(int32_t)o[u][j]) * kB;
            gr[u][j] = *(const T *)(zb + gb);
            gw[u][j] = *(const T *)(wb + gb);
            gs[u][j] = *(const T *)(sb + gb);
        }
    }
    double a64, b64;
    if (!red_check(t, prior_stop, scal, ctl, ring, parity, sh, bc, &go, a64, b64)) return;
    const T alpha = (T)a64, beta = (T)b64;
    auto row = [&](uint32_t i, const T *vv, const T *rr, const T *ww, const T *ss) {
        T sum = 0;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const T sn = ww[j] + beta * ss[j];
            const T rn = rr[j] - alpha * sn;
            sum += vv[j] * rn;
        }
        y[i] = sum;
    };
#pragma unroll
    for (int u = 0; u < kU; ++u)
        if (i0 + u * stride < hi) row(i0 + u * stride, v[u], gr[u], gw[u], gs[u]);
    for (uint32_t i = i0 + kU * stride; i < hi; i += stride) {
        T vv[W], rr[W], ww[W], ss[W];
        load_row(i, vv, rr, ww, ss);
        row(i, vv, rr, ww, ss);
    }
}

// k_red_kb: the previous iteration's update of the second colour (p, s, x,
// r in place, k_red_step's operations; alpha and beta as k_red_ka left them
// in scal), then w = KB [y0; r_new] and the three inner products of the
// next check, exactly as k_ell with DOT computes them.
template <class T, int W, bool PAT>
__global__ __launch_bounds__(256) void k_red_kb(uint32_t n1, const int16_t *__restrict__ off,
                                                const uint8_t *__restrict__ pid, const int16_t *__restrict__ ptab,
                                                const T *__restrict__ val, int64_t cbase, const T *__restrict__ dg,
                                                T *z, uint32_t rbase, T *__restrict__ w, T *__restrict__ p,
                                                T *__restrict__ s, T *__restrict__ x, const T *__restrict__ wgt,
                                                const double *scal, double *partials, const int *stop) {
#pragma clang fp contract(off)
    __shared__ double sh[3 * (kBlock / 64)];
    if (stop && *stop) return;
    const T alpha = (T)scal[S_CG_ALPHA], beta = (T)scal[S_CG_BETA];
    double dacc = 0.0, dacc2 = 0.0, dacc3 = 0.0;
    const uint32_t nb = (n1 + kBlock - 1) / kBlock;
    const XcdRange xr = xcd_range(0, nb);
    for (uint32_t b = xr.first; b < xr.end; b += xr.step) {
        const uint32_t i = b * kBlock + threadIdx.x;
        if (i < n1) {
            int16_t o[W];
            T v[W], xv[W];
            if constexpr (PAT) ell_pattern<W>(ptab, pid[i], o);
#pragma unroll
            for (int j = 0; j < W; ++j) {
                const size_t e = (size_t)j * n1 + i;
                if constexpr (!PAT) o[j] = off[e];
                v[j] = val[e];
            }
            const T dgi = dg[i], ri = z[rbase + i], pi = p[i], wi = w[i], si = s[i], xi = x[i], wg = wgt[i];
#pragma unroll
            for (int j = 0; j < W; ++j) xv[j] = z[(int64_t)i + cbase + o[j]];
            const T pn = ri + beta * pi;
            const T sn = wi + beta * si;
            const T rn = ri - alpha * sn;
            p[i] = pn;
            s[i] = sn;
            x[i] = xi + alpha * pn;
            z[rbase + i] = rn;
            T sum = 0;
#pragma unroll
            for (int j = 0; j < W; ++j) sum += v[j] * xv[j];
            sum += dgi * rn;
            const T out = T(1) * sum;
            w[i] = out;
            dacc += (double)rn * (double)out;
            dacc2 += (double)rn * (double)rn;
            dacc3 += (double)rn * (double)rn * (double)wg;
        }
    }
    block_sum3(dacc, dacc2, dacc3, sh);
    publish_partial(dacc, partials);
    publish_partial(dacc2, partials + kMaxPartials);
    publish_partial(dacc3, partials + 2 * kMaxPartials);
}

template <class T>
void red_step(pp_ctx *ctx, unsigned np, int parity, bool publish, uint32_t n, const T *w, T *p, T *s, T *x, T *r) {
    // one 1024-thread workgroup per CU: each sums the partials once
    const unsigned g = std::max(1u, std::min<unsigned>((unsigned)ctx->cu_count,
                                                       (unsigned)((n + kFinishThreads * 4 - 1) / (kFinishThreads * 4))));
    k_red_step<T><<<g, kFinishThreads, 0, ctx->stream>>>(ctx->red.partials + (size_t)S_CG_DELTA * kMaxPartials, np,
                                                        ctx->scal, ctx->ctl, publish ? ctx->ring_dev : nullptr,
                                                        parity, n, w, p, s, x, r);
    PP_HIP(hipGetLastError());
}
template void red_step<double>(pp_ctx *, unsigned, int, bool, uint32_t, const double *, double *, double *,
                               double *, double *);
template void red_step<float>(pp_ctx *, unsigned, int, bool, uint32_t, const float *, float *, float *, float *,
                              float *);

bool red_fused_ok(const EllOp &ka, const EllOp &kb, int dtype) {
    const bool vals = dtype == PP_F64 ? (ka.v64 && kb.v64 && kb.dg64) : (ka.v32 && kb.v32 && kb.dg32);
    // k_red_ka addresses the vectors (n0 + n1 entries) by 32-bit byte offsets
    const bool small = ((uint64_t)ka.n + kb.n) * (dtype == PP_F64 ? 8 : 4) < (1ull << 32);
    return ka.ok && kb.ok && vals && small && ka.width >= 4 && ka.width <= 7 && kb.width >= 4 && kb.width <= 7 &&
           !ka.dg64 && !ka.dg32 && ka.n > 0 && kb.n > 0;
}

template <class T>
void red_ka(pp_ctx *ctx, const EllOp &E, const T *z, T *y, uint32_t rbase, const T *w, const T *s, unsigned np,
            int parity, bool publish) {
    // one 1024-thread workgroup per CU (each sums the partials once; all
    // resident at once, so the check runs in one round)
    PP_CHECK(np <= (unsigned)kFinishThreads, PP_ERR_INVALID, "reduced CG: more partials than k_red_ka sums");
    const unsigned want = (E.n + kFinishThreads - 1) / kFinishThreads;
    unsigned g = std::max(1u, std::min<unsigned>((unsigned)ctx->cu_count, want));
    if (g >= kXcds) g -= g % kXcds;
    const T *val = sizeof(T) == 8 ? (const T *)E.v64 : (const T *)E.v32;
    double *part = ctx->red.partials + (size_t)S_CG_DELTA * kMaxPartials;
    int *ring = publish ? ctx->ring_dev : nullptr;
#define PP_RKA(W, P)                                                                                              \
    k_red_ka<T, W, P, red_ka_rows<T, W>()><<<g, kFinishThreads, 0, ctx->stream>>>(                                \
        E.n, E.off, E.pid, E.ptab, val, E.cbase, z, y, rbase, w, s, part, np, ctx->scal, ctx->ctl, ring, parity)
#define PP_RKA_P(W) PP_RKA(W, true)
#define PP_RKA_N(W) PP_RKA(W, false)
    if (E.pid) {
        PP_ELL_PAT_WIDTHS(E.width, PP_RKA_P)
    } else {
        PP_ELL_PAT_WIDTHS(E.width, PP_RKA_N)
    }
#undef PP_RKA_P
#undef PP_RKA_N
#undef PP_RKA
    PP_HIP(hipGetLastError());
}

template <class T>
void red_kb(pp_ctx *ctx, const EllOp &E, T *z, uint32_t rbase, T *w, T *p, T *s, T *x, const T *wt, unsigned grid) {
    const T *val = sizeof(T) == 8 ? (const T *)E.v64 : (const T *)E.v32;
    const T *dg = sizeof(T) == 8 ? (const T *)E.dg64 : (const T *)E.dg32;
    double *part = ctx->red.partials + (size_t)S_CG_DELTA * kMaxPartials;
#define PP_RKB(W, P)                                                                                              \
    k_red_kb<T, W, P><<<grid, kBlock, 0, ctx->stream>>>(E.n, E.off, E.pid, E.ptab, val, E.cbase, dg, z, rbase, w, p, \
                                                       s, x, wt, ctx->scal, part, ctx->stop_word)
#define PP_RKB_P(W) PP_RKB(W, true)
#define PP_RKB_N(W) PP_RKB(W, false)
    if (E.pid) {
        PP_ELL_PAT_WIDTHS(E.width, PP_RKB_P)
    } else {
        PP_ELL_PAT_WIDTHS(E.width, PP_RKB_N)
    }
#undef PP_RKB_P
#undef PP_RKB_N
#undef PP_RKB
    PP_HIP(hipGetLastError());
}
template void red_ka<double>(pp_ctx *, const EllOp &, const double *, double *, uint32_t, const double *,
                             const double *, unsigned, int, bool);
template void red_ka<float>(pp_ctx *, const EllOp &, const float *, float *, uint32_t, const float *, const float *,
                            unsigned, int, bool);
template void red_kb<double>(pp_ctx *, const EllOp &, double *, uint32_t, double *, double *, double *, double *,
                             const double *, unsigned);
template void red_kb<float>(pp_ctx *, const EllOp &, float *, uint32_t, float *, float *, float *, float *,
                            const float *, unsigned);

__global__ void k_red_init(double *scal, int resume, double gscale, int last) {
    double *st = scal + S_RED0;
    if (resume) {
        st[RS_GOLD] = scal[S_CG_GOLD] * gscale;
        st[RS_AOLD] = scal[S_CG_ALPHA];
        st[RS_FIRST] = scal[S_CG_FIRST];
    } else {
        st[RS_GOLD] = 0.0;
        st[RS_AOLD] = 1.0;
        st[RS_FIRST] = 1.0;
    }
    st[RS_LAST] = last ? 1.0 : 0.0;
    st[RS_CHECK] = 1.0;
}

// Everything a reduced solve sets before its first iteration in one launch:
// the control words (check 1, last when maxit <= 1), the tolerance, ||rin||^2
// (when given, bb >= 0) and the iteration state (as k_red_init).
__global__ void k_red_begin(int *ctl, double *scal, int maxit, double tol, double bb, int resume, double gscale) {
    if (threadIdx.x < C_COUNT) ctl[threadIdx.x] = 0;
    __syncthreads();
    if (threadIdx.x != 0) return;
    ctl[C_CHECK] = 1;
    ctl[C_LAST] = maxit <= 1 ? 1 : 0;
    ctl[C_MAXIT] = maxit;
    scal[S_TOL] = tol;
    if (bb >= 0.0) scal[S_BB] = bb;
    double *st = scal + S_RED0;
    if (resume) {
        st[RS_GOLD] = scal[S_CG_GOLD] * gscale;
        st[RS_AOLD] = scal[S_CG_ALPHA];
        st[RS_FIRST] = scal[S_CG_FIRST];
    } else {
        st[RS_GOLD] = 0.0;
        st[RS_AOLD] = 1.0;
        st[RS_FIRST] = 1.0;
    }
    st[RS_LAST] = maxit <= 1 ? 1.0 : 0.0;
    st[RS_CHECK] = 1.0;
}

// x = 0; p, s = resume ? scale * (p, s) : 0
template <class T>
__global__ void k_red_vectors(uint32_t n, T *__restrict__ x, T *__restrict__ p, T *__restrict__ s, int resume,
                              T scale) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] = 0;
        p[i] = resume ? scale * p[i] : (T)0;
        s[i] = resume ? scale * s[i] : (T)0;
    }
}

void red_begin(pp_ctx *ctx, uint32_t maxit, double tol, double bb, bool resume, double scale, int dtype,
               uint32_t n1, void *x, void *p, void *s) {
    k_red_begin<<<1, 64, 0, ctx->stream>>>(ctx->ctl, ctx->scal, (int)maxit, tol, bb, resume ? 1 : 0, scale * scale);
    const unsigned g = grid_for(ctx, n1, 4);
    if (dtype == PP_F64)
        k_red_vectors<double><<<g, kBlock, 0, ctx->stream>>>(n1, (double *)x, (double *)p, (double *)s,
                                                              resume ? 1 : 0, scale);
    else
        k_red_vectors<float><<<g, kBlock, 0, ctx->stream>>>(n1, (float *)x, (float *)p, (float *)s, resume ? 1 : 0,
                                                            (float)scale);
    PP_HIP(hipGetLastError());
}

void red_init(pp_ctx *ctx, bool resume, double gscale, bool last) {
    k_red_init<<<1, 1, 0, ctx->stream>>>(ctx->scal, resume ? 1 : 0, gscale, last ? 1 : 0);
    PP_HIP(hipGetLastError());
}

// Back-substitution staging of the reduced solve: e1 = s1 * x into both
// z[n0..) and e[n0..); z[0..n0) = rin[0..n0) (the first colour's residual,
// which the elimination step D0^{-1} (r0 - A01 e1) reads next).
template <class T>
__global__ void k_red_post(uint32_t n0, uint32_t n1, const T *__restrict__ x, const T *__restrict__ s1,
                           const T *__restrict__ rin, T *__restrict__ z, T *__restrict__ e) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n0 + n1; i += gridDim.x * blockDim.x) {
        if (i < n0) {
            z[i] = rin[i];
        } else {
            const T v = x[i - n0] * s1[i - n0];
            z[i] = v;
            e[i] = v;
        }
    }
}

template <class T>
void red_post(pp_ctx *ctx, uint32_t n0, uint32_t n1, const T *x, const T *s1, const T *rin, T *z, T *e) {
    k_red_post<T><<<grid_for(ctx, (uint64_t)n0 + n1, 4), kBlock, 0, ctx->stream>>>(n0, n1, x, s1, rin, z, e);
    PP_HIP(hipGetLastError());
}
template void red_post<double>(pp_ctx *, uint32_t, uint32_t, const double *, const double *, const double *,
                               double *, double *);
template void red_post<float>(pp_ctx *, uint32_t, uint32_t, const float *, const float *, const float *, float *,
                              float *);

// Sets the control words for a fresh PCG solve (graph-safe: a kernel, no memset).
__global__ void k_ctl_init(int *ctl, int first_check, int last, int maxit) {
    if (threadIdx.x < C_COUNT) ctl[threadIdx.x] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        ctl[C_CHECK] = first_check;
        ctl[C_LAST] = last;
        ctl[C_MAXIT] = maxit;
    }
}

StopScope::StopScope(pp_ctx *ctx) : c(ctx), prev(ctx->stop_word) { ctx->stop_word = ctx->ctl + C_STOP; }

void ctl_init(pp_ctx *ctx, int first_check, int last, uint32_t maxit) {
    k_ctl_init<<<1, 64, 0, ctx->stream>>>(ctx->ctl, first_check, last, (int)maxit);
    PP_HIP(hipGetLastError());
}

__global__ void k_set_slot(double *scal, int slot, double v) { scal[slot] = v; }
void set_slot(pp_ctx *ctx, int slot, double v) {
    k_set_slot<<<1, 1, 0, ctx->stream>>>(ctx->scal, slot, v);
    PP_HIP(hipGetLastError());
}

__global__ void k_scale_slot(double *scal, int slot, double v) { scal[slot] *= v; }
void scale_slot(pp_ctx *ctx, int slot, double v) {
    k_scale_slot<<<1, 1, 0, ctx->stream>>>(ctx->scal, slot, v);
    PP_HIP(hipGetLastError());
}

// scal[dst] = scal[src] (used to keep (b,b) while reusing slots)
__global__ void k_copy_slot(double *scal, int dst, int src) { scal[dst] = scal[src]; }
void copy_slot(pp_ctx *ctx, int dst, int src) {
    k_copy_slot<<<1, 1, 0, ctx->stream>>>(ctx->scal, dst, src);
    PP_HIP(hipGetLastError());
}

}  // namespace pp

namespace pp {
template <class T>
__global__ void k_permute(uint32_t n, const uint32_t *__restrict__ perm, const T *__restrict__ src, T *__restrict__ dst,
                          int gather) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (gather) dst[k] = src[perm[k]];
        else dst[perm[k]] = src[k];
    }
}

// gather: dst[k] = src[perm[k]]; scatter: dst[perm[k]] = src[k]
void launch_permute(pp_ctx *ctx, uint32_t n, const uint32_t *perm, int dtype, const void *src, void *dst,
                    bool gather) {
    const unsigned g = grid_for(ctx, n);
    if (dtype == PP_F64)
        k_permute<double><<<g, kBlock, 0, ctx->stream>>>(n, perm, (const double *)src, (double *)dst, gather);
    else
        k_permute<float><<<g, kBlock, 0, ctx->stream>>>(n, perm, (const float *)src, (float *)dst, gather);
    PP_HIP(hipGetLastError());
}
}  // namespace pp
