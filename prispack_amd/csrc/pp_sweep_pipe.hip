// pp_sweep_pipe.hip -- pipelined one-workgroup-per-chunk triangular sweeps.
//
// A chunk (PP_TRSV_BLOCK storage: a dependency-closed row range whose part
// of the vector fits in LDS) with many narrow levels -- an incomplete factor
// in natural order has hundreds to a thousand levels of 15-40 rows -- is
// bound by the latency of each level, not by bytes.  The plain chunked
// kernel separates levels with two workgroup barriers and reads each
// level's entries from global memory one level ahead; here one wavefront
// does all the arithmetic with no barrier at all, and the rest of the
// workgroup keeps the level data it needs waiting in LDS:
//
//   wave 0 (the consumer): the levels in order, one row per lane (64 rows a
//     pass), the lane's entries (up to kPipeK) held in registers.  While a
//     level computes, the next level's rows and entries are read from the
//     LDS ring into the other register set.  LDS operations of one wave
//     complete in issue order, so a level's solution values are in place
//     before the next level reads them without any wait or fence.
//   waves 1..P (the producers): wave w copies levels w-1, w-1+P, ... from
//     global memory into the ring.  The global loads are issued before the
//     wave waits for ring space, so their latency overlaps the consumer's
//     work on earlier levels; a level is published with a workgroup-scope
//     release store of its number into a descriptor slot.
//
// The ring is indexed by the absolute entry / row position modulo its size.
// A level is staged only when it fits in the ring together with the level
// before it (the consumer holds both at once); otherwise it is marked
// direct and the consumer reads its entries from global memory.  Every wait
// is for a condition that the consumer's progress through earlier levels
// satisfies, so no wait can block forever; each still gives up after
// kPipeSpinCap polls (the workgroup then stops and ctl[C_SPIN] is set, which
// the caller turns into an error).
//
// Per row the operations and their order are those of k_sweep_level
// (pre-division for mode 3, the products each rounded and subtracted in
// storage order, then the row's diagonal mode), so the results are the same
// bit for bit.
#include <hip/hip_runtime.h>

#include "pp_internal.h"

namespace pp {
namespace {

constexpr int kPipeK = 24;           // entries of a row whose columns the consumer reads ahead
constexpr int kPipeB = 12;           // entries per producer lane loaded before the ring wait
constexpr uint32_t kPipeSlots = 32;  // level descriptors in flight
constexpr uint32_t kPipeDirect = 0x80000000u;
constexpr int kPipeSpinCap = 1 << 22;

__device__ inline int lds_acquire(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_release(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The LDS carve of one workgroup: the chunk's vector, the entry ring (values
// and chunk-local columns), the row ring (row, first entry, length, diagonal),
// the level descriptors and the progress words.
template <class T>
struct PipeLds {
    T *xs, *rval, *rd;
    uint32_t *re0, *meta;
    uint16_t *rci, *rrow, *rlen;
    int *ready, *ctr;  // ctr: [0] last consumed level, [1] consumed entries, [2] consumed rows, [3] abort
    uint32_t me, mr;   // ring sizes minus one (powers of two)
};

template <class T>
__device__ inline PipeLds<T> pipe_carve(unsigned char *raw, uint32_t max_rows, uint32_t re, uint32_t rr) {
    PipeLds<T> P;
    P.xs = reinterpret_cast<T *>(raw);
    P.rval = P.xs + max_rows;
    P.rd = P.rval + re;
    P.re0 = reinterpret_cast<uint32_t *>(P.rd + rr);
    P.rci = reinterpret_cast<uint16_t *>(P.re0 + rr);
    P.rrow = P.rci + re;
    P.rlen = P.rrow + rr;
    P.meta = reinterpret_cast<uint32_t *>(P.rlen + rr);
    P.ready = reinterpret_cast<int *>(P.meta + 4 * kPipeSlots);
    P.ctr = P.ready + kPipeSlots;
    P.me = re - 1;
    P.mr = rr - 1;
    return P;
}

// Polls until *p >= want (or == want for eq); false after the cap (the
// abort word is set then, and any wave that sees it stops).
template <bool EQ>
__device__ inline bool pipe_wait(const int *p, int want, int *abort_word, int *ctl, bool sleep) {
    for (int spins = 0;; ++spins) {
        const int v = lds_acquire(p);
        if (EQ ? v == want : v >= want) return true;
        if (lds_acquire(abort_word)) return false;
        if (spins >= kPipeSpinCap) {
            if (__lane_id() == 0) {
                lds_release(abort_word, 1);
                atomicOr(&ctl[C_SPIN], 1);
            }
            return false;
        }
        if (sleep) __builtin_amdgcn_s_sleep(2);
    }
}

// One level's descriptor as the consumer holds it (wave-uniform).
struct PipeDesc {
    uint32_t L0, L1, E0, E1;
    bool direct;
};

// A lane's row of one level, read ahead from the ring: the row, its entry
// range, its diagonal and its first kPipeK columns (two 16-bit columns a
// word).  The values are read when the row computes, together with the
// vector entries those columns select (one LDS round trip for both).
template <class T>
struct PipeRow {
    uint32_t i, e0, len;
    T d;
    uint32_t cp[kPipeK / 2];
};

template <class T>
__device__ inline void pipe_read_desc(const PipeLds<T> &P, uint32_t l, PipeDesc &D) {
    const uint32_t *m = P.meta + 4 * (l & (kPipeSlots - 1));
    D.L0 = m[0];
    D.L1 = m[1];
    D.E0 = m[2];
    D.E1 = m[3] & ~kPipeDirect;
    D.direct = (m[3] & kPipeDirect) != 0;
}

// Row k of a staged level from the ring into registers.  Slots past the
// row's length hold other data (or none yet): their column becomes 0, so the
// vector read they cause stays inside the chunk, and their value is never
// used.
template <class T>
__device__ inline void pipe_read_row(const PipeLds<T> &P, uint32_t k, PipeRow<T> &R) {
    const uint32_t r = k & P.mr;
    R.i = P.rrow[r];
    R.e0 = P.re0[r];
    R.len = P.rlen[r];
    R.d = P.rd[r];
#pragma unroll
    for (int j = 0; j < kPipeK; j += 2) {
        const uint32_t c0 = P.rci[(R.e0 + (uint32_t)j) & P.me], c1 = P.rci[(R.e0 + (uint32_t)j + 1) & P.me];
        R.cp[j / 2] = ((uint32_t)j < R.len ? c0 : 0u) | (((uint32_t)j + 1 < R.len ? c1 : 0u) << 16);
    }
}

template <class T>
__device__ inline void pipe_finish(T s, uint32_t i, T d, int mode, T *xs) {
#pragma clang fp contract(off)
    if (mode == 0) s = s / d;
    else if (mode == 2) s = s * d;
    xs[i] = s;
}

// One row: the first kPipeK entries with their columns from registers, the
// rest eight at a time from the ring.
template <class T>
__device__ inline void pipe_row(const PipeLds<T> &P, const PipeRow<T> &R, int mode) {
#pragma clang fp contract(off)
    T vv[kPipeK], xv[kPipeK];
#pragma unroll
    for (int j = 0; j < kPipeK; ++j) {
        vv[j] = P.rval[(R.e0 + (uint32_t)j) & P.me];
        xv[j] = P.xs[(R.cp[j / 2] >> (16 * (j & 1))) & 0xffffu];
    }
    T s = P.xs[R.i];
    if (mode == 3) s = s / R.d;
#pragma unroll
    for (int j = 0; j < kPipeK; ++j) {
        const T p = vv[j] * xv[j];
        s = s - ((uint32_t)j < R.len ? p : T(0));  // x - (+0) is x: entries past the row change nothing
    }
    for (uint32_t j = kPipeK; j < R.len; j += 8) {
        T tv[8], tx[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t pos = (R.e0 + min(j + u, R.len - 1)) & P.me;
            tv[u] = P.rval[pos];
            tx[u] = P.xs[P.rci[pos]];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const T p = tv[u] * tx[u];
            s = s - (j + u < R.len ? p : T(0));
        }
    }
    pipe_finish(s, R.i, R.d, mode, P.xs);
}

// A direct level: each lane's rows straight from the global arrays.
template <class T>
__device__ inline void pipe_direct(const PipeLds<T> &P, const DevChunkSweep<T> &S, const PipeDesc &D, int mode) {
#pragma clang fp contract(off)
    for (uint32_t k = D.L0 + __lane_id(); k < D.L1; k += 64) {
        const uint32_t i = S.row[k];
        const T d = S.pdiag ? S.pdiag[k] : T(1);
        T s = P.xs[i];
        if (mode == 3) s = s / d;
        for (uint32_t e = S.rp[k], e1 = S.rp[k + 1]; e < e1; ++e) {
            const T p = S.val[e] * P.xs[S.ci[e]];
            s = s - p;
        }
        pipe_finish(s, i, d, mode, P.xs);
    }
}

// Consumer step: read level l+1 ahead (descriptor, then the lane's row) into
// (dn, rn), compute level l from (dc, rc), publish l as consumed.
template <class T>
__device__ inline bool pipe_step(const PipeLds<T> &P, const DevChunkSweep<T> &S, uint32_t l, uint32_t nlev,
                                 int mode, const PipeDesc &dc, const PipeRow<T> &rc, PipeDesc &dn, PipeRow<T> &rn,
                                 int *ctl) {
    const uint32_t lane = __lane_id();
    if (l + 1 < nlev) {
        if (!pipe_wait<true>(P.ready + ((l + 1) & (kPipeSlots - 1)), (int)(l + 2), P.ctr + 3, ctl, false))
            return false;
        pipe_read_desc(P, l + 1, dn);
        if (!dn.direct && dn.L0 + lane < dn.L1) pipe_read_row(P, dn.L0 + lane, rn);
    }
    if (dc.direct) {
        pipe_direct(P, S, dc, mode);
    } else {
        if (dc.L0 + lane < dc.L1) pipe_row(P, rc, mode);
        for (uint32_t k = dc.L0 + 64 + lane; k < dc.L1; k += 64) {  // rows past the first 64
            PipeRow<T> R;
            pipe_read_row(P, k, R);
            pipe_row(P, R, mode);
        }
    }
    if (lane == 0) {
        // the ring slots of level l are free once its last reads are done;
        // the solution values stay in LDS (only this wave reads them)
        lds_release(P.ctr + 1, (int)dc.E1);
        lds_release(P.ctr + 2, (int)dc.L1);
        lds_release(P.ctr, (int)l);
    }
    return true;
}

template <class T>
__device__ void pipe_consume(const PipeLds<T> &P, const DevChunkSweep<T> &S, uint32_t nlev, int mode, int *ctl) {
    if (nlev == 0) return;
    PipeDesc da, db;
    PipeRow<T> ra, rb;
    if (!pipe_wait<true>(P.ready, 1, P.ctr + 3, ctl, false)) return;
    pipe_read_desc(P, 0, da);
    if (!da.direct && da.L0 + __lane_id() < da.L1) pipe_read_row(P, da.L0 + __lane_id(), ra);
    for (uint32_t l = 0; l < nlev; l += 2) {
        if (!pipe_step(P, S, l, nlev, mode, da, ra, db, rb, ctl)) return;
        if (l + 1 < nlev && !pipe_step(P, S, l + 1, nlev, mode, db, rb, da, ra, ctl)) return;
    }
}

// Producer wave w of np: levels w, w + np, ... of the sweep whose level
// boundaries are lp/le[0..nlev].
template <class T>
__device__ void pipe_produce(const PipeLds<T> &P, const DevChunkSweep<T> &S, const uint32_t *lp, const uint32_t *le,
                             uint32_t nlev, uint32_t w, uint32_t np, uint32_t re, uint32_t rr, int *ctl) {
    const uint32_t lane = __lane_id();
    for (uint32_t l = w; l < nlev; l += np) {
        const uint32_t L0 = lp[l], L1 = lp[l + 1], E0 = le[l], E1 = le[l + 1];
        const uint32_t Lp = l ? lp[l - 1] : L0, Ep = l ? le[l - 1] : E0;
        // staged only when the ring holds this level and the one before it at once
        const bool fits = E1 - Ep <= re && L1 - Lp <= rr;
        if (fits) {
            uint16_t c[kPipeB];
            T v[kPipeB];
#pragma unroll
            for (int t = 0; t < kPipeB; ++t) {
                const uint32_t e = E0 + lane + 64u * t;
                const uint32_t es = e < E1 ? e : 0u;  // index 0 exists in every array
                c[t] = S.ci[es];
                v[t] = S.val[es];
            }
            const uint32_t k = L0 + lane, ks = k < L1 ? k : 0u;
            const uint16_t row = S.row[ks];
            const uint32_t e0 = S.rp[ks], e1 = S.rp[ks + 1];
            const T d = S.pdiag ? S.pdiag[ks] : T(1);
            // ring space: every entry and row before E1 - re / L1 - rr consumed
            if ((int)(E1 - re) > 0 && !pipe_wait<false>(P.ctr + 1, (int)(E1 - re), P.ctr + 3, ctl, true)) return;
            if ((int)(L1 - rr) > 0 && !pipe_wait<false>(P.ctr + 2, (int)(L1 - rr), P.ctr + 3, ctl, true)) return;
#pragma unroll
            for (int t = 0; t < kPipeB; ++t) {
                const uint32_t e = E0 + lane + 64u * t;
                if (e < E1) {
                    P.rci[e & P.me] = c[t];
                    P.rval[e & P.me] = v[t];
                }
            }
            for (uint32_t e = E0 + lane + 64u * kPipeB; e < E1; e += 64) {
                P.rci[e & P.me] = S.ci[e];
                P.rval[e & P.me] = S.val[e];
            }
            if (k < L1) {
                const uint32_t r = k & P.mr;
                P.rrow[r] = row;
                P.re0[r] = e0;
                P.rlen[r] = (uint16_t)(e1 - e0);
                P.rd[r] = d;
            }
            for (uint32_t q = k + 64; q < L1; q += 64) {
                const uint32_t r = q & P.mr;
                P.rrow[r] = S.row[q];
                P.re0[r] = S.rp[q];
                P.rlen[r] = (uint16_t)(S.rp[q + 1] - S.rp[q]);
                P.rd[r] = S.pdiag ? S.pdiag[q] : T(1);
            }
        }
        // the descriptor slot: level l - kPipeSlots consumed
        if (l >= kPipeSlots && !pipe_wait<false>(P.ctr, (int)(l - kPipeSlots), P.ctr + 3, ctl, true)) return;
        if (lane == 0) {
            uint32_t *m = P.meta + 4 * (l & (kPipeSlots - 1));
            m[0] = L0;
            m[1] = L1;
            m[2] = E0;
            m[3] = E1 | (fits ? 0u : kPipeDirect);
        }
        // release: this wave's ring and descriptor stores before the flag
        if (lane == 0) lds_release(P.ready + (l & (kPipeSlots - 1)), (int)(l + 1));
    }
}

template <class T>
__global__ __launch_bounds__(512) void k_sweep_chunks_pipe(const uint32_t *__restrict__ start, DevChunkSweep<T> lo,
                                                           int lo_mode, DevChunkSweep<T> up, int up_mode, const T *x,
                                                           T *y, uint32_t max_rows, uint32_t re, uint32_t rr,
                                                           const int *stop, int *ctl) {
    extern __shared__ unsigned char lds_raw[];
    if (stop && *stop) return;
    const PipeLds<T> P = pipe_carve<T>(lds_raw, max_rows, re, rr);
    const uint32_t c = blockIdx.x, r0 = start[c], r1 = start[c + 1];
    const uint32_t wave = threadIdx.x / 64, nwaves = blockDim.x / 64;
    for (uint32_t j = threadIdx.x; j < r1 - r0; j += blockDim.x) P.xs[j] = x[r0 + j];
    for (int s = 0; s < 2; ++s) {
        const DevChunkSweep<T> &S = s == 0 ? lo : up;
        const int mode = s == 0 ? lo_mode : up_mode;
        const uint32_t b0 = S.lev_off[c], nlev = S.lev_off[c + 1] - b0 - 1;
        const uint32_t *lp = S.lp + b0, *le = S.le + b0;
        if (threadIdx.x < kPipeSlots) P.ready[threadIdx.x] = 0;
        if (threadIdx.x == 0) {
            P.ctr[0] = -1;
            P.ctr[1] = (int)le[0];
            P.ctr[2] = (int)lp[0];
            P.ctr[3] = 0;
        }
        __syncthreads();
        if (wave == 0) pipe_consume<T>(P, S, nlev, mode, ctl);
        else pipe_produce<T>(P, S, lp, le, nlev, wave - 1, nwaves - 1, re, rr, ctl);
        __syncthreads();
    }
    for (uint32_t j = threadIdx.x; j < r1 - r0; j += blockDim.x) y[r0 + j] = P.xs[j];
}

}  // namespace

// LDS bytes of the pipelined kernel for rings of re entries and rr rows.
template <class T>
static size_t pipe_lds_bytes(uint32_t max_rows, uint32_t re, uint32_t rr) {
    return (size_t)max_rows * sizeof(T) + (size_t)re * (sizeof(T) + 2) + (size_t)rr * (sizeof(T) + 8) +
           kPipeSlots * 4 * sizeof(uint32_t) + kPipeSlots * sizeof(int) + 4 * sizeof(int);
}

// The ring sizes the chunk set gets (the largest entry ring that fits next
// to the vector and a 256-row ring); false when under 512 entries fit.
template <class T>
bool sweep_pipe_rings(pp_ctx *ctx, const ChunkSet &C, uint32_t &re, uint32_t &rr) {
    rr = 256;
    re = 0;
    for (uint32_t e = 1u << 15; e >= 512; e >>= 1)
        if (pipe_lds_bytes<T>(C.max_rows, e, rr) <= ctx->max_lds) {
            re = e;
            break;
        }
    return re != 0;
}
template bool sweep_pipe_rings<double>(pp_ctx *, const ChunkSet &, uint32_t &, uint32_t &);
template bool sweep_pipe_rings<float>(pp_ctx *, const ChunkSet &, uint32_t &, uint32_t &);

template <class T>
void launch_sweep_pair_pipe(pp_ctx *ctx, const ChunkSet &C, const DevChunkSweep<T> &lo, int mlo,
                            const DevChunkSweep<T> &up, int mup, const T *x, T *y) {
    uint32_t re = 0, rr = 0;
    PP_CHECK(C.nchunk && lo.row && up.row && sweep_pipe_rings<T>(ctx, C, re, rr), PP_ERR_INVALID,
             "pipelined chunk sweep: storage missing or chunk exceeds LDS");
    const size_t lds = pipe_lds_bytes<T>(C.max_rows, re, rr);
    if (lds > 65536)
        PP_HIP(hipFuncSetAttribute((const void *)k_sweep_chunks_pipe<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    k_sweep_chunks_pipe<T><<<C.nchunk, 512, lds, ctx->stream>>>(C.dstart, lo, mlo, up, mup, x, y, C.max_rows, re, rr,
                                                               ctx->stop_word, ctx->ctl);
    PP_HIP(hipGetLastError());
}
template void launch_sweep_pair_pipe<double>(pp_ctx *, const ChunkSet &, const DevChunkSweep<double> &, int,
                                             const DevChunkSweep<double> &, int, const double *, double *);
template void launch_sweep_pair_pipe<float>(pp_ctx *, const ChunkSet &, const DevChunkSweep<float> &, int,
                                            const DevChunkSweep<float> &, int, const float *, float *);

}  // namespace pp
