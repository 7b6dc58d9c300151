// pp_setup_dev.hip -- IC(0) and ILU(0) in multicolour order set up on the device.
//
// The host setup (pp_precond.cpp: setup_triangular, upload_sweep) colours
// the matrix, forms P A P^T, factors it and lays the two sweeps out in
// level order; every step is pattern work over the whole matrix and the
// arrays then cross PCIe.  Here the same steps run on the device from the
// matrix's device arrays, with the host's results:
//   colouring   greedy in row order (each row the smallest colour none of its
//               lower neighbours has), computed in rounds: a row is coloured
//               once all its lower neighbours are, so every row sees the
//               colours the serial pass gives it;
//   order       rows stably sorted by colour (hipCUB radix sort);
//   P A P^T     rows gathered, columns renumbered and sorted;
//   factor      the strict lower pattern, its dependency levels (rounds
//               again), the numeric IC(0) or ILU(0) level by level
//               (k_ic0_level, k_ilu0_level: the host's operation order);
//   sweeps      forward = L's rows, backward = L^T's rows (IC(0): the
//               matrix's upper rows, values found in L by binary search) or
//               U's rows (ILU(0): the upper entries themselves), each in
//               level order with the row blocks of the wide levels and the
//               sliced-column levels; the values reach them through maps
//               (factor entry per slot) that the re-factorization reuses.
// The row blocks' 16-bit column offsets are computed on the device too
// (k_two_base).  The host copies of the factor, of its sweeps and of
// P A P^T's columns and values are not made; a caller that needs them
// (the factor export, the chunked schedules) runs the host setup then.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <memory>
#include <vector>

#include "pp_internal.h"

namespace pp {
namespace {

constexpr int kB = 256;
constexpr uint32_t kMaxRow = 64;  // longest row the device path sorts in place

unsigned grid_of(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kB - 1) / kB, 32768)); }

// ---- hipCUB (setup only: temporary storage allocated per call)
void exclusive_scan(pp_ctx *ctx, const uint32_t *in, uint32_t *out, uint32_t n) {
    size_t bytes = 0;
    PP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, ctx->stream));
    unsigned char *tmp = dalloc<unsigned char>(std::max<size_t>(bytes, 1));
    const hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    dfree(tmp);
    PP_HIP(e);
}

// values stably sorted by their keys (keys below 2^bits)
void sort_pairs(pp_ctx *ctx, const uint32_t *kin, uint32_t *kout, const uint32_t *vin, uint32_t *vout, uint32_t n,
                int bits) {
    size_t bytes = 0;
    PP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, bits, ctx->stream));
    unsigned char *tmp = dalloc<unsigned char>(std::max<size_t>(bytes, 1));
    const hipError_t e =
        hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, (int)n, 0, bits, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    dfree(tmp);
    PP_HIP(e);
}

int bits_for(uint32_t m) {  // bits to hold 0..m
    int b = 1;
    while (b < 32 && (m >> b)) ++b;
    return b;
}

// ---- kernels (one thread per row unless stated)

// Reductions over a workgroup of kB threads: the wavefronts combine by
// shuffles, then the workgroup through LDS, so one atomic per workgroup
// reaches the global word (per-row or per-wavefront atomics on one word
// queue behind each other).
template <class T, class Op>
__device__ T block_reduce(T v, Op op, T *lds) {
    for (int o = 32; o > 0; o >>= 1) v = op(v, (T)__shfl_xor(v, o));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < kB / 64; ++k) v = op(v, lds[k]);
    return v;  // complete in thread 0
}

// fail bits: 1 columns not strictly ascending, 2 pattern not symmetric,
// 4 no diagonal, 8 a row longer than kMaxRow
__global__ void k_sd_check(uint32_t n, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci, int *fail) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        const uint32_t k0 = rp[i], k1 = rp[i + 1];
        int f = k1 - k0 > kMaxRow ? 8 : 0;
        bool diag = false;
        for (uint32_t k = k0; k < k1; ++k) {
            const uint32_t c = ci[k];
            if (k > k0 && !(ci[k - 1] < c)) f |= 1;
            if (c == i) {
                diag = true;
                continue;
            }
            uint32_t lo = rp[c], hi = rp[c + 1];
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2;
                if (ci[mid] < i) lo = mid + 1;
                else hi = mid;
            }
            if (lo == rp[c + 1] || ci[lo] != i) f |= 2;
        }
        if (!diag) f |= 4;
        if (f) atomicOr(fail, f);
    }
}

__global__ void k_sd_fill(uint32_t n, int *a, int v) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) a[i] = v;
}

__global__ void k_sd_iota(uint32_t n, uint32_t *a) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) a[i] = i;
}

// The greedy colouring in one launch, sync-free: one thread per row (no
// grid stride, so a row waits only on rows of its own or earlier
// workgroups, which the in-order dispatch has started), each reading its
// lower neighbours' colours as they appear (agent-scope atomics, visible
// across the XCDs' caches) and taking the smallest colour none of them
// has.  fail bit 16: all 64 colours taken; bit 128: a wait outlasted the
// deadline (the host path then colours).
__global__ __launch_bounds__(kB) void k_sd_colour(uint32_t n, const uint32_t *__restrict__ rp,
                                                  const uint32_t *__restrict__ ci, int *colour, long long deadline,
                                                  int *fail) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    bool done = i >= n;
    uint32_t e = 0, e1 = 0;
    unsigned long long m = 0;
    if (!done) {
        e = rp[i];
        e1 = rp[i + 1];
    }
    const long long t0 = wall_clock64();
    for (unsigned iter = 0;; ++iter) {
        if (!done) {
            while (e < e1) {
                const uint32_t j = ci[e];
                if (j < i) {
                    const int c = __hip_atomic_load(&colour[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (c < 0) break;
                    m |= 1ull << c;
                }
                ++e;
            }
            if (e == e1) {
                int c = 0;
                if (m == ~0ull) atomicOr(fail, 16);
                else c = __ffsll((long long)~m) - 1;
                __hip_atomic_store(&colour[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                done = true;
            }
        }
        if (__all(done)) break;
        if (wall_clock64() - t0 > deadline) {
            if (!done) atomicOr(fail, 128);
            break;
        }
        // another wave already gave up: its dependents would wait in vain
        if ((iter & 63) == 63 && (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 128)) break;
        __builtin_amdgcn_s_sleep(1);
    }
}

// one round of dependency levels: level(i) = 1 + the largest level of the
// rows it depends on (upper: the columns above i, else below)
template <bool UPPER>
__global__ void k_sd_level(uint32_t n, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci, int *lev) {
    volatile int *vl = lev;
    for (uint32_t t = blockIdx.x * kB + threadIdx.x; t < n; t += gridDim.x * kB) {
        const uint32_t i = UPPER ? n - 1 - t : t;  // rows in their dependency direction
        if (vl[i] >= 0) continue;
        int l = 0;
        bool ready = true;
        for (uint32_t k = rp[i]; k < rp[i + 1] && ready; ++k) {
            const uint32_t j = ci[k];
            if (UPPER ? j <= i : j >= i) continue;
            const int c = vl[j];
            if (c < 0) ready = false;
            else l = max(l, c + 1);
        }
        if (ready) vl[i] = l;
    }
}

// flag: a negative entry remains; mx: the largest entry
__global__ void k_sd_scan_state(uint32_t n, const int *__restrict__ a, int *flag, int *mx) {
    __shared__ int lds[kB / 64];
    int m = -1, neg = 0;
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        const int v = a[i];
        neg |= v < 0;
        m = max(m, v);
    }
    m = block_reduce(m, [](int x, int y) { return max(x, y); }, lds);
    neg = block_reduce(neg, [](int x, int y) { return x | y; }, lds);
    if (threadIdx.x == 0) {
        if (neg) atomicOr(flag, 1);
        atomicMax(mx, m);
    }
}

__global__ void k_sd_inverse(uint32_t n, const uint32_t *__restrict__ order, uint32_t *inv) {
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) inv[order[k]] = k;
}

// len[k] = the length of row order[k] (order null: of row k), len[n] = 0
__global__ void k_sd_row_len(uint32_t n, const uint32_t *__restrict__ order, const uint32_t *__restrict__ rp,
                             uint32_t *len) {
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k <= n; k += gridDim.x * kB) {
        if (k == n) {
            len[n] = 0;
            continue;
        }
        const uint32_t i = order ? order[k] : k;
        len[k] = rp[i + 1] - rp[i];
    }
}

// len[k] = cnt[perm[k]], len[n] = 0
__global__ void k_sd_perm_len(uint32_t n, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ cnt,
                              uint32_t *len) {
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k <= n; k += gridDim.x * kB) len[k] = k == n ? 0 : cnt[perm[k]];
}

// P A P^T: row k = row order[k], columns renumbered through inv and sorted
// (values and source entries along); src: the entry of A each entry holds
__global__ void k_sd_permute(uint32_t n, const uint32_t *__restrict__ order, const uint32_t *__restrict__ inv,
                             const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                             const double *__restrict__ v, const uint32_t *__restrict__ prp, uint32_t *pci,
                             double *pv, uint32_t *src) {
    constexpr uint32_t kReg = 8;  // rows this short are sorted in registers
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        const uint32_t i = order[k], a0 = rp[i], len = rp[i + 1] - a0, p0 = prp[k];
        if (len <= kReg) {
            uint32_t c[kReg], e[kReg];
#pragma unroll
            for (uint32_t t = 0; t < kReg; ++t) {
                c[t] = t < len ? inv[ci[a0 + t]] : UINT32_MAX;
                e[t] = a0 + t;
            }
#pragma unroll
            for (uint32_t t = 1; t < kReg; ++t)  // insertion sort over a fixed network of compares
#pragma unroll
                for (uint32_t q = t; q > 0; --q)
                    if (c[q - 1] > c[q]) {
                        const uint32_t tc = c[q - 1], te = e[q - 1];
                        c[q - 1] = c[q];
                        e[q - 1] = e[q];
                        c[q] = tc;
                        e[q] = te;
                    }
#pragma unroll
            for (uint32_t t = 0; t < kReg; ++t)
                if (t < len) {
                    pci[p0 + t] = c[t];
                    pv[p0 + t] = v[e[t]];
                    src[p0 + t] = e[t];
                }
            continue;
        }
        for (uint32_t t = 0; t < len; ++t) {  // insertion into the sorted prefix
            const uint32_t c = inv[ci[a0 + t]];
            uint32_t q = t;
            while (q > 0 && pci[p0 + q - 1] > c) {
                pci[p0 + q] = pci[p0 + q - 1];
                pv[p0 + q] = pv[p0 + q - 1];
                src[p0 + q] = src[p0 + q - 1];
                --q;
            }
            pci[p0 + q] = c;
            pv[p0 + q] = v[a0 + t];
            src[p0 + q] = a0 + t;
        }
    }
}

// cnt[i] = entries of row i left of (UPPER: right of) the diagonal, cnt[n] = 0
template <bool UPPER>
__global__ void k_sd_count_side(uint32_t n, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                                uint32_t *cnt) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i <= n; i += gridDim.x * kB) {
        if (i == n) {
            cnt[n] = 0;
            continue;
        }
        uint32_t c = 0;
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) c += UPPER ? ci[k] > i : ci[k] < i;
        cnt[i] = c;
    }
}

// L's strict lower pattern (columns, source entries) and each row's diagonal entry
__global__ void k_sd_lower(uint32_t n, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                           const uint32_t *__restrict__ lrp, uint32_t *lci, uint32_t *lsrc, uint32_t *dsrc) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        uint32_t q = lrp[i];
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const uint32_t c = ci[k];
            if (c < i) {
                lci[q] = c;
                lsrc[q++] = k;
            } else if (c == i) {
                dsrc[i] = k;
            }
        }
    }
}

// forward sweep in level order: position k holds L's row perm[k]; map: the
// factor entry (index into L's strict lower part) of every slot
__global__ void k_sd_gather_lower(uint32_t n, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ lrp,
                                  const uint32_t *__restrict__ lci, const uint32_t *__restrict__ rp, uint32_t *ci,
                                  uint32_t *map) {
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        const uint32_t i = perm[k], l0 = lrp[i], q0 = rp[k], len = lrp[i + 1] - l0;
        for (uint32_t e = 0; e < len; ++e) {
            ci[q0 + e] = lci[l0 + e];
            map[q0 + e] = l0 + e;
        }
    }
}

// backward sweep in level order: position k holds row j = perm[k] of L^T,
// i.e. the columns i > j of the matrix's row j (ascending), each value
// l_ij found in row i of L by binary search (fail bit 32: absent)
__global__ void k_sd_gather_upper(uint32_t n, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ prp,
                                  const uint32_t *__restrict__ pci, const uint32_t *__restrict__ lrp,
                                  const uint32_t *__restrict__ lci, const uint32_t *__restrict__ rp, uint32_t *ci,
                                  uint32_t *map, int *fail) {
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        const uint32_t j = perm[k];
        uint32_t q = rp[k];
        for (uint32_t e = prp[j]; e < prp[j + 1]; ++e) {
            const uint32_t i = pci[e];
            if (i <= j) continue;
            uint32_t lo = lrp[i], hi = lrp[i + 1];
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2;
                if (lci[mid] < j) lo = mid + 1;
                else hi = mid;
            }
            if (lo == lrp[i + 1] || lci[lo] != j) {
                atomicOr(fail, 32);
                lo = lrp[i];
            }
            ci[q] = i;
            map[q++] = lo;
        }
    }
}

// ILU(0)'s backward sweep in level order: position k holds U's row j =
// perm[k], the columns i > j of the matrix's row j with their own entries
// (the factor is computed in place on the matrix's pattern)
__global__ void k_sd_gather_upper_lu(uint32_t n, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ prp,
                                     const uint32_t *__restrict__ pci, const uint32_t *__restrict__ rp, uint32_t *ci,
                                     uint32_t *map) {
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < n; k += gridDim.x * kB) {
        const uint32_t j = perm[k];
        uint32_t q = rp[k];
        for (uint32_t e = prp[j]; e < prp[j + 1]; ++e)
            if (pci[e] > j) {
                ci[q] = pci[e];
                map[q++] = e;
            }
    }
}

// map[q] = via[map[q]]
__global__ void k_sd_compose(uint32_t cnt, const uint32_t *__restrict__ via, uint32_t *map) {
    for (uint32_t q = blockIdx.x * kB + threadIdx.x; q < cnt; q += gridDim.x * kB) map[q] = via[map[q]];
}

// the pivots' positions in the factor array: nnz + row
__global__ void k_sd_diag_map(uint32_t n, uint32_t nnz, uint32_t *map) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) map[i] = nnz + i;
}

// level bounds from the sorted keys: start[l] = the first position of a key
// >= l (start[nlev] = n), written where the key steps up
__global__ void k_sd_bounds(uint32_t n, uint32_t nlev, const uint32_t *__restrict__ key, uint32_t *start) {
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k <= n; k += gridDim.x * kB) {
        const uint32_t lo = k == 0 ? 0 : key[k - 1] + 1, hi = k == n ? nlev : key[k];
        for (uint32_t l = lo; l <= hi && l <= nlev; ++l) start[l] = k;
    }
}

// the first row whose first entry lies left of the diagonal (the colour split)
__global__ void k_sd_split(uint32_t n, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci, uint32_t *s) {
    __shared__ uint32_t lds[kB / 64];
    uint32_t m = UINT32_MAX;
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
        if (rp[i] < rp[i + 1] && ci[rp[i]] < i) m = min(m, i);
    m = block_reduce(m, [](uint32_t x, uint32_t y) { return min(x, y); }, lds);
    if (threadIdx.x == 0 && m != UINT32_MAX) atomicMin(s, m);
}

// offsets of one sliced-column sweep level (the host's upload_sweep loop):
// row t = original row row0 + t, entry j at [j * nr + t]; fail bit 64: an
// offset beyond 16 bits
__global__ void k_sd_ell_level(uint32_t nr, uint32_t w, const uint32_t *__restrict__ rp,
                               const uint32_t *__restrict__ ci, int64_t row0, int64_t cbase, int16_t *off,
                               int *fail) {
    for (uint32_t t = blockIdx.x * kB + threadIdx.x; t < nr; t += gridDim.x * kB) {
        const uint32_t k0 = rp[t], k1 = rp[t + 1];
        int16_t last = 0;
        bool bad = false;
        for (uint32_t j = 0; j < w; ++j) {
            if (k0 + j < k1) {
                const int64_t o = (int64_t)ci[k0 + j] - (row0 + (int64_t)t) - cbase;
                bad |= o < INT16_MIN || o > INT16_MAX;
                last = (int16_t)o;
            }
            off[(size_t)j * nr + t] = last;
        }
        if (bad) atomicOr(fail, 64);
    }
}

// 16-bit two-base column offsets of one row block per workgroup (the host's
// two_base_offsets): the block's columns sorted in LDS (bitonic), one
// window when they span at most kOffMask, else a split at the widest gap
// between consecutive distinct columns (the first of equal gaps), each
// window within kOffMask or the block keeps 32-bit columns (kNoBase)
constexpr int kTB = 256;
__global__ __launch_bounds__(kTB) void k_two_base(const uint32_t *__restrict__ blocks, const uint32_t *__restrict__ rp,
                                                  const uint32_t *__restrict__ ci, uint16_t *off16, uint32_t *base,
                                                  unsigned long long *n16) {
    __shared__ uint32_t u[kSpmvNnzBlock];
    __shared__ unsigned long long best;
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    const uint32_t e0 = rp[blocks[b]], e1 = rp[blocks[b + 1]], cnt = e1 - e0;
    if (cnt == 0 || cnt > kSpmvNnzBlock) {
        if (tid == 0) base[2 * b] = base[2 * b + 1] = kNoBase;
        return;
    }
    for (uint32_t t = tid; t < kSpmvNnzBlock; t += kTB) u[t] = t < cnt ? ci[e0 + t] : UINT32_MAX;
    if (tid == 0) best = 0;
    __syncthreads();
    for (uint32_t k = 2; k <= kSpmvNnzBlock; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = tid; t < kSpmvNnzBlock; t += kTB) {
                const uint32_t p = t ^ j;
                if (p > t) {
                    const uint32_t a = u[t], c = u[p];
                    if ((a > c) == ((t & k) == 0)) {
                        u[t] = c;
                        u[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    const uint32_t lo = u[0], top = u[cnt - 1];
    bool two = false, ok = true;
    uint32_t hi = lo;
    if (top - lo > kOffMask) {
        unsigned long long mine = 0;  // (gap, first position) as the largest key
        for (uint32_t t = tid + 1; t < cnt; t += kTB) {
            const uint32_t g = u[t] - u[t - 1];
            if (g) mine = max(mine, ((unsigned long long)g << 32) | (0xffffffffu - t));
        }
        atomicMax(&best, mine);
        __syncthreads();
        const uint32_t ts = 0xffffffffu - (uint32_t)(best & 0xffffffffu);
        hi = u[ts];
        two = true;
        ok = u[ts - 1] - lo <= kOffMask && top - hi <= kOffMask;
    }
    if (!ok) {
        if (tid == 0) base[2 * b] = base[2 * b + 1] = kNoBase;
        return;
    }
    for (uint32_t e = e0 + tid; e < e1; e += kTB) {
        const uint32_t c = ci[e];
        off16[e] = two && c >= hi ? (uint16_t)(kHiBase | (c - hi)) : (uint16_t)(c - lo);
    }
    if (tid == 0) {
        base[2 * b] = lo;
        base[2 * b + 1] = hi;
        atomicAdd(n16, (unsigned long long)cnt);
    }
}

// statistics of one level (positions [k0, k1)): st[0] the longest row,
// st[1] the shortest, st[2] set when the rows are not consecutive in the
// original numbering
__global__ void k_sd_level_stats(uint32_t k0, uint32_t k1, const uint32_t *__restrict__ perm,
                                 const uint32_t *__restrict__ rp, uint32_t *st) {
    uint32_t mx = 0, mn = UINT32_MAX, gap = 0;
    for (uint32_t k = k0 + blockIdx.x * kB + threadIdx.x; k < k1; k += gridDim.x * kB) {
        const uint32_t len = rp[k + 1] - rp[k];
        mx = max(mx, len);
        mn = min(mn, len);
        if (k > k0 && perm[k] != perm[k - 1] + 1) gap = 1;
    }
    __shared__ uint32_t lds[kB / 64];
    mx = block_reduce(mx, [](uint32_t x, uint32_t y) { return max(x, y); }, lds);
    mn = block_reduce(mn, [](uint32_t x, uint32_t y) { return min(x, y); }, lds);
    gap = block_reduce(gap, [](uint32_t x, uint32_t y) { return x | y; }, lds);
    if (threadIdx.x == 0) {
        atomicMax(st, mx);
        atomicMin(st + 1, mn);
        if (gap) atomicOr(st + 2, 1u);
    }
}

// per level l: heads[3l] = perm[lp[l]], heads[3l+1] = rp[lp[l]],
// heads[3l+2] = ci[rp[lp[l]]] (when the level has entries); heads[3 nlev] = rp[n]
__global__ void k_sd_level_heads(uint32_t nlev, uint32_t n, const uint32_t *__restrict__ lp,
                                 const uint32_t *__restrict__ perm, const uint32_t *__restrict__ rp,
                                 const uint32_t *__restrict__ ci, uint32_t *heads) {
    const uint32_t l = blockIdx.x * kB + threadIdx.x;
    if (l > nlev) return;
    if (l == nlev) {
        heads[3 * nlev] = rp[n];
        return;
    }
    const uint32_t k = lp[l];
    heads[3 * l] = k < n ? perm[k] : 0;
    heads[3 * l + 1] = rp[k];
    heads[3 * l + 2] = k < lp[l + 1] && rp[k] < rp[lp[l + 1]] ? ci[rp[k]] : 0;
}

// descriptors of nb row blocks (bounds blk [nb + 1]): first position, first
// entry, the 16-bit bases (kNoBase without them); the last closes the final block
__global__ void k_sd_descs(uint32_t nb, const uint32_t *__restrict__ blk, const uint32_t *__restrict__ rp,
                           const uint32_t *__restrict__ base, BlockDesc *desc) {
    for (uint32_t j = blockIdx.x * kB + threadIdx.x; j <= nb; j += gridDim.x * kB) {
        const bool has = j < nb && base;
        desc[j] = {blk[j], rp[blk[j]], has ? base[2 * j] : kNoBase, has ? base[2 * j + 1] : kNoBase};
    }
}

// flag when a[0, cnt) and b[0, cnt) differ
__global__ void k_sd_differ(uint32_t cnt, const uint32_t *__restrict__ a, const uint32_t *__restrict__ b, int *flag) {
    for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < cnt; k += gridDim.x * kB)
        if (a[k] != b[k]) atomicOr(flag, 1);
}

// ---- owned device arrays, freed unless handed over
// The setup's temporaries: returned to the context's pool at the end
// (kept ones leave with their new owner)
struct Arena {
    pp_ctx *ctx;
    std::vector<std::pair<void *, size_t>> p;
    explicit Arena(pp_ctx *c) : ctx(c) {}
    template <class T>
    T *get(size_t n) {
        n = std::max<size_t>(n, 1);
        T *a = dalloc<T>(n);
        p.emplace_back(a, n * sizeof(T));
        return a;
    }
    template <class T>
    T *keep(T *a) {  // no longer freed here
        for (auto &q : p)
            if (q.first == a) q.first = nullptr;
        return a;
    }
    ~Arena() {
        for (auto &q : p) pool_put(ctx, q.first, q.second);
    }
};

int read_int(pp_ctx *ctx, const int *d) {
    int v = 0;
    PP_HIP(hipMemcpyAsync(&v, d, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return v;
}

// rounds of k (colouring or levels) until no entry is negative; false when
// they do not finish within n rounds; *mx: the largest entry
template <class Round>
bool rounds(pp_ctx *ctx, uint32_t n, int *a, int *st, Round round, int &mx) {
    // a dependency chain needs one round per link at worst: long chains
    // (deep natural-order graphs) stay with the host's serial pass
    for (uint32_t done = 0; done <= std::min<uint32_t>(n, 4096);) {
        const uint32_t batch = done < 8 ? 2 : 16;
        for (uint32_t b = 0; b < batch; ++b) round();
        done += batch;
        PP_HIP(hipMemsetAsync(st, 0, sizeof(int), ctx->stream));
        PP_HIP(hipMemsetAsync(st + 1, 0xff, sizeof(int), ctx->stream));
        k_sd_scan_state<<<grid_of(n), kB, 0, ctx->stream>>>(n, a, st, st + 1);
        PP_HIP(hipGetLastError());
        int h[2];
        PP_HIP(hipMemcpyAsync(h, st, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        if (!h[0]) {
            mx = h[1];
            return true;
        }
    }
    return false;
}

// rows stably sorted by level (lev < nlev): perm and the host level bounds
void level_order(pp_ctx *ctx, Arena &ar, uint32_t n, int *lev, uint32_t nlev, const uint32_t *iota, uint32_t *perm,
                 std::vector<uint32_t> &lptr) {
    uint32_t *kout = ar.get<uint32_t>(n);
    sort_pairs(ctx, (const uint32_t *)lev, kout, iota, perm, n, bits_for(nlev));
    uint32_t *start = ar.get<uint32_t>(nlev + 1);
    k_sd_bounds<<<grid_of(n + 1), kB, 0, ctx->stream>>>(n, nlev, kout, start);
    PP_HIP(hipGetLastError());
    lptr.assign(nlev + 1, 0);
    PP_HIP(hipMemcpyAsync(lptr.data(), start, (nlev + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
}

// The level-ordered sweep d from its device arrays (perm, rp, ci taken
// over; values gathered through map from the factor array): the host
// upload_sweep's layout, its per-level facts reduced on the device (the
// row pointers cross to the host only for a wide level whose rows do not
// all fit the uniform row blocks).
void finish_sweep(pp_ctx *ctx, Arena &ar, DevSweep<double> &d, uint32_t n, const std::vector<uint32_t> &lp,
                  uint32_t *perm, uint32_t *rp, uint32_t *ci, const uint32_t *map, const double *fval, int *fail,
                  bool with_diag) {
    d.free_all();
    d.n = n;
    d.nlev = (uint32_t)lp.size() - 1;
    d.level_ptr = lp;
    const uint32_t nlev = d.nlev;
    d.perm = ar.keep(perm);
    d.rp = ar.keep(rp);
    d.ci = ar.keep(ci);
    // per-level statistics and heads
    uint32_t *st = ar.get<uint32_t>(3 * (size_t)nlev), *dlp = ar.get<uint32_t>(nlev + 1);
    uint32_t *heads = ar.get<uint32_t>(3 * (size_t)nlev + 1);
    std::vector<uint32_t> init(3 * (size_t)nlev);
    for (uint32_t l = 0; l < nlev; ++l) {
        init[3 * l] = 0;
        init[3 * l + 1] = UINT32_MAX;
        init[3 * l + 2] = 0;
    }
    PP_HIP(hipMemcpyAsync(st, init.data(), init.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    PP_HIP(hipMemcpyAsync(dlp, lp.data(), lp.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    for (uint32_t l = 0; l < nlev; ++l)
        if (lp[l + 1] > lp[l])
            k_sd_level_stats<<<std::min<uint32_t>((lp[l + 1] - lp[l] + kB - 1) / kB, 2048), kB, 0, ctx->stream>>>(
                lp[l], lp[l + 1], perm, rp, st + 3 * l);
    k_sd_level_heads<<<(nlev + kB) / kB, kB, 0, ctx->stream>>>(nlev, n, dlp, perm, rp, ci, heads);
    PP_HIP(hipGetLastError());
    std::vector<uint32_t> hst(3 * (size_t)nlev), hh(3 * (size_t)nlev + 1);
    PP_HIP(hipMemcpyAsync(hst.data(), st, hst.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipMemcpyAsync(hh.data(), heads, hh.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));  // init leaves scope
    d.nnz = hh[3 * nlev];
    auto level_entries = [&](uint32_t l) { return (l + 1 < nlev ? hh[3 * (l + 1) + 1] : d.nnz) - hh[3 * l + 1]; };
    d.val = dalloc<double>(std::max<uint32_t>(d.nnz, 1));
    launch_gather_map(ctx, d.nnz, map, fval, PP_F64, d.val);
    d.max_level_rows = 0;
    for (uint32_t l = 0; l < nlev; ++l) d.max_level_rows = std::max(d.max_level_rows, lp[l + 1] - lp[l]);
    d.level_off.assign(nlev, DevSweep<double>::kNoOffset);
    for (uint32_t l = 0; l < nlev; ++l)
        if (lp[l + 1] > lp[l] && !hst[3 * l + 2]) d.level_off[l] = (int64_t)hh[3 * l] - (int64_t)lp[l];
    // row blocks of the wide levels (the host's rule; uniform when every row
    // of the level fits kSpmvNnzBlock / kSpmvRowsBlock entries)
    std::vector<uint32_t> blk, hrp;
    d.level_blk.assign(2 * (size_t)nlev, 0);
    for (uint32_t l = 0; l < nlev; ++l) {
        if (lp[l + 1] - lp[l] < kStreamLevelRows) continue;
        uint32_t k = lp[l];
        if (blk.empty() || blk.back() != k) blk.push_back(k);
        d.level_blk[2 * l] = (uint32_t)blk.size() - 1;
        if (hst[3 * l] <= kSpmvNnzBlock / kSpmvRowsBlock) {
            for (k += kSpmvRowsBlock; k < lp[l + 1]; k += kSpmvRowsBlock) blk.push_back(k);
            blk.push_back(lp[l + 1]);
        } else {
            if (hrp.empty()) {
                hrp.resize((size_t)n + 1);
                PP_HIP(hipMemcpy(hrp.data(), rp, ((size_t)n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
            }
            while (k < lp[l + 1]) {
                const uint32_t start = k, e0 = hrp[k];
                if (hrp[k + 1] - e0 > kSpmvNnzBlock) {
                    ++k;
                } else {
                    while (k < lp[l + 1] && k - start < kSpmvRowsBlock && hrp[k + 1] - e0 <= kSpmvNnzBlock) ++k;
                }
                blk.push_back(k);
            }
        }
        d.level_blk[2 * l + 1] = (uint32_t)blk.size() - 1;
    }
    if (blk.size() > 1) {
        d.blocks = dalloc<uint32_t>(blk.size());
        PP_HIP(hipMemcpyAsync(d.blocks, blk.data(), blk.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              ctx->stream));
        d.nblk = blk.size() - 1;
        std::vector<uint32_t> base;
        uint16_t *off16 = nullptr;
        uint32_t *dbase = nullptr;
        d.nnz16 = two_base_device(ctx, (uint32_t)d.nblk, d.blocks, d.rp, d.ci, d.nnz, off16, dbase, base);
        if (d.nnz16) {
            d.ci16 = off16;
            d.blk_base = dbase;
        }
        d.desc = dalloc<BlockDesc>(blk.size());
        k_sd_descs<<<grid_of(blk.size()), kB, 0, ctx->stream>>>((uint32_t)d.nblk, d.blocks, d.rp, d.blk_base, d.desc);
        PP_HIP(hipGetLastError());
        PP_HIP(hipStreamSynchronize(ctx->stream));  // blk leaves scope
    }
    for (uint32_t l = 0; l < nlev; ++l)
        if (d.level_off[l] == DevSweep<double>::kNoOffset) d.perm_rows += lp[l + 1] - lp[l];
    // the pivots (original numbering), gathered by the caller; a unit
    // diagonal keeps none
    if (with_diag) d.diag = dalloc<double>(n);
    // sliced-column levels (factors large enough for the setup to time them)
    d.ell.assign(nlev, DevSweep<double>::EllLevel());
    for (uint32_t l = 0; l < nlev && n >= kTuneSweepRows; ++l) {
        const uint32_t k0 = lp[l], k1 = lp[l + 1], nr = k1 - k0;
        if (nr < kStreamLevelRows || d.level_off[l] == DevSweep<double>::kNoOffset) continue;
        const int64_t row0 = (int64_t)k0 + d.level_off[l];
        const uint32_t w = hst[3 * l], entries = level_entries(l);
        if (hst[3 * l + 1] == 0 || w > 16 || (double)w * nr > 1.25 * (double)entries) continue;
        const int64_t cbase = (int64_t)hh[3 * l + 2] - row0;
        PP_HIP(hipMemsetAsync(fail, 0, sizeof(int), ctx->stream));
        int16_t *off = dalloc<int16_t>((size_t)w * nr);
        k_sd_ell_level<<<grid_of(nr), kB, 0, ctx->stream>>>(nr, w, d.rp + k0, d.ci, row0, cbase, off, fail);
        PP_HIP(hipGetLastError());
        if (read_int(ctx, fail)) {
            dfree(off);
            continue;
        }
        auto &E = d.ell[l];
        E.nrows = nr;
        E.entries = entries;
        E.width = w;
        E.row0 = row0;
        E.cbase = cbase;
        E.off = off;
        E.val = dalloc<double>((size_t)w * nr);
        E.ok = true;
    }
}

}  // namespace

uint64_t two_base_device(pp_ctx *ctx, uint32_t nblk, const uint32_t *blocks, const uint32_t *rp, const uint32_t *ci,
                         uint32_t nnz, uint16_t *&off16, uint32_t *&dbase, std::vector<uint32_t> &base) {
    off16 = nullptr;
    dbase = nullptr;
    base.assign(2 * (size_t)nblk, kNoBase);
    if (!nblk) return 0;
    unsigned long long *cnt = nullptr;
    uint64_t n16 = 0;
    try {
        off16 = dalloc<uint16_t>(std::max<uint32_t>(nnz, 1));
        dbase = dalloc<uint32_t>(2 * (size_t)nblk);
        cnt = dalloc<unsigned long long>(1);
        PP_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), ctx->stream));
        k_two_base<<<nblk, kTB, 0, ctx->stream>>>(blocks, rp, ci, off16, dbase, cnt);
        PP_HIP(hipGetLastError());
        PP_HIP(hipMemcpyAsync(base.data(), dbase, base.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipMemcpyAsync(&n16, cnt, sizeof(n16), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        dfree(off16);
        dfree(dbase);
        dfree(cnt);
        off16 = nullptr;
        dbase = nullptr;
        throw;
    }
    dfree(cnt);
    if (!n16) {  // no block fits: the 32-bit columns only
        dfree(off16);
        dfree(dbase);
        off16 = nullptr;
        dbase = nullptr;
    }
    return n16;
}

bool setup_mc_device(pp_ctx *ctx, pp_precond *M) {
    const bool ic = M->opts.kind == PP_PC_IC0;
    pp_csr *A = M->A;
    const uint32_t n = A->nrow;
    if (n != A->ncol || n == 0 || A->h_rp.size() != (size_t)n + 1) return false;
    Arena ar(ctx);
    int *st = ar.get<int>(2);
    int *fail = ar.get<int>(1);
    PP_HIP(hipMemsetAsync(fail, 0, sizeof(int), ctx->stream));
    k_sd_check<<<grid_of(n), kB, 0, ctx->stream>>>(n, A->rp, A->ci, fail);
    PP_HIP(hipGetLastError());
    if (read_int(ctx, fail)) return false;  // unsorted, unsymmetric, no diagonal or long rows: the host path
    auto t0 = std::chrono::steady_clock::now();
    // colouring and the colour order
    int *colour = ar.get<int>(n);
    k_sd_fill<<<grid_of(n), kB, 0, ctx->stream>>>(n, colour, -1);
    k_sd_colour<<<(n + kB - 1) / kB, kB, 0, ctx->stream>>>(n, A->rp, A->ci, colour, ctx->spin_deadline, fail);
    PP_HIP(hipGetLastError());
    // more than 64 colours (the host's other colouring) or a wait that gave up
    if (read_int(ctx, fail)) return false;
    int maxc = 0;
    PP_HIP(hipMemsetAsync(st, 0, sizeof(int), ctx->stream));
    PP_HIP(hipMemsetAsync(st + 1, 0xff, sizeof(int), ctx->stream));
    k_sd_scan_state<<<grid_of(n), kB, 0, ctx->stream>>>(n, colour, st, st + 1);
    PP_HIP(hipGetLastError());
    {
        int h[2];
        PP_HIP(hipMemcpyAsync(h, st, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        if (h[0]) return false;  // a row left without a colour
        maxc = h[1];
    }
    const uint32_t ncolors = (uint32_t)maxc + 1;
    uint32_t *iota = ar.get<uint32_t>(n);
    k_sd_iota<<<grid_of(n), kB, 0, ctx->stream>>>(n, iota);
    uint32_t *order = ar.get<uint32_t>(n), *inv = ar.get<uint32_t>(n), *kout = ar.get<uint32_t>(n);
    sort_pairs(ctx, (const uint32_t *)colour, kout, iota, order, n, bits_for(ncolors));
    k_sd_inverse<<<grid_of(n), kB, 0, ctx->stream>>>(n, order, inv);
    // P A P^T
    uint32_t *len = ar.get<uint32_t>(n + 1), *prp = ar.get<uint32_t>(n + 1);
    k_sd_row_len<<<grid_of(n + 1), kB, 0, ctx->stream>>>(n, order, A->rp, len);
    PP_HIP(hipGetLastError());
    exclusive_scan(ctx, len, prp, n + 1);
    const uint32_t nnz = A->nnz;
    uint32_t *pci = ar.get<uint32_t>(nnz), *asrc = ar.get<uint32_t>(nnz);
    double *pv = ar.get<double>(nnz);
    k_sd_permute<<<grid_of(n), kB, 0, ctx->stream>>>(n, order, inv, A->rp, A->ci, A->v64, prp, pci, pv, asrc);
    PP_HIP(hipGetLastError());
    auto ms = [&] {
        PP_HIP(hipStreamSynchronize(ctx->stream));
        const auto now = std::chrono::steady_clock::now();
        const float v = (float)std::chrono::duration<double, std::milli>(now - t0).count();
        t0 = now;
        return v;
    };
    const float ms_order = ms();
    // L's strict lower pattern, its levels (the forward sweep's) and the
    // backward sweep's levels
    uint32_t *lrp = ar.get<uint32_t>(n + 1);
    k_sd_count_side<false><<<grid_of(n + 1), kB, 0, ctx->stream>>>(n, prp, pci, len);
    exclusive_scan(ctx, len, lrp, n + 1);
    uint32_t nl = 0;
    PP_HIP(hipMemcpyAsync(&nl, lrp + n, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    uint32_t *lci = ar.get<uint32_t>(nl), *lsrc = ar.get<uint32_t>(nl), *dsrc = ar.get<uint32_t>(n);
    k_sd_lower<<<grid_of(n), kB, 0, ctx->stream>>>(n, prp, pci, lrp, lci, lsrc, dsrc);
    PP_HIP(hipGetLastError());
    int *lev_lo = ar.get<int>(n), *lev_up = ar.get<int>(n);
    k_sd_fill<<<grid_of(n), kB, 0, ctx->stream>>>(n, lev_lo, -1);
    k_sd_fill<<<grid_of(n), kB, 0, ctx->stream>>>(n, lev_up, -1);
    int maxlo = 0, maxup = 0;
    if (!rounds(ctx, n, lev_lo, st, [&] { k_sd_level<false><<<grid_of(n), kB, 0, ctx->stream>>>(n, lrp, lci, lev_lo); },
                maxlo) ||
        !rounds(ctx, n, lev_up, st, [&] { k_sd_level<true><<<grid_of(n), kB, 0, ctx->stream>>>(n, prp, pci, lev_up); },
                maxup))
        return false;
    const uint32_t nlev_lo = (uint32_t)maxlo + 1, nlev_up = (uint32_t)maxup + 1;
    uint32_t *perm_lo = ar.get<uint32_t>(n), *perm_up = ar.get<uint32_t>(n);
    std::vector<uint32_t> lp_lo, lp_up;
    level_order(ctx, ar, n, lev_lo, nlev_lo, iota, perm_lo, lp_lo);
    level_order(ctx, ar, n, lev_up, nlev_up, iota, perm_up, lp_up);
    // the numeric factor, level by level (the forward sweep's levels are
    // L's): IC(0) into L's strict lower part then the pivots l_ii; ILU(0)
    // in place on the matrix's pattern then the inverse pivots 1/u_ii
    const uint32_t nf = ic ? nl : nnz;  // factor entries before the pivots
    double *fval = ar.get<double>((size_t)nf + n);
    uint32_t *rows = ar.get<uint32_t>(n);
    PP_HIP(hipMemcpyAsync(rows, perm_lo, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
    const int big = INT32_MAX;
    int *ffail = ar.get<int>(1);
    PP_HIP(hipMemcpyAsync(ffail, &big, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    for (uint32_t l = 0; l < nlev_lo; ++l) {
        if (ic)
            launch_ic0_level(ctx, rows + lp_lo[l], lp_lo[l + 1] - lp_lo[l], lrp, lci, lsrc, dsrc, pv, fval,
                             fval + nf, ffail);
        else
            launch_ilu0_level(ctx, rows + lp_lo[l], lp_lo[l + 1] - lp_lo[l], prp, pci, dsrc, pv, fval, fval + nf,
                              ffail);
    }
    if (read_int(ctx, ffail) != big) return false;  // a failing pivot: the host path reports it
    const float ms_factor = ms();
    // the sweeps
    uint32_t *srp_lo = ar.get<uint32_t>(n + 1), *sci_lo = ar.get<uint32_t>(nl), *map_lo = ar.get<uint32_t>(nl);
    k_sd_row_len<<<grid_of(n + 1), kB, 0, ctx->stream>>>(n, perm_lo, lrp, len);
    exclusive_scan(ctx, len, srp_lo, n + 1);
    k_sd_gather_lower<<<grid_of(n), kB, 0, ctx->stream>>>(n, perm_lo, lrp, lci, srp_lo, sci_lo, map_lo);
    if (!ic)  // ILU(0): L's entries are the matrix's own
        k_sd_compose<<<grid_of(nl), kB, 0, ctx->stream>>>(nl, lsrc, map_lo);
    PP_HIP(hipGetLastError());
    uint32_t *ucnt = ar.get<uint32_t>(n + 1), *srp_up = ar.get<uint32_t>(n + 1);
    uint32_t *sci_up = ar.get<uint32_t>(nl), *map_up = ar.get<uint32_t>(nl);
    k_sd_count_side<true><<<grid_of(n + 1), kB, 0, ctx->stream>>>(n, prp, pci, ucnt);
    k_sd_perm_len<<<grid_of(n + 1), kB, 0, ctx->stream>>>(n, perm_up, ucnt, len);
    PP_HIP(hipGetLastError());
    exclusive_scan(ctx, len, srp_up, n + 1);
    PP_HIP(hipMemsetAsync(fail, 0, sizeof(int), ctx->stream));
    if (ic)
        k_sd_gather_upper<<<grid_of(n), kB, 0, ctx->stream>>>(n, perm_up, prp, pci, lrp, lci, srp_up, sci_up,
                                                              map_up, fail);
    else
        k_sd_gather_upper_lu<<<grid_of(n), kB, 0, ctx->stream>>>(n, perm_up, prp, pci, srp_up, sci_up, map_up);
    PP_HIP(hipGetLastError());
    if (read_int(ctx, fail)) return false;
    uint32_t *dmap = ar.get<uint32_t>(n);
    k_sd_diag_map<<<grid_of(n), kB, 0, ctx->stream>>>(n, nf, dmap);
    PP_HIP(hipGetLastError());
    // everything that can fail has run: hand the arrays to the preconditioner
    // IC(0): forward L (divide by l_ii), backward L^T (divide); ILU(0):
    // forward L (unit), backward U (multiply by 1/u_ii)
    finish_sweep(ctx, ar, M->low64, n, lp_lo, perm_lo, srp_lo, sci_lo, map_lo, fval, fail, ic);
    finish_sweep(ctx, ar, M->up64, n, lp_up, perm_up, srp_up, sci_up, map_up, fval, fail, true);
    for (DevSweep<double> *d : {&M->low64, &M->up64})
        if (d->diag) launch_gather_map(ctx, n, dmap, fval, PP_F64, d->diag);
    PP_HIP(hipStreamSynchronize(ctx->stream));  // ell_fill_rows runs on the null stream
    for (DevSweep<double> *d : {&M->low64, &M->up64})
        for (uint32_t l = 0; l < d->nlev; ++l)
            if (d->ell[l].ok)
                ell_fill_rows<double>(d->ell[l].nrows, d->ell[l].width, d->rp + d->level_ptr[l], d->val,
                                      d->ell[l].val);
    // the forward sweep's last level and the backward sweep's first level
    // holding the same rows share one launch (setup_triangular's rule)
    M->fuse_mid = false;
    M->fuse_rows = 0;
    if (lp_lo.size() >= 2 && lp_up.size() >= 2) {
        const uint32_t a0 = lp_lo[lp_lo.size() - 2], a1 = lp_lo.back(), b0 = lp_up[0], b1 = lp_up[1];
        if (a1 - a0 == b1 - b0) {
            PP_HIP(hipMemsetAsync(fail, 0, sizeof(int), ctx->stream));
            k_sd_differ<<<grid_of(a1 - a0), kB, 0, ctx->stream>>>(a1 - a0, M->low64.perm + a0, M->up64.perm + b0,
                                                                   fail);
            PP_HIP(hipGetLastError());
            if (!read_int(ctx, fail)) {
                M->fuse_mid = true;
                M->fuse_rows = a1 - a0;
            }
        }
    }
    // the host factor's count: IC(0) L^T with its pivots twice, + n; ILU(0) L + U + n
    M->mnnz = ic ? (uint32_t)(2 * ((uint64_t)nl + n) + n) : nnz;
    const float ms_sweeps = ms();
    // P A P^T as the solves' device matrix (its row pointers on the host)
    std::vector<uint32_t> hprp(n + 1);
    PP_HIP(hipMemcpyAsync(hprp.data(), prp, (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    csr_free(M->Aperm);
    M->Aperm = nullptr;
    M->Aperm = csr_from_device(ctx, n, n, ar.keep(prp), ar.keep(pci), ar.keep(pv), std::move(hprp));
    M->Aperm_src_uid = A->uid;
    M->Aperm_version = A->version;
    {
        pp_csr *P = M->Aperm;
        P->tuned32 = false;
        uint32_t *sp = ar.get<uint32_t>(1);
        PP_HIP(hipMemsetAsync(sp, 0xff, sizeof(uint32_t), ctx->stream));
        k_sd_split<<<grid_of(n), kB, 0, ctx->stream>>>(n, P->rp, P->ci, sp);
        uint32_t split = 0;
        PP_HIP(hipMemcpyAsync(&split, sp, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        if (n >= (1u << 14) && !ell_build_device(ctx, P, false, 0, P->ell, 1.1))
            ell_build_device(ctx, P, true, split == UINT32_MAX ? n : split, P->ell, 1.1);
        csr_tune(ctx, P);
    }
    // the re-factorization state: the pattern, levels and maps above
    std::unique_ptr<DevRefactor> R(new DevRefactor);
    R->n = n;
    R->nnz = nf;
    if (ic) {
        R->rp = ar.keep(lrp);
        R->ci = ar.keep(lci);
        R->src = ar.keep(lsrc);
    } else {  // ILU(0) factors on the matrix's own pattern (its own copy: P A P^T keeps its arrays)
        R->rp = dalloc<uint32_t>((size_t)n + 1);
        R->ci = dalloc<uint32_t>(std::max<uint32_t>(nnz, 1));
        PP_HIP(hipMemcpyAsync(R->rp, M->Aperm->rp, ((size_t)n + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                              ctx->stream));
        PP_HIP(hipMemcpyAsync(R->ci, M->Aperm->ci, (size_t)nnz * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                              ctx->stream));
    }
    R->dsrc = ar.keep(dsrc);
    R->val = ar.keep(fval);
    R->rows = ar.keep(rows);
    R->lptr = lp_lo;
    R->lo_map = ar.keep(map_lo);
    R->lo_nnz = nl;
    R->lo_diag = ic;
    if (ic) R->lo_dmap = ar.keep(dmap);
    R->up_map = ar.keep(map_up);
    R->up_nnz = nl;
    R->up_dmap = dalloc<uint32_t>(n);
    PP_HIP(hipMemcpyAsync(R->up_dmap, dmap, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
    R->up_diag = true;
    R->aperm_map = ar.keep(asrc);
    R->aperm_nnz = nnz;
    R->fail = ar.keep(ffail);
    delete M->dref;
    M->dref = R.release();
    // the colour order
    M->order.resize(n);
    PP_HIP(hipMemcpyAsync(M->order.data(), order, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    dfree(M->dorder);
    M->dorder = ar.keep(order);
    M->ncolors = ncolors;
    if (!M->xp64) M->xp64 = dalloc<double>(n);
    PP_HIP(hipStreamSynchronize(ctx->stream));
    M->phase_ms[0] = ms_order;
    M->phase_ms[1] = ms_factor;
    M->phase_ms[2] = ms_sweeps + ms();
    return true;
}

}  // namespace pp
