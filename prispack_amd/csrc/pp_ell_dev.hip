// pp_ell_dev.hip -- the sliced-column copy of a CSR matrix (EllOp), built on
// the device from the matrix's device arrays.
//
// The same form as the host construction (pp_reduced.cpp: ell_build with no
// base and no separate diagonal, ell_build_two_colour), slot for slot and
// pattern number for pattern number: one thread per row writes its 16-bit
// offsets (checking that they fit and, in the two-colour form, that every
// coupling crosses the colours and the diagonal sits where the form expects
// it), the values are gathered by ell_fill_values, and the rows' offset
// tuples are numbered in order of first appearance:
//   1. every row hashes its tuple; the lanes of a wavefront sharing a hash
//      are represented by their lowest lane, which enters the hash into an
//      open-addressing table (CAS) and lowers the slot's first row (atomicMin);
//   2. the host orders the occupied slots by first row (at most 256 kept);
//   3. every row finds its slot, takes its number and compares its tuple with
//      the slot's first row's, so a hash collision fails the build instead of
//      merging two patterns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <vector>

#include "pp_internal.h"

namespace pp {
namespace {

constexpr int kB = 256;
constexpr uint32_t kSlots = 1024;  // hash slots (patterns beyond 256 are not kept)

struct Form {
    uint32_t n, width, split;
    int two;  // 1: two-colour form split at `split`
};

unsigned grid_rows(uint32_t n) { return std::max(1u, std::min<uint32_t>((n + kB - 1) / kB, 16384u)); }

// 16-bit offsets of every row (padding repeats the row's last offset);
// fail: an offset out of range, a coupling inside a colour, a diagonal
// away from its place
__global__ void k_eb_offsets(Form f, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ ci,
                             int16_t *__restrict__ off, int *fail) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < f.n; i += gridDim.x * kB) {
        uint32_t k0 = rp[i], k1 = rp[i + 1];
        int64_t cb = 0;
        const bool lo = i < f.split;
        bool bad = false;
        if (f.two) {
            const uint32_t kd = lo ? k0 : k1 - 1;
            bad = ci[kd] != i;
            if (lo) ++k0;
            else --k1;
            cb = lo ? (int64_t)f.split : -(int64_t)f.split;
        }
        int16_t last = 0;
        for (uint32_t j = 0; j < f.width; ++j) {
            if (k0 + j < k1) {
                const uint32_t c = ci[k0 + j];
                const int64_t o = (int64_t)c - (int64_t)i - cb;
                bad |= o < INT16_MIN || o > INT16_MAX;
                if (f.two) bad |= (c < f.split) == lo;
                last = (int16_t)o;
            }
            off[(size_t)j * f.n + i] = last;
        }
        if (bad) atomicOr(fail, 1);
    }
}

__device__ inline unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// the hash of row i's offset tuple (never 0: 0 marks an empty slot)
__device__ inline unsigned long long row_hash(const int16_t *__restrict__ off, uint32_t n, uint32_t width,
                                              uint32_t i) {
    unsigned long long lo = 0, hi = 0;
    for (uint32_t j = 0; j < width; ++j) {
        const unsigned long long u = (uint16_t)off[(size_t)j * n + i];
        if (j < 4) lo |= u << (16 * j);
        else hi |= u << (16 * (j - 4));
    }
    return mix64(lo ^ mix64(hi + 0x9e3779b97f4a7c15ull)) | 1ull;
}

__global__ void k_eb_insert(uint32_t n, uint32_t width, const int16_t *__restrict__ off, unsigned long long *tab,
                            uint32_t *first, int *fail) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t base = blockIdx.x * kB + (threadIdx.x & ~63u); base < n; base += gridDim.x * kB) {
        const uint32_t i = base + lane;
        const bool valid = i < n;
        const unsigned long long h = valid ? row_hash(off, n, width, i) : 0ull;
        unsigned long long pending = __ballot(valid);
        while (pending) {
            const int leader = __ffsll((long long)pending) - 1;
            const unsigned long long lh = ((unsigned long long)(uint32_t)__shfl((int)(uint32_t)(h >> 32), leader)
                                           << 32) |
                                          (uint32_t)__shfl((int)(uint32_t)h, leader);
            const unsigned long long same = __ballot(valid && h == lh) & pending;
            if ((int)lane == leader) {
                uint32_t s = (uint32_t)lh & (kSlots - 1);
                bool placed = false;
                for (uint32_t probe = 0; probe < kSlots && !placed; ++probe) {
                    const unsigned long long cur = atomicCAS(tab + s, 0ull, lh);
                    if (cur == 0ull || cur == lh) {
                        atomicMin(first + s, i);
                        placed = true;
                    } else {
                        s = (s + 1) & (kSlots - 1);
                    }
                }
                if (!placed) atomicOr(fail, 2);  // more distinct tuples than slots
            }
            pending &= ~same;
        }
    }
}

// ptab[p * 8 + j]: pattern p's offsets (its first row's), zeros past the width
__global__ void k_eb_ptab(uint32_t npat, uint32_t n, uint32_t width, const uint32_t *__restrict__ reps,
                          const int16_t *__restrict__ off, int16_t *ptab) {
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    if (t >= npat * 8) return;
    const uint32_t p = t / 8, j = t % 8;
    ptab[t] = j < width ? off[(size_t)j * n + reps[p]] : (int16_t)0;
}

// every row's pattern number, its tuple compared with its slot's first row's
__global__ void k_eb_assign(uint32_t n, uint32_t width, const int16_t *__restrict__ off,
                            const unsigned long long *__restrict__ tab, const uint32_t *__restrict__ first,
                            const uint8_t *__restrict__ num, uint8_t *pid, int *fail) {
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
        const unsigned long long h = row_hash(off, n, width, i);
        uint32_t s = (uint32_t)h & (kSlots - 1);
        bool found = false;
        for (uint32_t probe = 0; probe < kSlots && !found; ++probe) {
            if (tab[s] == h) found = true;
            else s = (s + 1) & (kSlots - 1);
        }
        bool bad = !found;
        if (found) {
            const uint32_t r = first[s];
            for (uint32_t j = 0; j < width; ++j) bad |= off[(size_t)j * n + i] != off[(size_t)j * n + r];
            pid[i] = num[s];
        }
        if (bad) atomicOr(fail, 4);
    }
}

}  // namespace

bool ell_build_device(pp_ctx *ctx, const pp_csr *A, bool two_colour, uint32_t split, EllOp &E, double max_fill) {
    E.free_all();
    const uint32_t n = A->nrow;
    const std::vector<uint32_t> &rp = A->h_rp;
    if (n == 0 || rp.size() != (size_t)n + 1) return false;
    // the host construction's conditions on the row lengths
    if (two_colour && (split == 0 || split >= n)) return false;
    const uint32_t tail = two_colour ? 1 : 0;
    std::atomic<uint32_t> wmax{0};
    std::atomic<bool> empty_row{false};  // a row with no sliced entry
    parallel_for(n, 1u << 16, [&](size_t i0, size_t i1) {
        uint32_t m = 0;
        bool e = false;
        for (size_t i = i0; i < i1; ++i) {
            const uint32_t len = rp[i + 1] - rp[i];
            e |= len <= tail;
            m = std::max(m, len);
        }
        if (e) empty_row = true;
        uint32_t cur = wmax.load();
        while (m > cur && !wmax.compare_exchange_weak(cur, m)) {
        }
    });
    if (empty_row.load()) return false;
    const uint32_t width = wmax.load() - tail;
    if (width > 16 || (uint64_t)width * n >= (1ull << 32)) return false;
    if ((double)width * n > max_fill * (double)(rp[n] - (uint64_t)tail * n)) return false;
    const Form f{n, width, two_colour ? split : 0u, two_colour ? 1 : 0};
    int16_t *off = nullptr;
    int *fail = nullptr;
    unsigned long long *tab = nullptr;
    uint32_t *first = nullptr;
    uint32_t *reps = nullptr;
    auto cleanup = [&] {
        dfree(off);
        dfree(fail);
        dfree(tab);
        dfree(first);
        dfree(reps);
    };
    try {
        off = dalloc<int16_t>((size_t)width * n);
        fail = dalloc<int>(1);
        PP_HIP(hipMemsetAsync(fail, 0, sizeof(int), ctx->stream));
        k_eb_offsets<<<grid_rows(n), kB, 0, ctx->stream>>>(f, A->rp, A->ci, off, fail);
        PP_HIP(hipGetLastError());
        int bad = 0;
        PP_HIP(hipMemcpyAsync(&bad, fail, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        if (bad) {
            cleanup();
            return false;
        }
        E.n = n;
        E.width = width;
        if (two_colour) {
            E.cbase = split;
            E.cbase1 = -(int64_t)split;
            E.csplit = split;
            E.dsplit = split;
        }
        // row patterns for the exactly instantiated widths
        bool patterns = false;
        if (width >= 4 && width <= 7) {
            tab = dalloc<unsigned long long>(kSlots);
            first = dalloc<uint32_t>(kSlots);
            PP_HIP(hipMemsetAsync(tab, 0, kSlots * sizeof(unsigned long long), ctx->stream));
            PP_HIP(hipMemsetAsync(first, 0xff, kSlots * sizeof(uint32_t), ctx->stream));
            k_eb_insert<<<grid_rows(n), kB, 0, ctx->stream>>>(n, width, off, tab, first, fail);
            PP_HIP(hipGetLastError());
            std::vector<unsigned long long> htab(kSlots);
            std::vector<uint32_t> hfirst(kSlots);
            PP_HIP(hipMemcpyAsync(htab.data(), tab, kSlots * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  ctx->stream));
            PP_HIP(hipMemcpyAsync(hfirst.data(), first, kSlots * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  ctx->stream));
            PP_HIP(hipMemcpyAsync(&bad, fail, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            PP_HIP(hipStreamSynchronize(ctx->stream));
            std::vector<std::pair<uint32_t, uint32_t>> used;  // (first row, slot)
            for (uint32_t s = 0; s < kSlots; ++s)
                if (htab[s]) used.emplace_back(hfirst[s], s);
            if (!bad && used.size() <= 256) {
                std::sort(used.begin(), used.end());
                const uint32_t npat = (uint32_t)used.size();
                std::vector<uint8_t> num(kSlots, 0);
                std::vector<uint32_t> hreps(npat);
                for (uint32_t p = 0; p < npat; ++p) {
                    num[used[p].second] = (uint8_t)p;
                    hreps[p] = used[p].first;
                }
                reps = dalloc<uint32_t>(npat);
                uint8_t *num_d = dalloc<uint8_t>(kSlots);
                E.pid = dalloc<uint8_t>((size_t)n + 1);  // one more byte: a two-row load at the last even row
                E.ptab = dalloc<int16_t>((size_t)npat * 8);
                E.npat = npat;
                PP_HIP(hipMemcpyAsync(num_d, num.data(), kSlots, hipMemcpyHostToDevice, ctx->stream));
                PP_HIP(hipMemcpyAsync(reps, hreps.data(), npat * sizeof(uint32_t), hipMemcpyHostToDevice,
                                      ctx->stream));
                k_eb_ptab<<<(npat * 8 + kB - 1) / kB, kB, 0, ctx->stream>>>(npat, n, width, reps, off, E.ptab);
                k_eb_assign<<<grid_rows(n), kB, 0, ctx->stream>>>(n, width, off, tab, first, num_d, E.pid, fail);
                PP_HIP(hipGetLastError());
                PP_HIP(hipMemcpyAsync(&bad, fail, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
                PP_HIP(hipStreamSynchronize(ctx->stream));  // num, hreps leave scope
                dfree(num_d);
                patterns = !bad;
                if (!patterns) {  // a hash collision: the offsets stay per slot
                    dfree(E.pid);
                    dfree(E.ptab);
                    E.pid = nullptr;
                    E.ptab = nullptr;
                    E.npat = 0;
                }
            }
        }
        if (patterns) {
            dfree(off);
        } else {
            E.off = off;
        }
        off = nullptr;
        E.v64 = dalloc<double>((size_t)width * n);
        if (two_colour) E.dg64 = dalloc<double>(n);
        ell_fill_values(ctx, E, PP_F64, A->rp, A->v64);
        PP_HIP(hipStreamSynchronize(ctx->stream));
        E.ok = true;
    } catch (...) {
        cleanup();
        E.free_all();
        throw;
    }
    cleanup();
    return true;
}

}  // namespace pp
