// pp_precond.cpp -- preconditioner objects of the C ABI: host factorization,
// conversion to gather-form sweeps, dependency-level analysis and upload
// (rows stored in level order so each level is a contiguous, coalesced range).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cstring>

#include "pp_internal.h"

namespace pp {
// reverse Cuthill-McKee order of a CSR pattern (pp_io.cpp)
std::vector<uint32_t> rcm_order(uint32_t n, const uint32_t *rp, const uint32_t *ci);
// multilevel preconditioners (pp_multilevel.hip)
void ml_setup(pp_ctx *ctx, pp_precond *M);
const pp_precond *ml_block_factor(const pp_precond *M);
void ml_apply(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y);
void ml_free(pp_precond *M);
void ml_need_f32(pp_ctx *ctx, pp_precond *M);

template <class T>
void DevSweep<T>::free_all() {
    if (!borrowed) {
        dfree(perm);
        dfree(rp);
        dfree(ci);
        dfree(blocks);
        dfree(ci16);
        dfree(blk_base);
        dfree(desc);
    }
    for (auto &e : ell) {
        if (!borrowed) dfree(e.off);
        dfree(e.val);
    }
    ell.clear();
    borrowed = false;
    dfree(val);
    dfree(diag);
    blocks = nullptr;
    ci16 = nullptr;
    blk_base = nullptr;
    desc = nullptr;
    level_blk.clear();
    perm = rp = ci = nullptr;
    val = diag = nullptr;
    level_ptr.clear();
    n = nnz = nlev = 0;
    nnz16 = nblk = perm_rows = 0;
}
template struct DevSweep<double>;
template struct DevSweep<float>;

// Levels of the dependency DAG: level(i) = 1 + max level of its dependencies.
static void compute_levels(const HostSweep &s, std::vector<uint32_t> &perm, std::vector<uint32_t> &level_ptr) {
    const uint32_t n = s.n;
    std::vector<uint32_t> lev(n, 0);
    uint32_t maxlev = 0;
    auto visit = [&](uint32_t i) {
        uint32_t l = 0;
        for (uint32_t k = s.rp[i]; k < s.rp[i + 1]; ++k) l = std::max(l, lev[s.ci[k]] + 1);
        lev[i] = l;
        maxlev = std::max(maxlev, l);
    };
    if (s.lower)
        for (uint32_t i = 0; i < n; ++i) visit(i);
    else
        for (uint32_t i = n; i-- > 0;) visit(i);
    const uint32_t nlev = n ? maxlev + 1 : 0;
    level_ptr.assign(nlev + 1, 0);
    for (uint32_t i = 0; i < n; ++i) level_ptr[lev[i] + 1]++;
    for (uint32_t l = 0; l < nlev; ++l) level_ptr[l + 1] += level_ptr[l];
    std::vector<uint32_t> fill(level_ptr.begin(), level_ptr.end() - (nlev ? 1 : 0));
    perm.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i) perm[fill[lev[i]]++] = i;
}

// The sweep's levels, computed on first use and kept with it.
static void levels_of(const HostSweep &s, std::vector<uint32_t> &perm, std::vector<uint32_t> &level_ptr) {
    if (s.lev_ptr.empty() || s.lev_perm.size() != s.n) compute_levels(s, s.lev_perm, s.lev_ptr);
    perm = s.lev_perm;
    level_ptr = s.lev_ptr;
}

template <class T>
void upload_sweep(const HostSweep &s, DevSweep<T> &d) {
    d.free_all();
    std::vector<uint32_t> perm, lp;
    levels_of(s, perm, lp);
    const uint32_t n = s.n;
    std::vector<uint32_t> rp(n + 1, 0);
    for (uint32_t k = 0; k < n; ++k) rp[k + 1] = rp[k] + (s.rp[perm[k] + 1] - s.rp[perm[k]]);
    std::vector<uint32_t> ci(rp[n]);
    std::vector<T> val(rp[n]);
    parallel_for(n, 1u << 14, [&](size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; ++k) {
            const uint32_t i = perm[k];
            for (uint32_t e = s.rp[i], q = rp[k]; e < s.rp[i + 1]; ++e, ++q) {
                ci[q] = s.ci[e];
                val[q] = (T)s.val[e];
            }
        }
    });
    d.n = n;
    d.nnz = (uint32_t)ci.size();
    d.nlev = lp.empty() ? 0 : (uint32_t)lp.size() - 1;
    d.level_ptr = lp;
    d.max_level_rows = 0;
    for (uint32_t l = 0; l < d.nlev; ++l) d.max_level_rows = std::max(d.max_level_rows, lp[l + 1] - lp[l]);
    d.level_off.assign(d.nlev, DevSweep<T>::kNoOffset);
    for (uint32_t l = 0; l < d.nlev; ++l) {
        bool contiguous = true;
        for (uint32_t k = lp[l] + 1; k < lp[l + 1] && contiguous; ++k) contiguous = perm[k] == perm[k - 1] + 1;
        if (contiguous && lp[l + 1] > lp[l]) d.level_off[l] = (int64_t)perm[lp[l]] - (int64_t)lp[l];
    }
    // row blocks of the wide levels (positions, level order)
    std::vector<uint32_t> blk;
    d.level_blk.assign(2 * (size_t)d.nlev, 0);
    for (uint32_t l = 0; l < d.nlev; ++l) {
        if (lp[l + 1] - lp[l] >= kStreamLevelRows) {
            uint32_t k = lp[l];
            if (blk.empty() || blk.back() != k) blk.push_back(k);
            d.level_blk[2 * l] = (uint32_t)blk.size() - 1;
            while (k < lp[l + 1]) {
                const uint32_t start = k, e0 = rp[k];
                if (rp[k + 1] - e0 > kSpmvNnzBlock) {
                    ++k;  // a long row alone
                } else {
                    while (k < lp[l + 1] && k - start < kSpmvRowsBlock && rp[k + 1] - e0 <= kSpmvNnzBlock) ++k;
                }
                blk.push_back(k);
            }
            d.level_blk[2 * l + 1] = (uint32_t)blk.size() - 1;
        }
    }
    if (blk.size() > 1) {
        d.blocks = dalloc<uint32_t>(blk.size());
        PP_HIP(hipMemcpy(d.blocks, blk.data(), blk.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        // 16-bit two-base offsets where a block's dependency columns fit
        std::vector<uint32_t> base;
        std::vector<uint16_t> off16(ci.size(), 0);
        d.nnz16 = two_base_blocks(blk, rp, ci.data(), off16, base);
        d.nblk = blk.size() - 1;
        const std::vector<BlockDesc> ds = block_descs(blk, rp, base);
        d.desc = dalloc<BlockDesc>(ds.size());
        PP_HIP(hipMemcpy(d.desc, ds.data(), ds.size() * sizeof(BlockDesc), hipMemcpyHostToDevice));
        if (d.nnz16) {
            d.ci16 = dalloc<uint16_t>(off16.size());
            d.blk_base = dalloc<uint32_t>(base.size());
            PP_HIP(hipMemcpy(d.ci16, off16.data(), off16.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
            PP_HIP(hipMemcpy(d.blk_base, base.data(), base.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        }
    }
    for (uint32_t l = 0; l < d.nlev; ++l)
        if (d.level_off[l] == DevSweep<T>::kNoOffset) d.perm_rows += lp[l + 1] - lp[l];
    d.perm = dalloc<uint32_t>(n);
    d.rp = dalloc<uint32_t>(n + 1);
    d.ci = dalloc<uint32_t>(d.nnz);
    d.val = dalloc<T>(d.nnz);
    PP_HIP(hipMemcpy(d.perm, perm.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    PP_HIP(hipMemcpy(d.rp, rp.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (d.nnz) {
        PP_HIP(hipMemcpy(d.ci, ci.data(), d.nnz * sizeof(uint32_t), hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(d.val, val.data(), d.nnz * sizeof(T), hipMemcpyHostToDevice));
    }
    if (s.mode != HostSweep::UNIT) {
        std::vector<T> dg(n);
        for (uint32_t i = 0; i < n; ++i) dg[i] = (T)s.diag[i];
        d.diag = dalloc<T>(n);
        PP_HIP(hipMemcpy(d.diag, dg.data(), n * sizeof(T), hipMemcpyHostToDevice));
    }
    // sliced-column form of the wide levels whose rows are contiguous, for
    // factors large enough for the setup to time it against the staged
    // kernel (tune_sweeps); smaller ones would never run it
    d.ell.assign(d.nlev, typename DevSweep<T>::EllLevel());
    for (uint32_t l = 0; l < d.nlev && n >= kTuneSweepRows; ++l) {
        const uint32_t k0 = lp[l], k1 = lp[l + 1], nr = k1 - k0;
        if (nr < kStreamLevelRows || d.level_off[l] == DevSweep<T>::kNoOffset) continue;
        const int64_t row0 = (int64_t)k0 + d.level_off[l];
        uint32_t w = 0;
        bool ok = true;
        for (uint32_t k = k0; k < k1 && ok; ++k) {
            w = std::max(w, rp[k + 1] - rp[k]);
            ok = rp[k + 1] > rp[k];
        }
        if (!ok || w > 16 || (double)w * nr > 1.25 * (double)(rp[k1] - rp[k0])) continue;
        const int64_t cbase = (int64_t)ci[rp[k0]] - row0;
        std::vector<int16_t> off((size_t)w * nr);
        std::atomic<bool> fits{true};
        parallel_for(nr, 1u << 15, [&](size_t t0, size_t t1) {
            for (size_t t = t0; t < t1; ++t) {
                const uint32_t k = k0 + (uint32_t)t;
                int16_t last = 0;
                for (uint32_t j = 0; j < w; ++j) {
                    const size_t e = (size_t)j * nr + t;
                    if (rp[k] + j < rp[k + 1]) {
                        const int64_t o = (int64_t)ci[rp[k] + j] - (row0 + (int64_t)t) - cbase;
                        if (o < INT16_MIN || o > INT16_MAX) {
                            fits = false;
                            return;
                        }
                        last = (int16_t)o;
                    }
                    off[e] = last;
                }
            }
        });
        if (!fits) continue;
        auto &E = d.ell[l];
        E.nrows = nr;
        E.entries = rp[k1] - rp[k0];
        E.width = w;
        E.row0 = row0;
        E.cbase = cbase;
        E.off = dalloc<int16_t>(off.size());
        E.val = dalloc<T>(off.size());
        PP_HIP(hipMemcpy(E.off, off.data(), off.size() * sizeof(int16_t), hipMemcpyHostToDevice));
        // the values from the level's rows already on the device (zeros in the padding)
        ell_fill_rows<T>(nr, w, d.rp + k0, d.val, E.val);
        E.ok = true;
    }
}
// The FP32 sweep of the same rows as an uploaded FP64 sweep: the index
// arrays are shared, the values and pivots rounded to FP32 on the device
// (round to nearest, as a host cast).
static void derive_sweep_f32(pp_ctx *ctx, const DevSweep<double> &s, DevSweep<float> &d) {
    d.free_all();
    d.n = s.n;
    d.nnz = s.nnz;
    d.nlev = s.nlev;
    d.perm = s.perm;
    d.rp = s.rp;
    d.ci = s.ci;
    d.blocks = s.blocks;
    d.desc = s.desc;
    d.ci16 = s.ci16;
    d.blk_base = s.blk_base;
    d.borrowed = true;
    d.level_ptr = s.level_ptr;
    d.level_off = s.level_off;
    d.level_blk = s.level_blk;
    d.nnz16 = s.nnz16;
    d.nblk = s.nblk;
    d.perm_rows = s.perm_rows;
    d.max_level_rows = s.max_level_rows;
    d.val = dalloc<float>(s.nnz);
    if (s.nnz) launch_convert(ctx, s.nnz, PP_F64, s.val, PP_F32, d.val);
    d.ell.assign(s.ell.size(), DevSweep<float>::EllLevel());
    for (size_t l = 0; l < s.ell.size(); ++l) {
        const auto &S = s.ell[l];
        if (!S.ok) continue;
        auto &E = d.ell[l];
        E.nrows = S.nrows;
        E.entries = S.entries;
        E.width = S.width;
        E.row0 = S.row0;
        E.cbase = S.cbase;
        E.off = S.off;
        E.val = dalloc<float>((size_t)S.width * S.nrows);
        launch_convert(ctx, S.width * S.nrows, PP_F64, S.val, PP_F32, E.val);
        E.ok = true;
    }
    if (s.diag) {
        d.diag = dalloc<float>(s.n);
        launch_convert(ctx, s.n, PP_F64, s.diag, PP_F32, d.diag);
    }
    PP_HIP(hipStreamSynchronize(ctx->stream));
}

void ChunkSet::free_all() {
    dfree(dstart);
    dstart = nullptr;
    start.clear();
    nchunk = max_rows = 0;
}

template <class T>
void DevChunkSweep<T>::free_all() {
    dfree(row);
    dfree(rp);
    dfree(ci);
    dfree(val);
    dfree(pdiag);
    dfree(lp);
    dfree(le);
    dfree(lev_off);
    row = ci = nullptr;
    rp = lp = le = lev_off = nullptr;
    val = pdiag = nullptr;
    n = nnz = max_levels = max_level_nnz = 0;
}
template struct DevChunkSweep<double>;
template struct DevChunkSweep<float>;

// Contiguous chunks of at most cap rows that no dependency of the given
// sweeps crosses; false when a dependency-closed range is larger than cap.
static bool build_chunks(const std::vector<const HostSweep *> &sw, uint32_t n, uint32_t cap, ChunkSet &C) {
    // blocked[r] > 0: some dependency spans the cut between rows r-1 and r
    std::vector<int32_t> blocked(n + 1, 0);
    for (const HostSweep *s : sw)
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t a = i, b = i;
            for (uint32_t k = s->rp[i]; k < s->rp[i + 1]; ++k) {
                a = std::min(a, s->ci[k]);
                b = std::max(b, s->ci[k]);
            }
            if (a < b) {
                blocked[a + 1] += 1;
                blocked[b + 1] -= 1;
            }
        }
    C.start.assign(1, 0);
    C.max_rows = 0;
    int32_t run = 0;
    uint32_t comp0 = 0;  // first row of the current dependency-closed component
    for (uint32_t r = 1; r <= n; ++r) {
        run += blocked[r];
        if (r < n && run > 0) continue;  // no cut before row r
        // [comp0, r) is a component
        if (r - comp0 > cap) return false;
        if (r - C.start.back() > cap) C.start.push_back(comp0);  // close the chunk before this component
        comp0 = r;
    }
    C.start.push_back(n);
    C.nchunk = (uint32_t)C.start.size() - 1;
    for (uint32_t c = 0; c < C.nchunk; ++c) C.max_rows = std::max(C.max_rows, C.start[c + 1] - C.start[c]);
    return true;
}

// The sweep in chunk-major, level-ordered positions (DevChunkSweep layout).
template <class T>
static void upload_chunk_sweep(const HostSweep &s, const ChunkSet &C, DevChunkSweep<T> &d) {
    d.free_all();
    const uint32_t n = s.n;
    std::vector<uint32_t> gperm, glp;
    levels_of(s, gperm, glp);
    std::vector<uint32_t> lev(n), chunk_of(n);
    for (uint32_t l = 0; l + 1 < glp.size(); ++l)
        for (uint32_t k = glp[l]; k < glp[l + 1]; ++k) lev[gperm[k]] = l;
    for (uint32_t c = 0; c < C.nchunk; ++c)
        for (uint32_t i = C.start[c]; i < C.start[c + 1]; ++i) chunk_of[i] = c;
    std::vector<std::vector<uint32_t>> rows(C.nchunk);
    for (uint32_t k = 0; k < n; ++k) rows[chunk_of[gperm[k]]].push_back(gperm[k]);
    std::vector<uint16_t> row, ci;
    std::vector<T> val, pdiag;
    std::vector<uint32_t> lp, off(1, 0), rp(1, 0);
    row.reserve(n);
    for (uint32_t c = 0; c < C.nchunk; ++c) {
        const uint32_t r0 = C.start[c];
        uint32_t nl = 0;
        for (size_t q = 0; q < rows[c].size(); ++q) {
            const uint32_t i = rows[c][q];
            if (q == 0 || lev[i] != lev[rows[c][q - 1]]) {
                if (!lp.empty() && nl) d.max_level_nnz = std::max(d.max_level_nnz, rp.back() - rp[lp.back()]);
                lp.push_back((uint32_t)row.size());
                ++nl;
            }
            row.push_back((uint16_t)(i - r0));
            if (s.mode != HostSweep::UNIT) pdiag.push_back((T)s.diag[i]);
            for (uint32_t e = s.rp[i]; e < s.rp[i + 1]; ++e) {
                ci.push_back((uint16_t)(s.ci[e] - r0));
                val.push_back((T)s.val[e]);
            }
            rp.push_back((uint32_t)ci.size());
        }
        if (nl) d.max_level_nnz = std::max(d.max_level_nnz, rp.back() - rp[lp.back()]);
        lp.push_back((uint32_t)row.size());
        off.push_back((uint32_t)lp.size());
        d.max_levels = std::max(d.max_levels, nl);
    }
    d.n = n;
    d.nnz = (uint32_t)ci.size();
    auto up = [](const auto &v, auto *&dst) {
        using E = std::remove_const_t<std::remove_reference_t<decltype(v[0])>>;
        dst = dalloc<E>(std::max<size_t>(1, v.size()));
        if (!v.empty()) PP_HIP(hipMemcpy(dst, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice));
    };
    up(row, d.row);
    up(rp, d.rp);
    up(ci, d.ci);
    up(val, d.val);
    std::vector<uint32_t> le(lp.size());
    for (size_t j = 0; j < lp.size(); ++j) le[j] = rp[lp[j]];
    up(lp, d.lp);
    up(le, d.le);
    up(off, d.lev_off);
    if (s.mode != HostSweep::UNIT) up(pdiag, d.pdiag);
}

// PP_TRSV_BLOCK storage for M (FP64; FP32 too when the FP32 sweeps exist).
// false (nothing kept) when the rows do not chunk within the LDS size.
static void ensure_host_setup(pp_precond *M);

static bool build_block_sweeps(pp_ctx *ctx, pp_precond *M) {
    ensure_host_setup(M);  // the chunks are cut from the host sweeps
    M->chunks.free_all();
    M->clow64.free_all();
    M->cup64.free_all();
    M->clow32.free_all();
    M->cup32.free_all();
    const uint32_t cap = (uint32_t)std::min<size_t>(ctx->max_lds / sizeof(double), 65535);
    if (!build_chunks({&M->hlow, &M->hup}, M->n, cap, M->chunks)) {
        M->chunks.free_all();
        return false;
    }
    M->chunks.dstart = dalloc<uint32_t>(M->chunks.start.size());
    PP_HIP(hipMemcpy(M->chunks.dstart, M->chunks.start.data(), M->chunks.start.size() * sizeof(uint32_t),
                     hipMemcpyHostToDevice));
    upload_chunk_sweep(M->hlow, M->chunks, M->clow64);
    upload_chunk_sweep(M->hup, M->chunks, M->cup64);
    // the chunk's vector and both sweeps' level boundaries must fit in LDS
    const size_t need = (size_t)M->chunks.max_rows * sizeof(double) +
                        4 * (size_t)(2 * (M->clow64.max_levels + 1) + 2 * (M->cup64.max_levels + 1));
    if (need > ctx->max_lds) {
        M->chunks.free_all();
        M->clow64.free_all();
        M->cup64.free_all();
        return false;
    }
    if (M->have32) {
        upload_chunk_sweep(M->hlow, M->chunks, M->clow32);
        upload_chunk_sweep(M->hup, M->chunks, M->cup32);
    }
    return true;
}

template void upload_sweep<double>(const HostSweep &, DevSweep<double> &);
template void upload_sweep<float>(const HostSweep &, DevSweep<float> &);

// Gather form of a sweep given by the rows of R:
//   diag_first: the first entry of each row is its diagonal (IC L^T layout).
static HostSweep sweep_from_rows(const HostCsr &R, bool lower, int mode, bool diag_first,
                                 const std::vector<double> *diag_vec) {
    HostSweep s;
    s.n = R.n;
    s.lower = lower;
    s.mode = mode;
    s.rp.assign(R.n + 1, 0);
    s.diag.assign(R.n, 1.0);
    const uint32_t skip = diag_first ? 1 : 0;
    for (uint32_t i = 0; i < R.n; ++i) s.rp[i + 1] = s.rp[i] + (R.rp[i + 1] - R.rp[i] - skip);
    s.ci.resize(s.rp[R.n]);
    s.val.resize(s.rp[R.n]);
    parallel_for(R.n, 1u << 15, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) {
            const uint32_t b = R.rp[i] + skip;
            if (diag_first) s.diag[i] = R.v[R.rp[i]];
            std::copy(R.ci.begin() + b, R.ci.begin() + R.rp[i + 1], s.ci.begin() + s.rp[i]);
            std::copy(R.v.begin() + b, R.v.begin() + R.rp[i + 1], s.val.begin() + s.rp[i]);
        }
    });
    if (diag_vec) s.diag = *diag_vec;
    return s;
}

// Column j of a CSR with the diagonal first becomes row j of the transpose
// with the diagonal last; returns the gather-form lower sweep.
static HostSweep lower_from_upper_rows(const HostCsr &Ut, int mode, bool diag_first) {
    const HostCsr L = transpose(Ut);  // rows ascending in column
    HostSweep s;
    s.n = L.n;
    s.lower = true;
    s.mode = mode;
    s.rp.assign(L.n + 1, 0);
    s.diag.assign(L.n, 1.0);
    std::vector<uint32_t> cnt(L.n);
    parallel_for(L.n, 1u << 15, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) {
            uint32_t c = 0;
            for (uint32_t k = L.rp[i]; k < L.rp[i + 1]; ++k) c += L.ci[k] != i;
            cnt[i] = c;
        }
    });
    for (uint32_t i = 0; i < L.n; ++i) s.rp[i + 1] = s.rp[i] + cnt[i];
    s.ci.resize(s.rp[L.n]);
    s.val.resize(s.rp[L.n]);
    parallel_for(L.n, 1u << 15, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) {
            uint32_t q = s.rp[i];
            for (uint32_t k = L.rp[i]; k < L.rp[i + 1]; ++k) {
                if (L.ci[k] == i) {
                    if (diag_first) s.diag[i] = L.v[k];
                    continue;
                }
                s.ci[q] = L.ci[k];
                s.val[q] = L.v[k];
                ++q;
            }
        }
    });
    return s;
}

static double ms_since(std::chrono::steady_clock::time_point &t) {
    const auto now = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(now - t).count();
    t = now;
    return ms;
}

// The first level's rows (when contiguous) and the FP32 sweeps, once the
// FP64 sweeps are on the device.
static void finish_device_sweeps(pp_precond *M) {
    M->lvl0_contiguous = false;
    if (M->low64.nlev > 0 && M->low64.level_off[0] != DevSweep<double>::kNoOffset) {
        M->lvl0_contiguous = true;
        M->lvl0_lo = (uint32_t)((int64_t)M->low64.level_ptr[0] + M->low64.level_off[0]);
        M->lvl0_hi = (uint32_t)((int64_t)M->low64.level_ptr[1] + M->low64.level_off[0]);
    }
    if (M->opts.make_f32) {
        derive_sweep_f32(M->ctx, M->low64, M->low32);
        derive_sweep_f32(M->ctx, M->up64, M->up32);
        M->have32 = true;
    }
}

// perm_out (multicolour order): receives P A P^T, which the factor was
// computed from.
static void setup_triangular(pp_precond *M, const HostCsr &A0, bool upload, HostCsr *perm_out = nullptr) {
    auto tp = std::chrono::steady_clock::now();
    HostCsr Aperm;
    const HostCsr *Ap = &A0;
    M->order.clear();
    if (M->opts.ordering == PP_ORDER_MULTICOLOR) {
        M->order = multicolor_order(A0, M->ncolors);
        Aperm = permute_symmetric(A0, M->order);
        Ap = &Aperm;
    } else if (M->opts.ordering == PP_ORDER_RCM) {
        M->order = rcm_order(A0.n, A0.rp.data(), A0.ci.data());
        Aperm = permute_symmetric(A0, M->order);
        Ap = &Aperm;
    } else {
        PP_CHECK(M->opts.ordering == PP_ORDER_NATURAL, PP_ERR_INVALID, "unknown ordering");
    }
    M->phase_ms[0] = (float)ms_since(tp);
    const HostCsr &A = *Ap;
    const int kind = M->opts.kind;
    int st = PP_OK;
    uint32_t fail = 0;
    if (kind == PP_PC_ICT || kind == PP_PC_IC0 || kind == PP_PC_ILDLT) {
        HostCsr lt;
        std::vector<double> d;
        if (kind == PP_PC_ICT) st = factor_ict(A, M->opts.tau, false, lt, d, fail);
        else if (kind == PP_PC_ILDLT) st = factor_ict(A, M->opts.tau, true, lt, d, fail);
        else st = factor_ic0(A, lt, fail, &M->hlow);
        M->d = std::move(d);
        M->mnnz = (uint32_t)(lt.ci.size() * 2 + A.n);  // IncompleteCholesky.cpp:245 convention
        if (st == PP_OK) {
            if (kind == PP_PC_ILDLT) {
                M->hlow = lower_from_upper_rows(lt, HostSweep::UNIT, false);
                M->hup = sweep_from_rows(lt, false, HostSweep::PREDIVIDE, false, &M->d);
            } else {
                if (kind != PP_PC_IC0 || M->hlow.rp.empty())  // IC(0) of sorted rows: filled by the factorization
                    M->hlow = lower_from_upper_rows(lt, HostSweep::DIVIDE, true);
                M->hup = sweep_from_rows(lt, false, HostSweep::DIVIDE, true, nullptr);
            }
        }
        M->lt_rp = std::move(lt.rp);
        M->lt_ci = std::move(lt.ci);
        M->lt_v = std::move(lt.v);
        if (st != PP_OK) {
            M->fail_row = fail;
            throw Error(st, kind == PP_PC_ILDLT ? "ILDLT: zero pivot at row " + std::to_string(fail)
                                                : "IC: matrix is not positive definite (pivot row " +
                                                      std::to_string(fail) + ")");
        }
    } else if (kind == PP_PC_ILUT || kind == PP_PC_ILU0) {
        HostCsr L, U;
        std::vector<double> D;
        if (kind == PP_PC_ILUT) st = factor_ilut(A, M->opts.lfil, M->opts.tau, L, U, D, fail);
        else st = factor_ilu0(A, L, U, D, fail);
        M->l_rp = L.rp; M->l_ci = L.ci; M->l_v = L.v;
        M->u_rp = U.rp; M->u_ci = U.ci; M->u_v = U.v;
        M->d = D;
        M->mnnz = (uint32_t)(L.ci.size() + U.ci.size() + A.n);
        if (st != PP_OK) {
            M->fail_row = fail;
            throw Error(st, std::string(st == PP_ERR_ZERO_ROW ? "ILU: zero row " : "ILU: zero pivot at row ") +
                                std::to_string(fail));
        }
        M->hlow = sweep_from_rows(L, true, HostSweep::UNIT, false, nullptr);
        M->hup = sweep_from_rows(U, false, HostSweep::MULTIPLY, false, &M->d);
    } else {
        throw Error(PP_ERR_UNSUPPORTED, "unknown preconditioner kind");
    }
    M->phase_ms[1] = (float)ms_since(tp);
    // the forward sweep's last level and the backward sweep's first level
    // (rows without dependencies) hold the same rows: one kernel does both
    {
        std::vector<uint32_t> pl, ll, pu, lu;
        levels_of(M->hlow, pl, ll);
        levels_of(M->hup, pu, lu);
        M->fuse_mid = false;
        M->fuse_rows = 0;
        if (ll.size() >= 2 && lu.size() >= 2) {
            const uint32_t a0 = ll[ll.size() - 2], a1 = ll.back(), b0 = lu[0], b1 = lu[1];
            if (a1 - a0 == b1 - b0 && std::equal(pl.begin() + a0, pl.begin() + a1, pu.begin() + b0)) {
                M->fuse_mid = true;
                M->fuse_rows = a1 - a0;
            }
        }
    }
    if (!upload) return;
    if (!M->order.empty()) {
        M->dorder = dalloc<uint32_t>(A.n);
        PP_HIP(hipMemcpy(M->dorder, M->order.data(), A.n * sizeof(uint32_t), hipMemcpyHostToDevice));
        M->xp64 = dalloc<double>(A.n);
    }
    upload_sweep(M->hlow, M->low64);
    upload_sweep(M->hup, M->up64);
    finish_device_sweeps(M);
    if (perm_out && Ap == &Aperm) *perm_out = std::move(Aperm);
}

// The host structures of a preconditioner set up on the device (the
// factor, its sweeps), from the matrix's current values: the same results
// the device path computed.
static void ensure_host_setup(pp_precond *M) {
    if (!M->host_pending) return;
    float saved[4];
    std::copy(M->phase_ms, M->phase_ms + 4, saved);
    HostCsr A;
    A.n = M->A->nrow;
    A.rp.swap(M->A->h_rp);
    A.ci.swap(M->A->h_ci);
    A.v.swap(M->A->h_v);
    try {
        setup_triangular(M, A, false);
    } catch (...) {
        A.rp.swap(M->A->h_rp);
        A.ci.swap(M->A->h_ci);
        A.v.swap(M->A->h_v);
        throw;
    }
    A.rp.swap(M->A->h_rp);
    A.ci.swap(M->A->h_ci);
    A.v.swap(M->A->h_v);
    std::copy(saved, saved + 4, M->phase_ms);
    M->host_pending = false;
    M->host_factor_stale = false;
}

// ---------------------------------------------------------------------------
// Numeric re-factorization on the device (pp_precond_refactor)
// ---------------------------------------------------------------------------
void DevRefactor::free_all() {
    for (uint32_t **p : {&rp, &ci, &src, &dsrc, &rows, &lo_map, &lo_dmap, &up_map, &up_dmap, &aperm_map}) {
        dfree(*p);
        *p = nullptr;
    }
    dfree(val);
    val = nullptr;
    dfree(fail);
    fail = nullptr;
}

template <class T>
static T *upload_vec(const std::vector<T> &v) {
    T *d = dalloc<T>(std::max<size_t>(v.size(), 1));
    if (!v.empty()) PP_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// The value slots of a sweep built from `s` (val = the factor entry index,
// diag = the pivot index, both as exact doubles) as index maps: the same
// upload the real sweep went through, so every slot lines up with the real
// sweep's storage.
static void sweep_maps(pp_ctx *ctx, const HostSweep &s, uint32_t *&map, uint32_t &nmap, uint32_t *&dmap,
                       bool &has_diag) {
    DevSweep<double> tmp;
    try {
        upload_sweep(s, tmp);
        nmap = tmp.nnz;
        map = dalloc<uint32_t>(std::max<uint32_t>(nmap, 1));
        launch_to_index(ctx, nmap, tmp.val, map);
        has_diag = tmp.diag != nullptr;
        if (has_diag) {
            dmap = dalloc<uint32_t>(s.n);
            launch_to_index(ctx, s.n, tmp.diag, dmap);
        }
        PP_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        tmp.free_all();
        throw;
    }
    tmp.free_all();
}

// Dependency levels of the strict lower pattern (rp, ci): rows in level order.
static void lower_levels(uint32_t n, const std::vector<uint32_t> &rp, const std::vector<uint32_t> &ci,
                         std::vector<uint32_t> &rows, std::vector<uint32_t> &lptr) {
    std::vector<uint32_t> lev(n, 0);
    uint32_t nlev = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t l = 0;
        for (uint32_t q = rp[i]; q < rp[i + 1]; ++q)
            if (ci[q] < i) l = std::max(l, lev[ci[q]] + 1);
        lev[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    lptr.assign(nlev + 1, 0);
    for (uint32_t i = 0; i < n; ++i) lptr[lev[i] + 1]++;
    for (uint32_t l = 0; l < nlev; ++l) lptr[l + 1] += lptr[l];
    rows.resize(n);
    std::vector<uint32_t> f(lptr.begin(), lptr.end() - 1);
    for (uint32_t i = 0; i < n; ++i) rows[f[lev[i]]++] = i;
}

// Builds the symbolic state from the preconditioner's host structures (the
// pattern the last host setup factored).  False when this factor cannot be
// re-factored on the device (unsorted or repeated columns, a missing
// diagonal, the one-workgroup schedule).
static bool devref_build(pp_ctx *ctx, pp_precond *M) {
    const int kind = M->opts.kind;
    if ((kind != PP_PC_IC0 && kind != PP_PC_ILU0) || M->trsv == PP_TRSV_BLOCK || M->trsv == PP_TRSV_PIPE ||
        M->multilevel)
        return false;
    const bool ordered = !M->order.empty();
    if (ordered && !M->Aperm) return false;
    const pp_csr *F = ordered ? M->Aperm : M->A;  // the matrix the factor was computed from
    const uint32_t n = F->nrow;
    const std::vector<uint32_t> &frp = F->h_rp, &fci = F->h_ci;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = frp[i] + 1; k < frp[i + 1]; ++k)
            if (!(fci[k - 1] < fci[k])) return false;  // the device rows merge sorted column lists
    std::vector<uint32_t> diagpos(n, UINT32_MAX);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = frp[i]; k < frp[i + 1]; ++k)
            if (fci[k] == i) diagpos[i] = k;
    for (uint32_t i = 0; i < n; ++i)
        if (diagpos[i] == UINT32_MAX) return false;
    std::unique_ptr<DevRefactor> R(new DevRefactor);
    R->n = n;
    std::vector<uint32_t> rows;
    if (kind == PP_PC_IC0) {
        // L's strict lower part in the order of the matrix's rows (factor_ic0)
        std::vector<uint32_t> lrp(n + 1, 0), lci, lsrc;
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t k = frp[i]; k < frp[i + 1]; ++k)
                if (fci[k] < i) {
                    lci.push_back(fci[k]);
                    lsrc.push_back(k);
                }
            lrp[i + 1] = (uint32_t)lci.size();
        }
        const uint32_t nnz = (uint32_t)lci.size();
        R->nnz = nnz;
        lower_levels(n, lrp, lci, rows, R->lptr);
        // the sweeps' value slots: forward sweep = L's rows (DIVIDE by l_ii),
        // backward sweep = L^T's rows with the pivot first
        if (M->hlow.rp != lrp || M->hlow.ci != lci) return false;  // not the sorted-rows factor layout
        HostSweep lo = M->hlow;
        for (uint32_t t = 0; t < nnz; ++t) lo.val[t] = (double)t;
        for (uint32_t i = 0; i < n; ++i) lo.diag[i] = (double)(nnz + i);
        HostCsr lt;
        lt.n = n;
        lt.rp = M->lt_rp;
        lt.ci = M->lt_ci;
        lt.v.resize(lt.ci.size());
        for (uint32_t j = 0; j < n; ++j) {
            PP_CHECK(lt.rp[j] < lt.rp[j + 1] && lt.ci[lt.rp[j]] == j, PP_ERR_INVALID, "L^T row without its pivot");
            lt.v[lt.rp[j]] = (double)(nnz + j);
            for (uint32_t q = lt.rp[j] + 1; q < lt.rp[j + 1]; ++q) {
                const uint32_t i = lt.ci[q];
                const auto b = lci.begin() + lrp[i], e = lci.begin() + lrp[i + 1];
                const auto it = std::lower_bound(b, e, j);
                PP_CHECK(it != e && *it == j, PP_ERR_INVALID, "L^T entry outside L's pattern");
                lt.v[q] = (double)(it - lci.begin());
            }
        }
        R->h_lt_map.resize(lt.v.size());
        for (size_t q = 0; q < lt.v.size(); ++q) R->h_lt_map[q] = (uint32_t)lt.v[q];
        HostSweep up = sweep_from_rows(lt, false, HostSweep::DIVIDE, true, nullptr);
        sweep_maps(ctx, lo, R->lo_map, R->lo_nnz, R->lo_dmap, R->lo_diag);
        sweep_maps(ctx, up, R->up_map, R->up_nnz, R->up_dmap, R->up_diag);
        R->rp = upload_vec(lrp);
        R->ci = upload_vec(lci);
        R->src = upload_vec(lsrc);
        R->dsrc = upload_vec(diagpos);
    } else {
        // ILU(0) on the matrix's own pattern: L strict lower (unit), U strict upper, D = 1/u_ii
        const uint32_t nnz = (uint32_t)fci.size();
        R->nnz = nnz;
        lower_levels(n, frp, fci, rows, R->lptr);
        HostCsr L, U;
        L.n = U.n = n;
        L.rp.assign(n + 1, 0);
        U.rp.assign(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t k = frp[i]; k < frp[i + 1]; ++k) {
                if (fci[k] < i) {
                    L.ci.push_back(fci[k]);
                    L.v.push_back((double)k);
                } else if (fci[k] > i) {
                    U.ci.push_back(fci[k]);
                    U.v.push_back((double)k);
                }
            }
            L.rp[i + 1] = (uint32_t)L.ci.size();
            U.rp[i + 1] = (uint32_t)U.ci.size();
        }
        std::vector<double> dix(n);
        for (uint32_t i = 0; i < n; ++i) dix[i] = (double)(nnz + i);
        HostSweep lo = sweep_from_rows(L, true, HostSweep::UNIT, false, nullptr);
        HostSweep up = sweep_from_rows(U, false, HostSweep::MULTIPLY, false, &dix);
        R->h_l_map.assign(L.v.begin(), L.v.end());
        R->h_u_map.assign(U.v.begin(), U.v.end());
        if (lo.rp != M->hlow.rp || lo.ci != M->hlow.ci || up.rp != M->hup.rp || up.ci != M->hup.ci) return false;
        sweep_maps(ctx, lo, R->lo_map, R->lo_nnz, R->lo_dmap, R->lo_diag);
        sweep_maps(ctx, up, R->up_map, R->up_nnz, R->up_dmap, R->up_diag);
        R->rp = upload_vec(frp);
        R->ci = upload_vec(fci);
        R->dsrc = upload_vec(diagpos);
    }
    PP_CHECK(R->lo_nnz == M->low64.nnz && R->up_nnz == M->up64.nnz, PP_ERR_INVALID, "sweep maps out of step");
    R->rows = upload_vec(rows);
    R->val = dalloc<double>((size_t)R->nnz + n);
    R->fail = dalloc<int>(1);
    if (ordered) {
        // entry k of P A P^T <- entry aperm_map[k] of A: the permutation of the entry indices
        HostCsr H;
        H.n = M->A->nrow;
        H.rp = M->A->h_rp;
        H.ci = M->A->h_ci;
        H.v.resize(H.ci.size());
        for (size_t k = 0; k < H.v.size(); ++k) H.v[k] = (double)k;
        const HostCsr P = permute_symmetric(H, M->order);
        PP_CHECK(P.rp == frp && P.ci == fci, PP_ERR_INVALID, "colour-ordered pattern out of step");
        std::vector<uint32_t> amap(P.v.size());
        for (size_t k = 0; k < amap.size(); ++k) amap[k] = (uint32_t)P.v[k];
        R->aperm_map = upload_vec(amap);
        R->aperm_nnz = (uint32_t)amap.size();
    }
    delete M->dref;
    M->dref = R.release();
    return true;
}

// The FP32 sweep's values from the FP64 sweep's, into the arrays it has
// (the same storage the first derivation allocated: cached graphs stay valid).
static void refresh_sweep_f32(pp_ctx *ctx, const DevSweep<double> &s, DevSweep<float> &d) {
    if (s.nnz) launch_convert(ctx, s.nnz, PP_F64, s.val, PP_F32, d.val);
    for (size_t l = 0; l < s.ell.size() && l < d.ell.size(); ++l)
        if (s.ell[l].ok && d.ell[l].ok)
            launch_convert(ctx, s.ell[l].width * s.ell[l].nrows, PP_F64, s.ell[l].val, PP_F32, d.ell[l].val);
    if (s.diag && d.diag) launch_convert(ctx, s.n, PP_F64, s.diag, PP_F32, d.diag);
}

static void refill_ell_levels(DevSweep<double> &d) {
    for (uint32_t l = 0; l < d.nlev && l < d.ell.size(); ++l) {
        auto &E = d.ell[l];
        if (!E.ok) continue;
        const uint32_t k0 = d.level_ptr[l];
        ell_fill_rows<double>(E.nrows, E.width, d.rp + k0, d.val, E.val);
    }
}

// Numeric IC(0) / ILU(0) of the matrix's current values on the device, into
// the existing sweeps.  Throws the factorization's failure like the host.
static void devref_run(pp_ctx *ctx, pp_precond *M) {
    DevRefactor &R = *M->dref;
    const bool ordered = !M->order.empty();
    const double *av = M->A->v64;
    if (ordered) {
        pp_csr *P = M->Aperm;
        launch_gather_map(ctx, R.aperm_nnz, R.aperm_map, M->A->v64, PP_F64, P->v64);
        if (P->ell.ok) ell_fill_values(ctx, P->ell, PP_F64, P->rp, P->v64);
        if (P->h_v.size() == R.aperm_nnz)  // a matrix built on the device keeps no host values
            PP_HIP(hipMemcpyAsync(P->h_v.data(), P->v64, (size_t)R.aperm_nnz * sizeof(double),
                                  hipMemcpyDeviceToHost, ctx->stream));
        P->f32_dirty = true;
        ++P->version;
        if (P->trans) {
            csr_free(P->trans);
            P->trans = nullptr;
        }
        M->Aperm_version = M->A->version;
        av = P->v64;
    }
    const int big = INT32_MAX;
    PP_HIP(hipMemcpyAsync(R.fail, &big, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    for (size_t l = 0; l + 1 < R.lptr.size(); ++l) {
        const uint32_t r0 = R.lptr[l], nr = R.lptr[l + 1] - r0;
        if (M->opts.kind == PP_PC_IC0)
            launch_ic0_level(ctx, R.rows + r0, nr, R.rp, R.ci, R.src, R.dsrc, av, R.val, R.val + R.nnz, R.fail);
        else
            launch_ilu0_level(ctx, R.rows + r0, nr, R.rp, R.ci, R.dsrc, av, R.val, R.val + R.nnz, R.fail);
    }
    int fail = big;
    PP_HIP(hipMemcpyAsync(&fail, R.fail, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    if (fail != big) {
        M->fail_row = (uint32_t)fail;
        M->ready = 0;
        if (M->opts.kind == PP_PC_IC0)
            throw Error(PP_ERR_NOT_SPD, "IC: matrix is not positive definite (pivot row " + std::to_string(fail) + ")");
        throw Error(PP_ERR_ZERO_PIVOT, "ILU: zero pivot at row " + std::to_string(fail));
    }
    // the factor into the sweeps' level-ordered storage
    launch_gather_map(ctx, R.lo_nnz, R.lo_map, R.val, PP_F64, M->low64.val);
    if (R.lo_diag) launch_gather_map(ctx, R.n, R.lo_dmap, R.val, PP_F64, M->low64.diag);
    launch_gather_map(ctx, R.up_nnz, R.up_map, R.val, PP_F64, M->up64.val);
    if (R.up_diag) launch_gather_map(ctx, R.n, R.up_dmap, R.val, PP_F64, M->up64.diag);
    PP_HIP(hipStreamSynchronize(ctx->stream));
    refill_ell_levels(M->low64);
    refill_ell_levels(M->up64);
    if (M->have32) {
        refresh_sweep_f32(ctx, M->low64, M->low32);
        refresh_sweep_f32(ctx, M->up64, M->up32);
    }
    PP_HIP(hipStreamSynchronize(ctx->stream));
    M->host_factor_stale = true;
    ++M->device_refactors;
}

// Times each staged-sweep load variant on one precision's sweeps (cold
// cache) and keeps the fastest for that precision.
template <class T>
int tune_sweep_pair(pp_ctx *ctx, pp_precond *M, DevSweep<T> &lo, DevSweep<T> &up, int start) {
    T *y = dalloc<T>(M->n);
    int best_var = start;
    try {
        launch_fill_zero(ctx, y, (size_t)M->n * sizeof(T));
        float best = 1e30f;
        const bool have16 = lo.ci16 || up.ci16;
        for (int var = 0; var < kVarCount; ++var) {
            if ((var & kVarIdx16) && !have16) continue;
            lo.variant = up.variant = var;
            const float ms = time_cold_ms(ctx, [&] {
                launch_sweep_pair<T>(ctx, lo, M->hlow.mode, up, M->hup.mode, M->fuse_mid, y, y);
            }, 3);
            if (ms < best) {
                best = ms;
                best_var = var;
            }
        }
        bool have_ell = false;
        for (const auto &e : lo.ell) have_ell |= e.ok;
        for (const auto &e : up.ell) have_ell |= e.ok;
        if (have_ell) {  // the sliced-column levels (the others keep the best staged variant)
            lo.variant = up.variant = kVarEll;
            lo.ell_fallback = up.ell_fallback = best_var;
            const float ms = time_cold_ms(ctx, [&] {
                launch_sweep_pair<T>(ctx, lo, M->hlow.mode, up, M->hup.mode, M->fuse_mid, y, y);
            }, 3);
            if (ms < best) {
                best = ms;
                best_var = kVarEll;
            }
        }
        lo.variant = up.variant = best_var;
    } catch (...) {
        dfree(y);
        throw;
    }
    dfree(y);
    return best_var;
}

void precond_need_f32(pp_precond *M) {
    if (M->opts.kind == PP_PC_AMSED || M->opts.kind == PP_PC_GMSLR) {
        ml_need_f32(M->ctx, M);
        return;
    }
    // every buffer an FP32 apply touches exists before any graph capture
    if (!M->order.empty() && !M->xp32) M->xp32 = dalloc<float>(M->n);
    if (M->trsv == PP_TRSV_SYNCFREE && !M->tmp32) M->tmp32 = dalloc<float>(M->n);
    if (M->have32) return;
    derive_sweep_f32(M->ctx, M->low64, M->low32);
    derive_sweep_f32(M->ctx, M->up64, M->up32);
    M->low32.variant = M->up32.variant = M->sweep_variant;
    if (M->trsv == PP_TRSV_LEVELS && M->n >= kTuneSweepRows && (M->low32.blocks || M->up32.blocks) && !M->sweep_pinned)
        tune_sweep_pair<float>(M->ctx, M, M->low32, M->up32, M->sweep_variant);
    if (M->chunks.nchunk) {
        upload_chunk_sweep(M->hlow, M->chunks, M->clow32);
        upload_chunk_sweep(M->hup, M->chunks, M->cup32);
    }
    M->have32 = true;
}

// P A P^T for a solve in the factor's row order (rebuilt when A or its values
// changed since the last build)
pp_csr *precond_permuted_matrix(pp_ctx *ctx, pp_precond *M, pp_csr *A) {
    if (M->Aperm && M->Aperm_src_uid == A->uid && M->Aperm_version == A->version) return M->Aperm;
    csr_free(M->Aperm);
    M->Aperm = nullptr;
    HostCsr H;
    H.n = A->nrow;
    H.rp = A->h_rp;
    H.ci = A->h_ci;
    H.v = A->h_v;
    M->Aperm = csr_create_host(ctx, permute_symmetric(H, M->order));  // the matrix changed since setup
    M->Aperm_src_uid = A->uid;
    M->Aperm_version = A->version;
    return M->Aperm;
}

// True when a solve may produce the forward sweep's first level itself
// (launch_level0 on the right-hand side) and call precond_apply_after_level0.
bool precond_level0_fusable(const pp_precond *M) {
    return M && M->ready && M->trsv == PP_TRSV_LEVELS && M->lvl0_contiguous && !M->multilevel &&
           M->opts.kind != PP_PC_NONE && (M->order.empty() || M->io_permuted);
}

// y[lvl0] = first forward level applied to x[lvl0]
void precond_level0(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y) {
    if (dtype == PP_F64)
        launch_level0<double>(ctx, M->lvl0_lo, M->lvl0_hi, M->hlow.mode, M->low64.diag, (const double *)x,
                              (double *)y);
    else
        launch_level0<float>(ctx, M->lvl0_lo, M->lvl0_hi, M->hlow.mode, M->low32.diag, (const float *)x,
                             (float *)y);
}

// y = M^{-1} x with the forward sweep's first level already in y
void precond_apply_after_level0(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y) {
    if (dtype == PP_F64)
        launch_sweep_pair<double>(ctx, M->low64, M->hlow.mode, M->up64, M->hup.mode, M->fuse_mid,
                                  (const double *)x, (double *)y, 1);
    else
        launch_sweep_pair<float>(ctx, M->low32, M->hlow.mode, M->up32, M->hup.mode, M->fuse_mid, (const float *)x,
                                 (float *)y, 1);
}

// y = M^{-1} x on the context stream (y may alias x).
void precond_apply(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y) {
    PP_CHECK(M->ready, PP_ERR_NOT_READY, "preconditioner used before setup()");
    const size_t esz = dtype == PP_F64 ? 8 : 4;
    if (M->opts.kind == PP_PC_NONE) {
        if (x != y) PP_HIP(hipMemcpyAsync(y, x, (size_t)M->n * esz, hipMemcpyDeviceToDevice, ctx->stream));
        return;
    }
    if (M->opts.kind == PP_PC_AMSED || M->opts.kind == PP_PC_GMSLR) {
        ml_apply(ctx, M, dtype, x, y);
        return;
    }
    if (!M->order.empty() && !M->io_permuted && x != (const void *)M->xp64 && x != (const void *)M->xp32) {
        // factor of P A P^T: y = P^T M^{-1} P x (the inner call sees x == xp)
        void *xp;
        if (dtype == PP_F64) {
            xp = M->xp64;
        } else {
            PP_CHECK(M->xp32, PP_ERR_HIP, "FP32 permutation buffer missing");  // from precond_need_f32
            xp = M->xp32;
        }
        launch_permute(ctx, M->n, M->dorder, dtype, x, xp, true);
        precond_apply(ctx, M, dtype, xp, xp);
        launch_permute(ctx, M->n, M->dorder, dtype, xp, y, false);
        return;
    }
    if (M->trsv == PP_TRSV_SYNCFREE) {
        // out of place: x -> tmp (lower) -> y (upper); y may alias x
        if (dtype == PP_F64) {
            PP_CHECK(M->tmp64, PP_ERR_HIP, "sync-free sweep buffer missing");  // allocated at setup
            launch_sweep_syncfree<double>(ctx, M->low64, M->hlow.mode, (const double *)x, M->tmp64);
            launch_sweep_syncfree<double>(ctx, M->up64, M->hup.mode, M->tmp64, (double *)y);
        } else {
            precond_need_f32(M);
            PP_CHECK(M->tmp32, PP_ERR_HIP, "sync-free sweep buffer missing");  // allocated by precond_need_f32
            launch_sweep_syncfree<float>(ctx, M->low32, M->hlow.mode, (const float *)x, M->tmp32);
            launch_sweep_syncfree<float>(ctx, M->up32, M->hup.mode, M->tmp32, (float *)y);
        }
        return;
    }
    if (M->trsv == PP_TRSV_BLOCK || M->trsv == PP_TRSV_PIPE) {
        const bool pipe = M->trsv == PP_TRSV_PIPE;
        if (dtype == PP_F64) {
            (pipe ? launch_sweep_pair_pipe<double> : launch_sweep_pair_chunks<double>)(
                ctx, M->chunks, M->clow64, M->hlow.mode, M->cup64, M->hup.mode, (const double *)x, (double *)y);
        } else {
            precond_need_f32(M);
            (pipe ? launch_sweep_pair_pipe<float> : launch_sweep_pair_chunks<float>)(
                ctx, M->chunks, M->clow32, M->hlow.mode, M->cup32, M->hup.mode, (const float *)x, (float *)y);
        }
        return;
    }
    if (dtype == PP_F64) {
        launch_sweep_pair<double>(ctx, M->low64, M->hlow.mode, M->up64, M->hup.mode, M->fuse_mid,
                                  (const double *)x, (double *)y);
    } else {
        precond_need_f32(M);
        launch_sweep_pair<float>(ctx, M->low32, M->hlow.mode, M->up32, M->hup.mode, M->fuse_mid,
                                 (const float *)x, (float *)y);
    }
}

}  // namespace pp

using namespace pp;

void set_host_error(const std::string &m);

template <class F>
static int guarded_pc(pp_ctx *ctx, F &&f) {
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

void pp_precond_opts_default(pp_precond_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->kind = PP_PC_ICT;
    o->tau = 1e-3;
    o->lfil = 20;
    o->trsv = PP_TRSV_AUTO;
    o->levels = 3;
    o->ndom = 16;
    o->rank = 32;
    o->eps = 0.8;
}

int pp_precond_create(pp_ctx *ctx, pp_csr *A, const pp_precond_opts *opts, pp_precond **out) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_pc(ctx, [&] {
        PP_CHECK(A && opts && out, PP_ERR_INVALID, "NULL argument");
        PP_CHECK(A->nrow == A->ncol, PP_ERR_INVALID, "preconditioner needs a square matrix");
        auto *M = new pp_precond;
        M->ctx = ctx;
        M->uid = next_uid();
        M->A = A;
        M->opts = *opts;
        M->n = A->nrow;
        *out = M;
    });
}

}  // extern "C"

// Times the staged-sweep load variants (ordinary / non-temporal loads, 32 /
// 16-bit dependency offsets) on one FP64 apply and keeps the fastest; the
// FP32 sweeps use the same choice.  Only when a level is wide enough to use
// the staged kernel and the factor is large.
static void tune_sweeps(pp_ctx *ctx, pp_precond *M) {
    if (M->n < kTuneSweepRows || (!M->low64.blocks && !M->up64.blocks)) return;
    M->sweep_variant = tune_sweep_pair<double>(ctx, M, M->low64, M->up64, M->sweep_variant);
}

// PP_TRSV_AUTO: the fastest of one launch per level and, when the rows chunk
// within LDS, the chunked one-workgroup schedules (barrier per level, and
// the pipelined form when its rings fit), timed here over a few applies
// each.  PP_TRSV_BLOCK / PP_TRSV_PIPE asked for explicitly (force) fall
// back to levels when the rows do not chunk, and PIPE to BLOCK when the
// rings do not fit.
static int choose_schedule(pp_ctx *ctx, pp_precond *M, int force) {
    const uint32_t nlev = M->low64.nlev + M->up64.nlev;
    if ((!force && nlev < 8) || !build_block_sweeps(ctx, M)) return PP_TRSV_LEVELS;
    uint32_t re = 0, rr = 0;
    const bool pipe_ok = sweep_pipe_rings<double>(ctx, M->chunks, re, rr);
    if (force == PP_TRSV_BLOCK || (force == PP_TRSV_PIPE && !pipe_ok)) return PP_TRSV_BLOCK;
    if (force == PP_TRSV_PIPE) return PP_TRSV_PIPE;
    double *v = dalloc<double>(M->n);
    const int cand[3] = {PP_TRSV_LEVELS, PP_TRSV_BLOCK, PP_TRSV_PIPE};
    float ms[3] = {0.f, 0.f, 0.f};
    const int ncand = pipe_ok ? 3 : 2;
    try {
        launch_fill_zero(ctx, v, (size_t)M->n * sizeof(double));
        launch_fill_zero(ctx, ctx->ctl, C_COUNT * sizeof(int));
        for (int k = 0; k < ncand; ++k) {
            auto apply = [&] {
                if (cand[k] == PP_TRSV_LEVELS)
                    launch_sweep_pair<double>(ctx, M->low64, M->hlow.mode, M->up64, M->hup.mode, M->fuse_mid, v, v);
                else if (cand[k] == PP_TRSV_BLOCK)
                    launch_sweep_pair_chunks<double>(ctx, M->chunks, M->clow64, M->hlow.mode, M->cup64, M->hup.mode,
                                                     v, v);
                else
                    launch_sweep_pair_pipe<double>(ctx, M->chunks, M->clow64, M->hlow.mode, M->cup64, M->hup.mode,
                                                   v, v);
            };
            apply();
            PP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
            for (int r = 0; r < 3; ++r) apply();
            PP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
            PP_HIP(hipEventSynchronize(ctx->ev1));
            PP_HIP(hipEventElapsedTime(&ms[k], ctx->ev0, ctx->ev1));
        }
        if (ncand == 3) {  // a pipelined sweep that gave up waiting is not a candidate
            int ctl[C_COUNT];
            read_ctl(ctx, ctl);
            if (ctl[C_SPIN]) {
                ms[2] = 1e30f;
                launch_fill_zero(ctx, ctx->ctl, C_COUNT * sizeof(int));
            }
        }
    } catch (...) {
        dfree(v);
        throw;
    }
    dfree(v);
    int best = 0;
    for (int k = 1; k < ncand; ++k)
        if (ms[k] < ms[best]) best = k;
    if (cand[best] != PP_TRSV_LEVELS) return cand[best];
    M->chunks.free_all();  // not used: give the memory back
    M->clow64.free_all();
    M->cup64.free_all();
    return PP_TRSV_LEVELS;
}

// IC(0) or ILU(0) in multicolour order on a matrix large enough for the
// device work to pay (and not PP_SETUP_HOST=1 in the environment: the host
// setup the device one is tested against)
static bool device_setup_applies(const pp_precond *M) {
    const char *host = std::getenv("PP_SETUP_HOST");
    const int t = M->opts.trsv;
    return (M->opts.kind == PP_PC_IC0 || M->opts.kind == PP_PC_ILU0) && M->opts.ordering == PP_ORDER_MULTICOLOR &&
           (t == PP_TRSV_AUTO || t == PP_TRSV_LEVELS || t == PP_TRSV_SYNCFREE) && M->A->nrow >= (1u << 15) &&
           !(host && host[0] == '1');
}

extern "C" {

int pp_precond_setup(pp_ctx *ctx, pp_precond *M) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_pc(ctx, [&] {
        PP_CHECK(M, PP_ERR_INVALID, "NULL preconditioner");
        if (M->ready) return;  // setup() is idempotent (IncompleteCholesky.cpp:105)
        PoolScope pool(ctx);
        auto t0 = std::chrono::steady_clock::now();
        if (M->opts.kind == PP_PC_NONE) {
            M->mnnz = 0;
        } else if (M->opts.kind == PP_PC_AMSED || M->opts.kind == PP_PC_GMSLR) {
            ml_setup(ctx, M);
        } else if (device_setup_applies(M) && setup_mc_device(ctx, M)) {
            // on the device (pp_setup_dev.hip): the host structures wait for a caller that needs them
            const bool ic = M->opts.kind == PP_PC_IC0;
            M->host_pending = true;
            M->dev_setup = true;
            M->hlow = HostSweep();
            M->hlow.mode = ic ? HostSweep::DIVIDE : HostSweep::UNIT;
            M->hup = HostSweep();
            M->hup.lower = false;
            M->hup.mode = ic ? HostSweep::DIVIDE : HostSweep::MULTIPLY;
            for (auto *v : {&M->lt_rp, &M->lt_ci, &M->l_rp, &M->l_ci, &M->u_rp, &M->u_ci}) v->clear();
            for (auto *v : {&M->lt_v, &M->l_v, &M->u_v, &M->d}) v->clear();
            M->fail_row = 0;
            finish_device_sweeps(M);
            auto tp = std::chrono::steady_clock::now();
            M->phase_ms[2] = (float)(std::chrono::duration<double, std::milli>(tp - t0).count() - M->phase_ms[0] -
                                     M->phase_ms[1]);
            const int want = M->opts.trsv;
            M->trsv = want == PP_TRSV_AUTO ? choose_schedule(ctx, M, want) : want;
            if (M->trsv == PP_TRSV_BLOCK || M->trsv == PP_TRSV_PIPE) {
                delete M->dref;  // the re-factorization does not refill the chunked sweeps
                M->dref = nullptr;
            }
            if (M->trsv == PP_TRSV_LEVELS) tune_sweeps(ctx, M);
            if (M->trsv == PP_TRSV_SYNCFREE && !M->tmp64) M->tmp64 = dalloc<double>(M->n);
            PP_HIP(hipStreamSynchronize(ctx->stream));
            M->phase_ms[3] = (float)ms_since(tp);
        } else {
            M->host_pending = false;
            M->dev_setup = false;
            // the matrix's host arrays are lent to the factorization, not copied
            HostCsr A, Ap;
            A.n = M->A->nrow;
            struct Lend {
                HostCsr &A;
                pp_csr *src;
                Lend(HostCsr &a, pp_csr *s) : A(a), src(s) { swap(); }
                ~Lend() { swap(); }
                void swap() {
                    A.rp.swap(src->h_rp);
                    A.ci.swap(src->h_ci);
                    A.v.swap(src->h_v);
                }
            };
            {
                Lend lend(A, M->A);
                setup_triangular(M, A, true, &Ap);
            }
            // P A P^T on the device for the solves (they run on the colour-ordered system)
            if (!Ap.rp.empty()) {
                csr_free(M->Aperm);
                M->Aperm = nullptr;
                M->Aperm = csr_create_host(ctx, std::move(Ap));
                M->Aperm_src_uid = M->A->uid;
                M->Aperm_version = M->A->version;
            }
            auto tp = std::chrono::steady_clock::now();
            M->phase_ms[2] = (float)(std::chrono::duration<double, std::milli>(tp - t0).count() - M->phase_ms[0] -
                                     M->phase_ms[1]);
            const int want = M->opts.trsv;
            M->trsv = want == PP_TRSV_AUTO || want == PP_TRSV_BLOCK || want == PP_TRSV_PIPE
                          ? choose_schedule(ctx, M, want)
                          : want;
            if (M->trsv == PP_TRSV_LEVELS) tune_sweeps(ctx, M);
            // the sync-free apply's buffer between its two sweeps: allocated here, so that no
            // apply allocates (a solve may capture its first apply into a graph)
            if (M->trsv == PP_TRSV_SYNCFREE && !M->tmp64) M->tmp64 = dalloc<double>(M->n);
            PP_HIP(hipStreamSynchronize(ctx->stream));
            M->phase_ms[3] = (float)ms_since(tp);
        }
        M->ready = 1;
        M->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int pp_precond_refactor(pp_ctx *ctx, pp_precond *M) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_pc(ctx, [&] {
        PP_CHECK(M, PP_ERR_INVALID, "NULL preconditioner");
        PP_CHECK(M->ready || M->dref, PP_ERR_NOT_READY, "refactor() before setup()");
        auto t0 = std::chrono::steady_clock::now();
        if (M->dref || devref_build(ctx, M)) {
            devref_run(ctx, M);
        } else {
            // the host path: everything setup() builds, from the current values
            M->ready = 0;
            M->low64.free_all();
            M->up64.free_all();
            M->low32.free_all();
            M->up32.free_all();
            M->have32 = false;
            dfree(M->dorder);
            M->dorder = nullptr;
            dfree(M->xp64);
            M->xp64 = nullptr;
            M->chunks.free_all();
            M->clow64.free_all();
            M->cup64.free_all();
            M->clow32.free_all();
            M->cup32.free_all();
            if (M->multilevel) ml_free(M);
            reduced_free(M);
            M->uid = next_uid();  // graphs captured with the freed arrays are never replayed
            const int st = pp_precond_setup(ctx, M);
            if (st != PP_OK) throw Error(st, ctx->err);
            if (M->opts.make_f32) precond_need_f32(M);
            M->host_factor_stale = false;
        }
        M->ready = 1;
        M->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

int pp_precond_apply(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_pc(ctx, [&] {
        PP_CHECK(M && x && y, PP_ERR_INVALID, "NULL argument");
        PP_CHECK(M->ready, PP_ERR_NOT_READY, "preconditioner used before setup()");
        if (dtype == PP_F32) precond_need_f32(M);  // FP32 copies, once, outside any capture
        launch_fill_zero(ctx, ctx->ctl, C_COUNT * sizeof(int));  // clear stop / error words (graph-safe)
        precond_apply(ctx, M, dtype, x, y);
        if (M->trsv == PP_TRSV_SYNCFREE || M->trsv == PP_TRSV_PIPE) {
            int ctl[C_COUNT];
            read_ctl(ctx, ctl);
            PP_CHECK(!ctl[C_SPIN], PP_ERR_HIP, "sync-free triangular sweep timed out waiting for a dependency");
        }
    });
}

int pp_precond_apply_part(pp_ctx *ctx, pp_precond *M, int part, int dtype, void *xuser) {
    if (!ctx) return PP_ERR_INVALID;
    return guarded_pc(ctx, [&] {
        PP_CHECK(M && xuser, PP_ERR_INVALID, "NULL argument");
        PP_CHECK(M->ready, PP_ERR_NOT_READY, "preconditioner used before setup()");
        PP_CHECK(M->opts.kind != PP_PC_AMSED && M->opts.kind != PP_PC_GMSLR && M->opts.kind != PP_PC_NONE,
                 PP_ERR_UNSUPPORTED, "not a triangular preconditioner");
        launch_fill_zero(ctx, ctx->ctl, sizeof(int));
        const bool ildlt = M->opts.kind == PP_PC_ILDLT;
        if (dtype == PP_F32) precond_need_f32(M);
        // a reordered factor works on P x (permuted in, permuted back out)
        void *x = xuser;
        if (!M->order.empty()) {
            x = dtype == PP_F64 ? (void *)M->xp64 : (void *)M->xp32;
            launch_permute(ctx, M->n, M->dorder, dtype, xuser, x, true);
        }
        struct Back {
            pp_ctx *c;
            pp_precond *m;
            int dt;
            void *from, *to;
            ~Back() {
                if (from != to) {
                    try {
                        launch_permute(c, m->n, m->dorder, dt, from, to, false);
                    } catch (...) {
                    }
                }
            }
        } back{ctx, M, dtype, x, xuser};
        if (part == 0) {
            if (dtype == PP_F64) launch_sweep<double>(ctx, M->low64, M->hlow.mode, (double *)x, (double *)x);
            else launch_sweep<float>(ctx, M->low32, M->hlow.mode, (float *)x, (float *)x);
        } else if (part == 1) {
            // D^{-1} alone: only ILDLT has a diagonal stage (identity otherwise)
            if (ildlt) {
                HostSweep diag;
                diag.n = M->n;
                diag.lower = true;
                diag.mode = HostSweep::PREDIVIDE;
                diag.rp.assign(M->n + 1, 0);
                diag.diag = M->d;
                if (dtype == PP_F64) {
                    DevSweep<double> ds;
                    upload_sweep(diag, ds);
                    launch_sweep<double>(ctx, ds, HostSweep::PREDIVIDE, (double *)x, (double *)x);
                    PP_HIP(hipStreamSynchronize(ctx->stream));
                    ds.free_all();
                } else {
                    DevSweep<float> ds;
                    upload_sweep(diag, ds);
                    launch_sweep<float>(ctx, ds, HostSweep::PREDIVIDE, (float *)x, (float *)x);
                    PP_HIP(hipStreamSynchronize(ctx->stream));
                    ds.free_all();
                }
            }
        } else if (part == 2) {
            // the ILDLT upper sweep folds D^{-1} in; alone it must not
            int mode = ildlt ? HostSweep::UNIT : M->hup.mode;
            if (dtype == PP_F64) launch_sweep<double>(ctx, M->up64, mode, (double *)x, (double *)x);
            else launch_sweep<float>(ctx, M->up32, mode, (float *)x, (float *)x);
        } else {
            throw Error(PP_ERR_INVALID, "part must be 0, 1 or 2");
        }
    });
}

// Bytes one sweep streams: values, dependency indices, row pointers, the
// level row map where read, the diagonal, the block tables, and the vectors
// (right-hand side read, solution written, dependency values read once).
}  // extern "C"
template <class T>
static uint64_t sweep_bytes(const DevSweep<T> &d) {
    const uint64_t esz = sizeof(T);
    if (d.variant == kVarEll) {
        // sliced-column levels: padded values and 16-bit offsets; the other
        // levels as the staged kernel with the default loads streams them
        uint64_t b = 0, nnz_e = 0, rows_e = 0;
        for (const auto &E : d.ell)
            if (E.ok) {
                b += (esz + 2) * (uint64_t)E.width * E.nrows;
                nnz_e += E.entries;
                rows_e += E.nrows;
            }
        const uint64_t nnz_s = d.nnz - nnz_e;
        b += (esz + 2) * nnz_s + 4ull * (d.n - rows_e + 1) + 4 * d.perm_rows;
        if (d.diag) b += esz * d.n;
        return b + 3 * esz * d.n;
    }
    const uint64_t n16 = (d.variant & kVarIdx16) ? d.nnz16 : 0;
    uint64_t b = esz * d.nnz + 2 * n16 + 4 * (d.nnz - n16) + 4ull * (d.n + 1) + 4 * d.perm_rows;
    if (d.diag) b += esz * d.n;
    if (d.nblk) b += 4 * (d.nblk + 1) + (n16 ? 8 * d.nblk : 0);
    return b + 3 * esz * d.n;
}
extern "C" {

int pp_precond_apply_bytes(const pp_precond *M, int dtype, int in_solve, uint64_t *bytes) {
    if (!M || !bytes) return PP_ERR_INVALID;
    if (!M->ready) return PP_ERR_NOT_READY;
    if (M->opts.kind == PP_PC_AMSED || M->opts.kind == PP_PC_GMSLR || M->opts.kind == PP_PC_NONE)
        return PP_ERR_UNSUPPORTED;
    uint64_t b;
    if (dtype == PP_F64) {
        b = sweep_bytes(M->low64) + sweep_bytes(M->up64);
    } else {
        if (!M->have32) return PP_ERR_NOT_READY;
        b = sweep_bytes(M->low32) + sweep_bytes(M->up32);
    }
    const uint64_t esz = dtype == PP_F64 ? 8 : 4;
    if (!M->order.empty() && !in_solve) b += 2 * (4ull * M->n + 2 * esz * M->n);  // gather in, scatter out
    if (M->trsv != PP_TRSV_SYNCFREE && M->fuse_mid)  // the fused level: no second row pointer, read and write
        b -= (uint64_t)M->fuse_rows * (4 + 2 * esz);
    *bytes = b;
    return PP_OK;
}

int pp_precond_info_get(const pp_precond *M, pp_precond_info *info) {
    if (!M || !info) return PP_ERR_INVALID;
    std::memset(info, 0, sizeof(*info));
    info->kind = M->opts.kind;
    info->ready = M->ready;
    info->n = M->n;
    info->annz = M->A->nnz;
    info->mnnz = M->mnnz;
    info->fill_ratio = M->A->nnz ? (double)M->mnnz / (double)M->A->nnz : 0.0;
    info->levels_lower = M->low64.nlev;
    info->levels_upper = M->up64.nlev;
    info->trsv = M->trsv;
    if (const pp_precond *B = ml_block_factor(M)) {
        // multilevel: the sweeps of the first level's block factor (the subdomains)
        info->levels_lower = B->low64.nlev;
        info->levels_upper = B->up64.nlev;
        info->trsv = B->trsv;
    }
    info->setup_ms = M->setup_ms;
    info->fail_row = M->fail_row;
    info->sweep_variant = M->sweep_variant;
    info->fused_levels = M->fuse_mid ? 1 : 0;
    for (int k = 0; k < 4; ++k) info->phase_ms[k] = M->phase_ms[k];
    info->device_refactors = M->device_refactors;
    info->setup_on_device = M->dev_setup ? 1 : 0;
    if (M->reduced && M->reduced->ok) {
        info->reduced_form = M->reduced->built;
        info->reduced_form_ms = M->reduced->build_ms;
    }
    return PP_OK;
}

int pp_precond_set_sweep_variant(pp_precond *M, int variant) {
    if (!M || variant < 0 || variant > kVarEll) return PP_ERR_INVALID;
    if (!M->ready || M->multilevel) return PP_ERR_NOT_READY;
    bool have16 = M->low64.ci16 || M->up64.ci16, have_ell = false;
    for (const auto &e : M->low64.ell) have_ell |= e.ok;
    for (const auto &e : M->up64.ell) have_ell |= e.ok;
    if (variant == kVarEll ? !have_ell : ((variant & kVarIdx16) && !have16)) return PP_ERR_INVALID;
    M->sweep_variant = variant;
    M->sweep_pinned = true;
    // variant 8: the levels without a sliced-column form run the staged kernel with the default loads
    for (DevSweep<double> *d : {&M->low64, &M->up64}) {
        d->variant = variant;
        if (variant == kVarEll) d->ell_fallback = kVarDefault;
    }
    for (DevSweep<float> *d : {&M->low32, &M->up32}) {
        d->variant = variant;
        if (variant == kVarEll) d->ell_fallback = kVarDefault;
    }
    return PP_OK;
}

// The host copy of a factor last computed on the device: its entries
// downloaded and placed through the maps the device state keeps.
static void refresh_host_factor(pp_precond *M) {
    DevRefactor &R = *M->dref;
    std::vector<double> v((size_t)R.nnz + R.n);
    PP_HIP(hipMemcpy(v.data(), R.val, v.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (M->opts.kind == PP_PC_IC0) {
        for (size_t q = 0; q < M->lt_v.size(); ++q) M->lt_v[q] = v[R.h_lt_map[q]];
    } else {
        for (size_t q = 0; q < M->l_v.size(); ++q) M->l_v[q] = v[R.h_l_map[q]];
        for (size_t q = 0; q < M->u_v.size(); ++q) M->u_v[q] = v[R.h_u_map[q]];
        M->d.assign(v.begin() + R.nnz, v.end());
    }
    M->host_factor_stale = false;
}

int pp_precond_factor(const pp_precond *M, int which, uint32_t *nnz, uint32_t *rowptr, uint32_t *colidx,
                      double *values, double *d) {
    if (!M || !nnz) return PP_ERR_INVALID;
    if (M->host_pending || (M->host_factor_stale && M->dref)) {
        try {
            PP_HIP(hipSetDevice(M->ctx->device));
            if (M->host_pending) ensure_host_setup(const_cast<pp_precond *>(M));
            else refresh_host_factor(const_cast<pp_precond *>(M));
        } catch (const Error &e) {
            set_host_error(e.what());
            return e.status;
        }
    }
    const std::vector<uint32_t> *rp, *ci;
    const std::vector<double> *v;
    if (which == 0) { rp = &M->lt_rp; ci = &M->lt_ci; v = &M->lt_v; }
    else if (which == 1) { rp = &M->l_rp; ci = &M->l_ci; v = &M->l_v; }
    else if (which == 2) { rp = &M->u_rp; ci = &M->u_ci; v = &M->u_v; }
    else return PP_ERR_INVALID;
    if (rp->empty()) return PP_ERR_NOT_READY;
    *nnz = (uint32_t)ci->size();
    if (rowptr) std::memcpy(rowptr, rp->data(), rp->size() * sizeof(uint32_t));
    if (colidx && !ci->empty()) std::memcpy(colidx, ci->data(), ci->size() * sizeof(uint32_t));
    if (values && !v->empty()) std::memcpy(values, v->data(), v->size() * sizeof(double));
    if (d && !M->d.empty()) std::memcpy(d, M->d.data(), M->d.size() * sizeof(double));
    return PP_OK;
}

void pp_precond_destroy(pp_precond *M) {
    if (!M) return;
    (void)hipSetDevice(M->ctx->device);
    (void)hipStreamSynchronize(M->ctx->stream);
    dfree(M->dorder);
    dfree(M->xp64);
    dfree(M->xp32);
    dfree(M->tmp64);
    dfree(M->tmp32);
    csr_free(M->Aperm);
    pp::reduced_free(M);
    M->chunks.free_all();
    M->clow64.free_all();
    M->cup64.free_all();
    M->clow32.free_all();
    M->cup32.free_all();
    M->low64.free_all();
    M->up64.free_all();
    M->low32.free_all();
    M->up32.free_all();
    delete M->dref;
    if (M->multilevel) ml_free(M);
    delete M;
}

}  // extern "C"

// ---------------------------------------------------------------------------- host-only factorization
struct pp_hfactor {
    pp_precond M;
    int status = PP_OK;
};

extern "C" {

int pp_factor_host(const pp_precond_opts *opts, uint32_t n, const uint32_t *rp, const uint32_t *ci, const double *v,
                   pp_hfactor **out) {
    if (!opts || !rp || !out) return PP_ERR_INVALID;
    auto *F = new pp_hfactor;
    *out = F;
    F->M.opts = *opts;
    F->M.n = n;
    try {
        HostCsr A;
        A.n = n;
        A.rp.assign(rp, rp + n + 1);
        A.ci.assign(ci, ci + rp[n]);
        A.v.assign(v, v + rp[n]);
        PP_CHECK(opts->kind >= PP_PC_ICT && opts->kind <= PP_PC_ILU0, PP_ERR_UNSUPPORTED,
                 "host factorization covers ICT, ILUT, ILDLT, IC0, ILU0");
        setup_triangular(&F->M, A, false);
        F->M.ready = 1;
    } catch (const Error &e) {
        F->status = e.status;
        set_host_error(e.what());
    }
    return F->status;
}

int pp_hfactor_get(const pp_hfactor *F, int which, uint32_t *nnz, uint32_t *rp, uint32_t *ci, double *v, double *d) {
    if (!F) return PP_ERR_INVALID;
    return pp_precond_factor(&F->M, which, nnz, rp, ci, v, d);
}

int pp_hfactor_info(const pp_hfactor *F, uint32_t *mnnz, uint32_t *fail_row, uint32_t *lev_lo, uint32_t *lev_up) {
    if (!F) return PP_ERR_INVALID;
    if (mnnz) *mnnz = F->M.mnnz;
    if (fail_row) *fail_row = F->M.fail_row;
    std::vector<uint32_t> perm, lp;
    if (lev_lo) {
        *lev_lo = 0;
        if (F->M.hlow.n) {
            levels_of(F->M.hlow, perm, lp);
            *lev_lo = (uint32_t)lp.size() - 1;
        }
    }
    if (lev_up) {
        *lev_up = 0;
        if (F->M.hup.n) {
            levels_of(F->M.hup, perm, lp);
            *lev_up = (uint32_t)lp.size() - 1;
        }
    }
    return PP_OK;
}

void pp_hfactor_destroy(pp_hfactor *F) { delete F; }

}  // extern "C"
