// pp_core.cpp -- context, device memory, CSR handles, SpMV/BLAS-1 entry points,
// model-problem generators and seeded inputs of the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

#include "pp_internal.h"

namespace {
std::mutex g_err_mu;
std::string g_err;

void set_global_error(const std::string &m) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = m;
}
}  // namespace

void set_host_error(const std::string &m) { set_global_error(m); }

// Runs f, turning exceptions into a status and a message on ctx.
template <class F>
static int guarded(pp_ctx *ctx, F &&f) {
    try {
        if (ctx) pp::set_device(ctx);
        f();
        return PP_OK;
    } catch (const pp::Error &e) {
        if (ctx) ctx->err = e.what();
        else set_global_error(e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->err = "host allocation failed";
        else set_global_error("host allocation failed");
        return PP_ERR_NOMEM;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        else set_global_error(e.what());
        return PP_ERR_INVALID;
    }
}

namespace pp {

void set_device(pp_ctx *ctx) { PP_HIP(hipSetDevice(ctx->device)); }

static thread_local pp_ctx *tl_pool = nullptr;
PoolScope::PoolScope(pp_ctx *ctx) : prev(tl_pool) { tl_pool = ctx; }
PoolScope::~PoolScope() { tl_pool = prev; }

static void *pool_take(pp_ctx *ctx, size_t bytes) {
    auto it = ctx->pool.lower_bound(bytes);
    if (it == ctx->pool.end() || it->first > bytes + bytes / 4) return nullptr;
    if (ctx->pool_dirty) {
        PP_HIP(hipStreamSynchronize(ctx->stream));
        ctx->pool_dirty = false;
    }
    void *p = it->second;
    ctx->pool_bytes -= it->first;
    ctx->pool.erase(it);
    return p;
}

void pool_put(pp_ctx *ctx, void *p, size_t bytes) {
    if (!p) return;
    if (ctx->pool_bytes + bytes > kPoolCap) {
        dfree(p);
        return;
    }
    ctx->pool.emplace(bytes, p);
    ctx->pool_bytes += bytes;
    ctx->pool_dirty = true;
}

void pool_release(pp_ctx *ctx) {
    for (auto &e : ctx->pool) dfree(e.second);
    ctx->pool.clear();
    ctx->pool_bytes = 0;
    ctx->pool_dirty = false;
}

template <class T>
T *dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    if (tl_pool)
        if (void *q = pool_take(tl_pool, n * sizeof(T))) return (T *)q;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) throw Error(PP_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return (T *)p;
}
template double *dalloc<double>(size_t);
template float *dalloc<float>(size_t);
template uint32_t *dalloc<uint32_t>(size_t);
template int *dalloc<int>(size_t);
template unsigned char *dalloc<unsigned char>(size_t);
template BlockDesc *dalloc<BlockDesc>(size_t);
template uint4 *dalloc<uint4>(size_t);
template int16_t *dalloc<int16_t>(size_t);
template unsigned long long *dalloc<unsigned long long>(size_t);

void dfree(void *p) {
    if (p) (void)hipFree(p);
}

void ensure_hist(pp_ctx *ctx, uint32_t cap) {
    if (ctx->hist_cap >= cap) return;
    dfree(ctx->hist);
    ctx->hist = dalloc<double>(cap);
    ctx->hist_cap = cap;
}

void *ws_get(pp_ctx *ctx, int idx, size_t bytes) {
    PP_CHECK(idx >= 0 && idx < 48, PP_ERR_INVALID, "workspace index");
    if (ctx->arena) {  // a zero-allocation call: slots are carved from the fixed scratch array
        if (ctx->arena_len[idx] < bytes) {
            const size_t off = (ctx->arena_used + 255) & ~(size_t)255;
            PP_CHECK(off <= ctx->arena_cap && bytes <= ctx->arena_cap - off, PP_ERR_INVALID,
                     "the solve's work vectors exceed the zero-allocation scratch array");
            ctx->arena_off[idx] = off;
            ctx->arena_len[idx] = bytes;
            ctx->arena_used = off + bytes;
        }
        return ctx->arena + ctx->arena_off[idx];
    }
    if (ctx->ws_bytes[idx] < bytes) {
        PP_HIP(hipStreamSynchronize(ctx->stream));
        dfree(ctx->ws[idx]);
        ctx->ws[idx] = dalloc<unsigned char>(bytes);
        ctx->ws_bytes[idx] = bytes;
    }
    return ctx->ws[idx];
}

double read_slot(pp_ctx *ctx, int slot) {
    PP_HIP(hipMemcpyAsync(ctx->pinned, ctx->scal + slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    return ctx->pinned[0];
}

void read_ctl(pp_ctx *ctx, int *out) {
    PP_HIP(hipMemcpyAsync(ctx->pinned_ctl, ctx->ctl, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(out, ctx->pinned_ctl, C_COUNT * sizeof(int));
}

void Graph::reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (g) (void)hipGraphDestroy(g);
    exec = nullptr;
    g = nullptr;
}

static void csr_upload_values_f32(pp_ctx *ctx, pp_csr *A) {
    if (!A->v32) A->v32 = dalloc<float>(A->nnz);
    launch_convert(ctx, A->nnz, PP_F64, A->v64, PP_F32, A->v32);
    A->f32_dirty = false;
}

static void csr_tune_spmv(pp_ctx *ctx, pp_csr *A, int dtype);

void csr_need_f32(pp_ctx *ctx, pp_csr *A) {
    if (A->f32_dirty || !A->v32) {
        csr_upload_values_f32(ctx, A);
        if (A->ell.ok) {
            if (!A->ell.v32) A->ell.v32 = dalloc<float>((size_t)A->ell.width * A->ell.n);
            if (A->ell.dg64 && !A->ell.dg32) A->ell.dg32 = dalloc<float>(A->ell.n);
            ell_fill_values(ctx, A->ell, PP_F32, A->rp, A->v32);
        }
    }
    if (!A->tuned32) {  // the FP32 loads are timed on their own (their best variant differs)
        A->tuned32 = true;
        csr_tune_spmv(ctx, A, PP_F32);
    }
}


uint64_t next_uid() {
    static std::atomic<uint64_t> counter{1};
    return counter.fetch_add(1);
}

void parallel_for(size_t n, size_t grain, const std::function<void(size_t, size_t)> &fn) {
    const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t parts = std::min(hw, grain ? std::max<size_t>(1, n / grain) : hw);
    if (parts <= 1) {
        if (n) fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(parts - 1);
    std::exception_ptr err;
    std::mutex m;
    auto run = [&](size_t b, size_t e) {
        try {
            fn(b, e);
        } catch (...) {
            std::lock_guard<std::mutex> g(m);
            if (!err) err = std::current_exception();
        }
    };
    for (size_t p = 1; p < parts; ++p) th.emplace_back(run, n * p / parts, n * (p + 1) / parts);
    run(0, n / parts);
    for (auto &t : th) t.join();
    if (err) std::rethrow_exception(err);
}

uint64_t two_base_blocks(const std::vector<uint32_t> &blk, const std::vector<uint32_t> &rp, const uint32_t *ci,
                         std::vector<uint16_t> &off16, std::vector<uint32_t> &base) {
    const size_t nb = blk.size() > 1 ? blk.size() - 1 : 0;
    base.assign(2 * nb, kNoBase);
    std::atomic<uint64_t> n16{0};
    parallel_for(nb, 256, [&](size_t b0, size_t b1) {
        uint64_t local = 0;
        for (size_t j = b0; j < b1; ++j) {
            const uint32_t e0 = rp[blk[j]], e1 = rp[blk[j + 1]];
            if (e1 == e0 || e1 - e0 > kSpmvNnzBlock) continue;
            uint32_t lo = 0, hi = 0;
            if (!two_base_offsets(ci + e0, e1 - e0, lo, hi, off16.data() + e0)) continue;
            base[2 * j] = lo;
            base[2 * j + 1] = hi;
            local += e1 - e0;
        }
        n16 += local;
    });
    return n16.load();
}

std::vector<BlockDesc> block_descs(const std::vector<uint32_t> &blk, const std::vector<uint32_t> &rp,
                                   const std::vector<uint32_t> &base) {
    std::vector<BlockDesc> d(blk.size());
    for (size_t j = 0; j < blk.size(); ++j) {
        const bool has = j + 1 < blk.size() && 2 * j + 1 < base.size();
        d[j] = {blk[j], rp[blk[j]], has ? base[2 * j] : kNoBase, has ? base[2 * j + 1] : kNoBase};
    }
    return d;
}

bool two_base_offsets(const uint32_t *cols, uint32_t cnt, uint32_t &base_lo, uint32_t &base_hi, uint16_t *out) {
    if (cnt == 0) return false;
    std::vector<uint32_t> u(cols, cols + cnt);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    // split at the widest gap between consecutive distinct columns
    size_t split = u.size();  // first column of the high window
    uint32_t widest = 0;
    for (size_t i = 1; i < u.size(); ++i)
        if (u[i] - u[i - 1] > widest) {
            widest = u[i] - u[i - 1];
            split = i;
        }
    if (u.back() - u.front() <= kOffMask) split = u.size();  // one window suffices
    base_lo = u.front();
    if (u[split - 1] - base_lo > kOffMask) return false;
    base_hi = split < u.size() ? u[split] : base_lo;
    if (split < u.size() && u.back() - base_hi > kOffMask) return false;
    for (uint32_t k = 0; k < cnt; ++k) {
        const bool hi = split < u.size() && cols[k] >= base_hi;
        out[k] = hi ? (uint16_t)(kHiBase | (cols[k] - base_hi)) : (uint16_t)(cols[k] - base_lo);
    }
    return true;
}

// The device CSR of a host matrix whose arrays it takes over; tune: time
// the SpMV load variants on it (for matrices in an iteration's hot loop):
// 0 no, 1 with cold caches, 2 with warm caches.
pp_csr *csr_from_vectors(pp_ctx *ctx, uint32_t nrow, uint32_t ncol, std::vector<uint32_t> &&rp,
                         std::vector<uint32_t> &&ci, std::vector<double> &&v, int tune) {
    PP_CHECK(rp.size() == (size_t)nrow + 1, PP_ERR_INVALID, "rowptr size");
    PP_CHECK(rp[0] == 0, PP_ERR_INVALID, "rowptr[0] must be 0");
    const uint32_t nnz = rp[nrow];
    PP_CHECK(ci.size() == nnz && v.size() == nnz, PP_ERR_INVALID, "colidx/values size");
    for (uint32_t i = 0; i < nrow; ++i) PP_CHECK(rp[i] <= rp[i + 1], PP_ERR_INVALID, "rowptr not monotone");
    parallel_for(nnz, 1u << 16, [&](size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; ++k) PP_CHECK(ci[k] < ncol, PP_ERR_INVALID, "column index out of range");
    });
    std::unique_ptr<pp_csr> A(new pp_csr);
    A->ctx = ctx;
    A->uid = next_uid();
    A->nrow = nrow;
    A->ncol = ncol;
    A->nnz = nnz;
    A->h_rp = std::move(rp);
    A->h_ci = std::move(ci);
    A->h_v = std::move(v);
    A->rp = dalloc<uint32_t>(nrow + 1);
    A->ci = dalloc<uint32_t>(nnz);
    A->v64 = dalloc<double>(nnz);
    PP_HIP(hipMemcpyAsync(A->rp, A->h_rp.data(), (nrow + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                          ctx->stream));
    if (nnz) {
        PP_HIP(hipMemcpyAsync(A->ci, A->h_ci.data(), nnz * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipMemcpyAsync(A->v64, A->h_v.data(), nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    std::vector<uint32_t> blocks;
    build_row_blocks(A->h_rp, blocks);
    A->nblocks = (uint32_t)blocks.size() - 1;
    A->row_blocks = dalloc<uint32_t>(blocks.size());
    PP_HIP(hipMemcpyAsync(A->row_blocks, blocks.data(), blocks.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                          ctx->stream));
    // 16-bit two-base column offsets where a block's columns fit two windows
    std::vector<uint32_t> base;
    std::vector<uint16_t> off16(nnz);
    A->nnz16 = two_base_blocks(blocks, A->h_rp, A->h_ci.data(), off16, base);
    A->nblocks16 = 0;
    for (uint32_t bi = 0; bi < A->nblocks; ++bi) A->nblocks16 += base[2 * bi] != kNoBase;
    if (A->nblocks16) {
        A->ci16 = dalloc<uint16_t>(nnz);
        A->blk_base = dalloc<uint32_t>(base.size());
        PP_HIP(hipMemcpyAsync(A->ci16, off16.data(), nnz * sizeof(uint16_t), hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipMemcpyAsync(A->blk_base, base.data(), base.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    {
        const std::vector<BlockDesc> ds = block_descs(blocks, A->h_rp, base);
        A->blk_desc = dalloc<BlockDesc>(ds.size());
        PP_HIP(hipMemcpyAsync(A->blk_desc, ds.data(), ds.size() * sizeof(BlockDesc), hipMemcpyHostToDevice,
                              ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));  // ds, blocks, off16 leave scope
    }
    A->tune_warm = tune == 2;
    // the sliced-column copy for matrices timed cold (the reduced CG's
    // operators, timed warm, keep their own)
    // (built on the device from the arrays just uploaded)
    if (tune == 1 && nrow >= (1u << 14) && !ell_build_device(ctx, A.get(), false, 0, A->ell, 1.1) &&
        nrow == ncol) {
        // a colour-ordered matrix (columns about n/2 away from their rows):
        // the two-colour form, the split at the first row with an entry left
        // of its diagonal
        uint32_t split = 0;
        while (split < nrow && (A->h_rp[split] == A->h_rp[split + 1] || A->h_ci[A->h_rp[split]] >= split)) ++split;
        ell_build_device(ctx, A.get(), true, split, A->ell, 1.1);
    }
    if (tune) csr_tune_spmv(ctx, A.get(), PP_F64);
    else A->tuned32 = true;  // neither precision is timed
    return A.release();
}

void csr_tune(pp_ctx *ctx, pp_csr *A) { csr_tune_spmv(ctx, A, PP_F64); }

void csr_ensure_host(pp_csr *A) {
    if (A->h_ci.size() == A->nnz && A->h_v.size() == A->nnz && A->h_rp.size() == (size_t)A->nrow + 1) return;
    PP_HIP(hipStreamSynchronize(A->ctx->stream));
    A->h_rp.resize((size_t)A->nrow + 1);
    A->h_ci.resize(A->nnz);
    A->h_v.resize(A->nnz);
    PP_HIP(hipMemcpy(A->h_rp.data(), A->rp, ((size_t)A->nrow + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (A->nnz) {
        PP_HIP(hipMemcpy(A->h_ci.data(), A->ci, (size_t)A->nnz * sizeof(uint32_t), hipMemcpyDeviceToHost));
        PP_HIP(hipMemcpy(A->h_v.data(), A->v64, (size_t)A->nnz * sizeof(double), hipMemcpyDeviceToHost));
    }
}

pp_csr *csr_from_device(pp_ctx *ctx, uint32_t nrow, uint32_t ncol, uint32_t *rp, uint32_t *ci, double *v,
                        std::vector<uint32_t> &&h_rp) {
    pp_csr *A = new pp_csr;
    A->rp = rp;
    A->ci = ci;
    A->v64 = v;
    try {
        PP_CHECK(h_rp.size() == (size_t)nrow + 1 && h_rp[0] == 0, PP_ERR_INVALID, "rowptr size");
        A->ctx = ctx;
        A->uid = next_uid();
        A->nrow = nrow;
        A->ncol = ncol;
        A->nnz = h_rp[nrow];
        A->h_rp = std::move(h_rp);
        std::vector<uint32_t> blocks;
        build_row_blocks(A->h_rp, blocks);
        A->nblocks = (uint32_t)blocks.size() - 1;
        A->row_blocks = dalloc<uint32_t>(blocks.size());
        PP_HIP(hipMemcpyAsync(A->row_blocks, blocks.data(), blocks.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              ctx->stream));
        // 16-bit two-base column offsets where a block's columns fit two windows
        std::vector<uint32_t> base;
        A->nnz16 = two_base_device(ctx, A->nblocks, A->row_blocks, A->rp, A->ci, A->nnz, A->ci16, A->blk_base, base);
        A->nblocks16 = 0;
        for (uint32_t bi = 0; bi < A->nblocks; ++bi) A->nblocks16 += base[2 * bi] != kNoBase;
        const std::vector<BlockDesc> ds = block_descs(blocks, A->h_rp, base);
        A->blk_desc = dalloc<BlockDesc>(ds.size());
        PP_HIP(hipMemcpyAsync(A->blk_desc, ds.data(), ds.size() * sizeof(BlockDesc), hipMemcpyHostToDevice,
                              ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));  // blocks, ds leave scope
        A->tuned32 = true;                         // neither precision is timed
    } catch (...) {
        csr_free(A);
        throw;
    }
    return A;
}

pp_csr *csr_new(pp_ctx *ctx, uint32_t nrow, uint32_t ncol, const uint32_t *rp, const uint32_t *ci,
                const double *v) {
    PP_CHECK(rp != nullptr, PP_ERR_INVALID, "rowptr is NULL");
    const uint32_t nnz = rp[nrow];
    PP_CHECK(nnz == 0 || (ci && v), PP_ERR_INVALID, "colidx/values are NULL");
    return csr_from_vectors(ctx, nrow, ncol, std::vector<uint32_t>(rp, rp + nrow + 1),
                            std::vector<uint32_t>(ci, ci + nnz), std::vector<double>(v, v + nnz), 1);
}

// Times each SpMV load variant (ordinary / non-temporal operand loads, 32-bit
// / 16-bit column offsets) on this matrix and keeps the fastest.  Only for
// matrices large enough for the choice to matter.
static void csr_tune_spmv(pp_ctx *ctx, pp_csr *A, int dtype) {
    if (A->nnz < (1u << 20)) return;
    int &variant = dtype == PP_F64 ? A->spmv_variant : A->spmv_variant32;
    double *x = dalloc<double>(A->ncol), *y = dalloc<double>(A->nrow);  // large enough for either type
    try {
        launch_fill_zero(ctx, x, (size_t)A->ncol * sizeof(double));
        float best = 1e30f;
        int best_var = variant;
        std::vector<std::pair<uint64_t, int>> cand;  // (bytes streamed, variant)
        for (int var = 0; var < kVarCount; ++var) {
            if ((var & kVarIdx16) && !A->ci16) continue;
            // non-temporal loads evict what the iteration's next launch reads
            if (A->tune_warm && (var & kVarNT)) continue;
            uint64_t bytes = 0;
            variant = var;
            pp_csr_spmv_bytes(A, dtype, &bytes);
            cand.emplace_back(bytes, var);
        }
        // against a sliced-column form only the two CSR variants streaming the
        // fewest bytes are timed (the others move more through HBM, and each
        // cold timing costs a cache flush)
        if (A->ell.ok && cand.size() > 2) {
            std::stable_sort(cand.begin(), cand.end(),
                             [](const std::pair<uint64_t, int> &a, const std::pair<uint64_t, int> &b) {
                                 return a.first < b.first;
                             });
            cand.resize(2);
        }
        for (const auto &c : cand) {
            const int var = c.second;
            variant = var;
            const float ms = time_cold_ms(
                ctx, [&] { launch_spmv(ctx, A, dtype, 1.0, x, 0.0, y, -1, nullptr, false); }, 3, !A->tune_warm);
            if (ms < best) {
                best = ms;
                best_var = var;
            }
        }
        if (A->ell.ok) {
            if (dtype == PP_F32 && !A->ell.v32) {
                A->ell.v32 = dalloc<float>((size_t)A->ell.width * A->ell.n);
                if (A->ell.dg64 && !A->ell.dg32) A->ell.dg32 = dalloc<float>(A->ell.n);
                ell_fill_values(ctx, A->ell, PP_F32, A->rp, A->v32);
            }
            uint64_t csr_bytes = 0, ell_bytes = 0;
            variant = best_var;
            pp_csr_spmv_bytes(A, dtype, &csr_bytes);
            variant = kVarEll;
            pp_csr_spmv_bytes(A, dtype, &ell_bytes);
            const float ms = time_cold_ms(
                ctx, [&] { launch_spmv(ctx, A, dtype, 1.0, x, 0.0, y, -1, nullptr, false); }, 3, !A->tune_warm);
            // within the timing noise (3%) the form that streams fewer bytes,
            // so the choice does not flip between runs
            if (ms < best || (ms < 1.03f * best && ell_bytes < csr_bytes)) {
                best = ms;
                best_var = kVarEll;
            }
        }
        variant = best_var;
    } catch (...) {
        dfree(x);
        dfree(y);
        throw;
    }
    dfree(x);
    dfree(y);
}

pp_csr *csr_create_host(pp_ctx *ctx, HostCsr &&H) {
    return csr_from_vectors(ctx, H.n, H.n, std::move(H.rp), std::move(H.ci), std::move(H.v), 1);
}

void csr_free(pp_csr *A) {
    if (!A) return;
    if (A->trans) csr_free(A->trans);
    dfree(A->rp);
    dfree(A->ci);
    dfree(A->v64);
    dfree(A->v32);
    dfree(A->row_blocks);
    dfree(A->ci16);
    dfree(A->blk_base);
    dfree(A->blk_desc);
    A->ell.free_all();
    delete A;
}

// Explicit transpose (gather form for A^T x, no atomics).
pp_csr *csr_transpose(pp_ctx *ctx, pp_csr *A) {
    if (A->trans) return A->trans;
    HostCsr H;
    H.n = A->nrow;
    H.rp = A->h_rp;
    H.ci = A->h_ci;
    H.v = A->h_v;
    PP_CHECK(A->nrow == A->ncol, PP_ERR_UNSUPPORTED, "transpose of a rectangular matrix");
    A->trans = csr_create_host(ctx, transpose(H));
    return A->trans;
}

}  // namespace pp

using namespace pp;

// ============================================================================ C ABI
extern "C" {

const char *pp_global_error(void) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err.c_str();
}

const char *pp_last_error(const pp_ctx *ctx) { return ctx ? ctx->err.c_str() : pp_global_error(); }

int pp_device_count(int *count) {
    return guarded(nullptr, [&] {
        PP_CHECK(count, PP_ERR_INVALID, "count is NULL");
        int c = 0;
        PP_HIP(hipGetDeviceCount(&c));
        *count = c;
    });
}

int pp_ctx_create(int device, pp_ctx **out) {
    return guarded(nullptr, [&] {
        PP_CHECK(out, PP_ERR_INVALID, "out is NULL");
        int ndev = 0;
        PP_HIP(hipGetDeviceCount(&ndev));
        PP_CHECK(device >= 0 && device < ndev, PP_ERR_INVALID, "no such device");
        std::unique_ptr<pp_ctx> c(new pp_ctx);
        c->device = device;
        PP_HIP(hipSetDevice(device));
        noalloc_setup(device);  // the zero-allocation solve's array address on this device
        PP_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
        c->stream = c->own_stream;
        hipDeviceProp_t prop;
        PP_HIP(hipGetDeviceProperties(&prop, device));
        c->cu_count = prop.multiProcessorCount;
        {
            int lds = 0;
            if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
                c->max_lds = (size_t)lds;
        }
        int clock_khz = 0;
        if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
            clock_khz <= 0)
            clock_khz = 100000;
        c->spin_deadline = (long long)clock_khz * 1000LL * 2;  // a dependency wait may last 2 s
        c->red.partials = dalloc<double>((size_t)S_COUNT * kMaxPartials);
        c->scal = dalloc<double>(S_COUNT);
        c->ctl = dalloc<int>(C_COUNT);
        PP_HIP(hipMemset(c->scal, 0, S_COUNT * sizeof(double)));
        PP_HIP(hipMemset(c->ctl, 0, C_COUNT * sizeof(int)));
        PP_HIP(hipHostMalloc((void **)&c->pinned, 640 * sizeof(double), hipHostMallocDefault));
        PP_HIP(hipHostMalloc((void **)&c->pinned_ctl, 64 * sizeof(int), hipHostMallocDefault));
        PP_HIP(hipHostMalloc((void **)&c->ring_host, pp_ctx::kRing * C_COUNT * sizeof(int),
                             hipHostMallocMapped | hipHostMallocCoherent));
        PP_HIP(hipHostGetDevicePointer((void **)&c->ring_dev, c->ring_host, 0));
        PP_HIP(hipEventCreate(&c->ev0));
        PP_HIP(hipEventCreate(&c->ev1));
        ensure_hist(c.get(), 4096);
        PP_HIP(hipDeviceSynchronize());
        *out = c.release();
    });
}

int pp_ctx_set_stream(pp_ctx *ctx, void *s) {
    if (!ctx) return PP_ERR_INVALID;
    // pooled blocks may still be in use by work queued on the old stream
    if (ctx->pool_dirty && hipStreamSynchronize(ctx->stream) != hipSuccess) return PP_ERR_HIP;
    ctx->pool_dirty = false;
    ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
    return PP_OK;
}

void *pp_ctx_stream(pp_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int pp_ctx_sync(pp_ctx *ctx) {
    return guarded(ctx, [&] { PP_HIP(hipStreamSynchronize(ctx->stream)); });
}

void pp_ctx_destroy(pp_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    pool_release(ctx);
    dfree(ctx->red.partials);
    dfree(ctx->scal);
    dfree(ctx->ctl);
    dfree(ctx->hist);
    dfree(ctx->flush_buf);
    for (int i = 0; i < 48; ++i) dfree(ctx->ws[i]);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->pinned_ctl) (void)hipHostFree(ctx->pinned_ctl);
    if (ctx->ring_host) (void)hipHostFree(ctx->ring_host);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

int pp_malloc(pp_ctx *ctx, size_t bytes, void **dptr) {
    return guarded(ctx, [&] {
        PP_CHECK(dptr, PP_ERR_INVALID, "dptr is NULL");
        *dptr = dalloc<unsigned char>(bytes);
    });
}

int pp_free(pp_ctx *ctx, void *dptr) {
    return guarded(ctx, [&] { PP_HIP(hipFree(dptr)); });
}

int pp_memcpy_h2d(pp_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guarded(ctx, [&] {
        PP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pp_memcpy_d2h(pp_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guarded(ctx, [&] {
        PP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int pp_memcpy_d2d(pp_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guarded(ctx, [&] { PP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream)); });
}

int pp_copy_stream(pp_ctx *ctx, const void *src, void *dst, size_t bytes, uint32_t groups) {
    return guarded(ctx, [&] {
        PP_CHECK(src && dst, PP_ERR_INVALID, "NULL pointer");
        PP_CHECK(((uintptr_t)src | (uintptr_t)dst) % 16 == 0, PP_ERR_INVALID, "pointers must be 16-byte aligned");
        launch_copy_stream(ctx, src, dst, bytes, groups);
    });
}

int pp_memset_zero(pp_ctx *ctx, void *dst, size_t bytes) {
    return guarded(ctx, [&] { launch_fill_zero(ctx, dst, bytes); });
}

int pp_csr_create(pp_ctx *ctx, uint32_t nrow, uint32_t ncol, const uint32_t *rp, const uint32_t *ci,
                  const double *v, pp_csr **out) {
    return guarded(ctx, [&] {
        PP_CHECK(ctx && out, PP_ERR_INVALID, "ctx/out is NULL");
        *out = csr_new(ctx, nrow, ncol, rp, ci, v);
    });
}

int pp_csr_set_values(pp_ctx *ctx, pp_csr *A, const double *values, int on_device) {
    return guarded(ctx, [&] {
        PP_CHECK(A && values, PP_ERR_INVALID, "A/values is NULL");
        if (on_device) {
            PP_HIP(hipMemcpyAsync(A->v64, values, A->nnz * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            PP_HIP(hipMemcpyAsync(A->h_v.data(), values, A->nnz * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
        } else {
            std::memcpy(A->h_v.data(), values, A->nnz * sizeof(double));
            PP_HIP(hipMemcpyAsync(A->v64, values, A->nnz * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        }
        if (A->ell.ok) ell_fill_values(ctx, A->ell, PP_F64, A->rp, A->v64);
        PP_HIP(hipStreamSynchronize(ctx->stream));
        A->f32_dirty = true;
        ++A->version;
        if (A->trans) {
            csr_free(A->trans);
            A->trans = nullptr;
        }
    });
}

int pp_csr_spmv_variant(const pp_csr *A, int dtype, int *variant) {
    if (!A || !variant || (dtype != PP_F64 && dtype != PP_F32)) return PP_ERR_INVALID;
    *variant = dtype == PP_F64 ? A->spmv_variant : A->spmv_variant32;
    return PP_OK;
}

int pp_csr_set_spmv_variant(pp_csr *A, int dtype, int variant) {
    if (!A || variant < 0 || variant > kVarEll || (dtype != PP_F64 && dtype != PP_F32)) return PP_ERR_INVALID;
    if (variant == kVarEll && !A->ell.ok) return PP_ERR_INVALID;
    if (variant != kVarEll && (variant & kVarIdx16) && !A->ci16) return PP_ERR_INVALID;
    if (variant == kVarEll && dtype == PP_F32 && !A->ell.v32 && A->v32 && !A->f32_dirty) {
        A->ell.v32 = dalloc<float>((size_t)A->ell.width * A->ell.n);
        if (A->ell.dg64 && !A->ell.dg32) A->ell.dg32 = dalloc<float>(A->ell.n);
        ell_fill_values(A->ctx, A->ell, PP_F32, A->rp, A->v32);
        PP_HIP(hipStreamSynchronize(A->ctx->stream));
    }
    if (dtype == PP_F64) {
        A->spmv_variant = variant;
    } else {
        A->spmv_variant32 = variant;
        A->tuned32 = true;  // pinned: the FP32 copy's creation keeps it
    }
    return PP_OK;
}

int pp_csr_spmv_bytes(const pp_csr *A, int dtype, uint64_t *bytes) {
    if (!A || !bytes) return PP_ERR_INVALID;
    const uint64_t esz = dtype == PP_F64 ? 8 : 4;
    if ((dtype == PP_F64 ? A->spmv_variant : A->spmv_variant32) == kVarEll) {
        // padded values and 16-bit offsets (or one pattern byte per row and
        // the pattern table), x read, y written
        const uint64_t wn = A->ell.width * (uint64_t)A->ell.n;
        const uint64_t idx = A->ell.pid ? A->ell.n + 16ull * A->ell.npat : 2 * wn;
        const uint64_t diag = A->ell.dg64 ? esz * A->ell.n : 0;  // a separate diagonal vector
        *bytes = esz * wn + idx + diag + esz * ((uint64_t)A->ncol + A->nrow);
        return PP_OK;
    }
    uint64_t b = esz * A->nnz;                                    // values
    const uint64_t n16 = ((dtype == PP_F64 ? A->spmv_variant : A->spmv_variant32) & kVarIdx16) ? A->nnz16 : 0;
    b += 2 * n16 + 4 * ((uint64_t)A->nnz - n16);                 // column indices
    b += 4ull * (A->nrow + 1) + 4ull * (A->nblocks + 1);          // row pointers, row blocks
    if (n16) b += 8ull * A->nblocks;                              // block bases
    b += esz * ((uint64_t)A->ncol + A->nrow);                     // x read, y written
    *bytes = b;
    return PP_OK;
}

int pp_csr_dims(const pp_csr *A, uint32_t *nrow, uint32_t *ncol, uint32_t *nnz) {
    if (!A) return PP_ERR_INVALID;
    if (nrow) *nrow = A->nrow;
    if (ncol) *ncol = A->ncol;
    if (nnz) *nnz = A->nnz;
    return PP_OK;
}

int pp_csr_device_ptrs(const pp_csr *A, const uint32_t **rowptr, const uint32_t **colidx, const double **v64,
                       const float **v32) {
    if (!A) return PP_ERR_INVALID;
    if (rowptr) *rowptr = A->rp;
    if (colidx) *colidx = A->ci;
    if (v64) *v64 = A->v64;
    if (v32) *v32 = A->f32_dirty ? nullptr : A->v32;
    return PP_OK;
}

void pp_csr_destroy(pp_csr *A) {
    if (!A) return;
    (void)hipSetDevice(A->ctx->device);
    (void)hipStreamSynchronize(A->ctx->stream);
    csr_free(A);
}

int pp_spmv(pp_ctx *ctx, const pp_csr *A, int dtype, double alpha, const void *x, double beta, void *y) {
    return guarded(ctx, [&] {
        PP_CHECK(A && x && y, PP_ERR_INVALID, "NULL argument");
        PP_CHECK(dtype == PP_F64 || dtype == PP_F32, PP_ERR_INVALID, "bad dtype");
        if (dtype == PP_F32) csr_need_f32(ctx, const_cast<pp_csr *>(A));
        launch_spmv(ctx, A, dtype, alpha, x, beta, y, -1, nullptr, false);
    });
}

int pp_spmv_t(pp_ctx *ctx, pp_csr *A, int dtype, double alpha, const void *x, double beta, void *y) {
    return guarded(ctx, [&] {
        PP_CHECK(A && x && y, PP_ERR_INVALID, "NULL argument");
        pp_csr *T = csr_transpose(ctx, A);
        if (dtype == PP_F32) csr_need_f32(ctx, T);
        launch_spmv(ctx, T, dtype, alpha, x, beta, y, -1, nullptr, false);
    });
}

int pp_dot(pp_ctx *ctx, int dtype, uint32_t n, const void *x, const void *y, double *result) {
    return guarded(ctx, [&] {
        PP_CHECK(x && y && result, PP_ERR_INVALID, "NULL argument");
        launch_dot(ctx, dtype, n, x, y, S_TMP1);
        *result = read_slot(ctx, S_TMP1);
    });
}

int pp_nrm2(pp_ctx *ctx, int dtype, uint32_t n, const void *x, double *result) {
    return guarded(ctx, [&] {
        PP_CHECK(x && result, PP_ERR_INVALID, "NULL argument");
        launch_dot(ctx, dtype, n, x, x, S_TMP1);
        *result = std::sqrt(read_slot(ctx, S_TMP1));
    });
}

int pp_axpby(pp_ctx *ctx, int dtype, uint32_t n, double a, const void *x, double b, const void *y, void *z) {
    return guarded(ctx, [&] {
        PP_CHECK(x && y && z, PP_ERR_INVALID, "NULL argument");
        launch_axpby(ctx, dtype, n, a, x, b, y, z);
    });
}

int pp_scale(pp_ctx *ctx, int dtype, uint32_t n, double a, void *x) {
    return guarded(ctx, [&] {
        PP_CHECK(x, PP_ERR_INVALID, "NULL argument");
        launch_axpby(ctx, dtype, n, a, x, 0.0, x, x);
    });
}

int pp_convert(pp_ctx *ctx, uint32_t n, int sdt, const void *s, int ddt, void *d) {
    return guarded(ctx, [&] {
        PP_CHECK(s && d, PP_ERR_INVALID, "NULL argument");
        launch_convert(ctx, n, sdt, s, ddt, d);
    });
}

int pp_csr_outer_on_pattern(pp_ctx *ctx, const pp_csr *A, double scale, const double *lambda, const double *x,
                            double *grad) {
    return guarded(ctx, [&] {
        PP_CHECK(A && lambda && x && grad, PP_ERR_INVALID, "NULL argument");
        launch_outer_on_pattern(ctx, A, scale, lambda, x, grad);
    });
}

// ---------------------------------------------------------------------------- inputs
// Seeded SplitMix64 stream (Steele, Lea, Flood 2014).
static inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

int pp_fill_random(int kind, uint64_t seed, uint64_t n, double *out) {
    return guarded(nullptr, [&] {
        PP_CHECK(out || n == 0, PP_ERR_INVALID, "out is NULL");
        uint64_t s = seed;
        const double inv53 = 1.0 / 9007199254740992.0;
        if (kind == 0) {
            for (uint64_t i = 0; i < n; ++i) out[i] = 2.0 * ((double)(splitmix64(s) >> 11) * inv53) - 1.0;
        } else {
            const double two_pi = 6.283185307179586476925286766559;
            for (uint64_t i = 0; i < n; i += 2) {
                double u1 = (double)((splitmix64(s) >> 11) + 1) * inv53;
                double u2 = (double)(splitmix64(s) >> 11) * inv53;
                double r = std::sqrt(-2.0 * std::log(u1));
                double th = two_pi * u2;
                out[i] = r * std::cos(th);
                if (i + 1 < n) out[i + 1] = r * std::sin(th);
            }
        }
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------- generators
namespace {
struct StencilOut {
    uint32_t *n, *nnz, *rp, *ci;
    double *v;
};

// 7-point (or 5-point when nz == 1) stencil on a lexicographic x-fastest grid;
// coef(dir) gives the off-diagonal for dir in {z-, y-, x-, x+, y+, z+}, written
// in ascending column order.
template <class C>
int stencil(int nx, int ny, int nz, double diag, C coef, StencilOut o) {
    return guarded(nullptr, [&] {
        PP_CHECK(nx > 0 && ny > 0 && nz > 0, PP_ERR_INVALID, "grid dims must be positive");
        const int64_t n = (int64_t)nx * ny * nz;
        PP_CHECK(n < (int64_t)UINT32_MAX / 8, PP_ERR_INVALID, "grid too large for 32-bit indices");
        const int64_t nxy = (int64_t)nx * ny;
        uint64_t nnz = 0;
        for (int64_t p = 0; p < n; ++p) {
            int64_t k = p / nxy, i = (p - k * nxy) / nx, j = p - k * nxy - i * nx;
            nnz += 1 + (k > 0) + (i > 0) + (j > 0) + (j < nx - 1) + (i < ny - 1) + (k < nz - 1);
        }
        if (o.n) *o.n = (uint32_t)n;
        if (o.nnz) *o.nnz = (uint32_t)nnz;
        if (!o.rp) return;
        PP_CHECK(o.ci && o.v, PP_ERR_INVALID, "colidx/values NULL");
        uint32_t q = 0;
        o.rp[0] = 0;
        for (int64_t p = 0; p < n; ++p) {
            int64_t k = p / nxy, i = (p - k * nxy) / nx, j = p - k * nxy - i * nx;
            auto put = [&](int64_t c, double v) {
                o.ci[q] = (uint32_t)c;
                o.v[q++] = v;
            };
            if (k > 0) put(p - nxy, coef(0));
            if (i > 0) put(p - nx, coef(1));
            if (j > 0) put(p - 1, coef(2));
            put(p, diag);
            if (j < nx - 1) put(p + 1, coef(3));
            if (i < ny - 1) put(p + nx, coef(4));
            if (k < nz - 1) put(p + nxy, coef(5));
            o.rp[p + 1] = q;
        }
    });
}
}  // namespace

extern "C" {

// generateLaplacianCOO + transCOO2CSR (ModelProblem.cpp:68-133, MatrixTools.h:16-62):
// 2D when a dimension is 1 (diagonal 4), else 3D (diagonal 6); centred
// convection: "+" neighbour -1 - alpha, "-" neighbour -1 + alpha; diagonal -= shift.
int pp_generate_laplacian(int nx, int ny, int nz, double ax, double ay, double az, double shift, uint32_t *n,
                          uint32_t *nnz, uint32_t *rp, uint32_t *ci, double *v) {
    int is2d = (nx == 1) + (ny == 1) + (nz == 1) > 0;
    double vd = (is2d ? 4.0 : 6.0) - shift;
    double c[6] = {-1.0 + az, -1.0 + ay, -1.0 + ax, -1.0 - ax, -1.0 - ay, -1.0 - az};
    return stencil(nx, ny, nz, vd, [&](int d) { return c[d]; }, StencilOut{n, nnz, rp, ci, v});
}

// First-order upwind convection-diffusion scaled by h^2, h = 1/(nx+1):
// diagonal 2*dim + h*sum(beta), upstream neighbour -1 - h*beta_k, downstream -1.
int pp_generate_upwind(int nx, int ny, int nz, double bx, double by, double bz, uint32_t *n, uint32_t *nnz,
                       uint32_t *rp, uint32_t *ci, double *v) {
    double h = 1.0 / (double)(nx + 1);
    double vd = nz > 1 ? 6.0 + h * (bx + by + bz) : 4.0 + h * (bx + by);
    double c[6] = {-1.0 - h * bz, -1.0 - h * by, -1.0 - h * bx, -1.0, -1.0, -1.0};
    return stencil(nx, ny, nz, vd, [&](int d) { return c[d]; }, StencilOut{n, nnz, rp, ci, v});
}

// -div(kappa grad u) on a side x side grid, Dirichlet; interior faces take the
// harmonic mean of the two nodal coefficients, boundary faces the node's own.
int pp_generate_vardiff2d(int side, const double *kappa, uint32_t *n, uint32_t *nnz, uint32_t *rp, uint32_t *ci,
                          double *v) {
    return guarded(nullptr, [&] {
        PP_CHECK(side > 0, PP_ERR_INVALID, "side must be positive");
        const int64_t N = (int64_t)side * side;
        uint64_t cnt = (uint64_t)N * 5 - 4 * (uint64_t)side;
        if (n) *n = (uint32_t)N;
        if (nnz) *nnz = (uint32_t)cnt;
        if (!rp) return;
        PP_CHECK(kappa && ci && v, PP_ERR_INVALID, "NULL argument");
        uint32_t q = 0;
        rp[0] = 0;
        for (int64_t p = 0; p < N; ++p) {
            int64_t j = p / side, i = p - j * side;
            double kp = kappa[p];
            auto hm = [&](int64_t o) { return 2.0 * kp * kappa[o] / (kp + kappa[o]); };
            double fs = j > 0 ? hm(p - side) : kp;
            double fw = i > 0 ? hm(p - 1) : kp;
            double fe = i < side - 1 ? hm(p + 1) : kp;
            double fn = j < side - 1 ? hm(p + side) : kp;
            if (j > 0) { ci[q] = (uint32_t)(p - side); v[q++] = -fs; }
            if (i > 0) { ci[q] = (uint32_t)(p - 1); v[q++] = -fw; }
            ci[q] = (uint32_t)p; v[q++] = ((fs + fw) + fe) + fn;
            if (i < side - 1) { ci[q] = (uint32_t)(p + 1); v[q++] = -fe; }
            if (j < side - 1) { ci[q] = (uint32_t)(p + side); v[q++] = -fn; }
            rp[p + 1] = q;
        }
    });
}

}  // extern "C"
