// pp_solve.cpp -- Krylov drivers of the C ABI: device-resident PCG (FP64, and
// mixed precision = FP32 inner PCG inside FP64 iterative refinement) and
// FGMRES(m) with CGS2 orthogonalisation.
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pp_internal.h"

namespace pp {
void csr_need_f32(pp_ctx *ctx, pp_csr *A);
pp_csr *csr_transpose(pp_ctx *ctx, pp_csr *A);
void precond_apply(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y);
void precond_need_f32(pp_precond *M);
void launch_scale_convert(pp_ctx *ctx, uint32_t n, const double *src, double scale, float *dst);
void launch_axpy_f32_to_f64(pp_ctx *ctx, uint32_t n, double a, const float *x, double *y);

namespace {

// Workspace indices in the context cache.
enum { W_R = 0, W_W, W_D, W_V, W_R32, W_W32, W_D32, W_V32, W_X32, W_B32, W_E32, W_VB, W_ZB, W_ZB32, W_T, W_T32,
       W_PB, W_PX, W_SR64, W_SR32 = W_SR64 + 5,  // W_SR*: five vectors each (up to slot 27)
       W_RD64 = kWsFree, W_RD32 = W_RD64 + 5,      // W_RD*: five vectors each (reduced CG)
       W_RIN64 = W_RD32 + 5, W_E64 };              // up to slot 43
static_assert(W_SR32 + 5 <= kWsGmresState && W_E64 < 48, "workspace slots overlap");

template <class T>
constexpr int dt() {
    return sizeof(T) == 8 ? PP_F64 : PP_F32;
}

// The load variants a body's kernels were captured with: the SpMV variants of
// A (both precisions) and the sweep variants of M's factor.  Part of every
// graph key, so that pinning a variant after a solve takes effect in the next
// solve instead of replaying a graph of the old launches.
uint64_t variant_key(const pp_csr *A, const pp_precond *M) {
    uint64_t k = A ? ((uint64_t)(A->spmv_variant & 0xff) | (uint64_t)(A->spmv_variant32 & 0xff) << 8) : 0;
    if (M)
        k |= (uint64_t)(M->low64.variant & 0xf) << 16 | (uint64_t)(M->up64.variant & 0xf) << 20 |
             (uint64_t)(M->low32.variant & 0xf) << 24 | (uint64_t)(M->up32.variant & 0xf) << 28 |
             (uint64_t)(M->low64.ell_fallback & 0xf) << 32 | (uint64_t)(M->up64.ell_fallback & 0xf) << 36 |
             (uint64_t)(M->low32.ell_fallback & 0xf) << 40 | (uint64_t)(M->up32.ell_fallback & 0xf) << 44 |
             (uint64_t)(M->trsv & 0xf) << 48;
    return k;
}

struct PcgResult {
    int32_t iters = 0;
    int conv = 0;
    int status = PP_OK;
    double final_error = -1;
    int nhist = 0;
};

// Runs iteration bodies until the stop word: the body (every launch of it:
// sweeps, updates, SpMV) is captured once into a hipGraph and replayed, or
// (use_graph false) launched on the stream each time.  kInFlight bodies stay in
// flight: the last convergence check of each body writes the control words
// into the next slot of a mapped pinned ring, sequence number last, and the
// host polls that number (no event or blocking wait between bodies, so
// consecutive graph launches run back to back), so the GPU never waits for
// the host; a body past convergence only runs kernels that return at once.
// check_first: the control words before the first body decide whether any
// body runs (they are read synchronously).
// key: everything the body's kernels are launched with; an instantiated
// graph with the same key is reused (capture and instantiation cost more
// than a few iterations).
// On return ctl holds the final control words and *norm the last check's
// relative norm (scal[S_TMP0]): where a body published the stop with its
// norm (C_NORM_SET), both come from the ring and the host waits once, for
// the bodies still in flight to drain; otherwise they are read after it.
// The cached graph of `body` under `key` (captured on first use, as in
// drive_bodies), launched once on the context's stream.
template <class Body>
void run_graph(pp_ctx *ctx, const std::vector<uint64_t> &key, Body &body) {
    Graph *G = nullptr;
    for (auto &e : ctx->graphs)
        if (e.first == key) G = e.second.get();
    if (!G) {
        std::unique_ptr<Graph> fresh(new Graph);
        PP_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t tmp = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &tmp);
            if (tmp) (void)hipGraphDestroy(tmp);
            throw;
        }
        PP_HIP(hipStreamEndCapture(ctx->stream, &fresh->g));
        PP_HIP(hipGraphInstantiate(&fresh->exec, fresh->g, nullptr, nullptr, 0));
        G = fresh.get();
        if (ctx->graphs.size() >= 8) ctx->graphs.erase(ctx->graphs.begin());
        ctx->graphs.emplace_back(key, std::move(fresh));
    }
    PP_HIP(hipGraphLaunch(G->exec, ctx->stream));
}

template <class Body>
void drive_bodies(pp_ctx *ctx, Body &body, bool use_graph, uint64_t max_replays, int *ctl, bool check_first,
                  const std::vector<uint64_t> &key, double *norm) {
    Graph *G = nullptr;
    std::unique_ptr<Graph> fresh;
    if (use_graph) {
        for (auto &e : ctx->graphs)
            if (e.first == key) G = e.second.get();
        if (!G) {
            fresh.reset(new Graph);
            PP_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                hipGraph_t tmp = nullptr;
                (void)hipStreamEndCapture(ctx->stream, &tmp);
                if (tmp) (void)hipGraphDestroy(tmp);
                throw;
            }
            PP_HIP(hipStreamEndCapture(ctx->stream, &fresh->g));
            PP_HIP(hipGraphInstantiate(&fresh->exec, fresh->g, nullptr, nullptr, 0));
            G = fresh.get();
            if (ctx->graphs.size() >= 8) ctx->graphs.erase(ctx->graphs.begin());
            ctx->graphs.emplace_back(key, std::move(fresh));
        }
    }
    constexpr int R = pp_ctx::kRing;
    volatile int *slots = ctx->ring_host;
    for (int i = 0; i < R * C_COUNT; ++i) slots[i] = 0;  // no sequence number of an earlier solve remains
    uint64_t issued = 0, checked = 0;
    auto issue = [&] {
        if (use_graph) PP_HIP(hipGraphLaunch(G->exec, ctx->stream));
        else body();
        ++issued;
    };
    // body k has published when its slot carries sequence number k + 1; the
    // stream is queried now and then so that a failed launch cannot leave
    // the host waiting (an idle stream with no publication ends the wait)
    auto wait_published = [&](uint64_t k) -> bool {
        volatile int *slot = slots + (k % R) * C_COUNT;
        for (uint64_t spin = 1;; ++spin) {
            if ((uint64_t)slot[C_PUB] == k + 1) return true;
            if ((spin & 1023) == 0) {
                const hipError_t e = hipStreamQuery(ctx->stream);
                if (e == hipSuccess) return (uint64_t)slot[C_PUB] == k + 1;
                if (e != hipErrorNotReady) PP_HIP(e);
            }
        }
    };
    bool run = true, published = false;
    if (check_first) {
        read_ctl(ctx, ctl);
        PP_CHECK(!ctl[C_SPIN], PP_ERR_HIP, "sync-free triangular sweep timed out waiting for a dependency");
        run = !ctl[C_STOP];
    }
    if (run) {
        for (;;) {
            while (issued < checked + pp_ctx::kInFlight && issued < max_replays) issue();
            if (!wait_published(checked)) break;  // read the words synchronously below
            for (int i = 0; i < C_COUNT; ++i) ctl[i] = slots[(checked % R) * C_COUNT + i];
            ++checked;
            PP_CHECK(!ctl[C_SPIN], PP_ERR_HIP, "sync-free triangular sweep timed out waiting for a dependency");
            if (ctl[C_STOP]) published = true;  // bodies after the stop leave the words as they are
            if (ctl[C_STOP] || checked >= max_replays) break;
        }
    }
    if (published && ctl[C_NORM_SET]) {
        PP_HIP(hipStreamSynchronize(ctx->stream));  // no later body may publish into the next solve's ring
        const uint64_t bits = (uint64_t)(uint32_t)ctl[C_NORM_HI] << 32 | (uint32_t)ctl[C_NORM_LO];
        std::memcpy(norm, &bits, sizeof(double));
        return;
    }
    PP_HIP(hipMemcpyAsync(ctx->pinned_ctl, ctx->ctl, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipMemcpyAsync(ctx->pinned, ctx->scal + S_TMP0, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PP_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(ctl, ctx->pinned_ctl, C_COUNT * sizeof(int));
    *norm = ctx->pinned[0];
}

// Result of a device PCG from its final control words and last norm.
PcgResult collect_pcg(pp_ctx *ctx, const int *ctl, double norm, uint32_t maxit, bool hist_on) {
    PcgResult res;
    res.conv = ctl[C_CONV];
    res.status = ctl[C_STATUS];
    const int m = ctl[C_M];
    res.iters = m + 1;  // CG.cpp:81
    res.final_error = norm;
    if (ctl[C_CHECK] <= 1 && !res.conv && maxit <= 1) res.final_error = -1;
    res.nhist = std::max(0, std::min<int>(m, (int)maxit));
    if (hist_on) {
        int cnt = res.conv ? m : std::max(0, ctl[C_CHECK] - 1);
        if (!res.conv && ctl[C_LAST]) cnt = (int)maxit - 1;
        cnt = std::max(0, std::min<int>(cnt, (int)ctx->hist_cap));
        ctx->last_hist.assign(cnt, 0.0);
        if (cnt) PP_HIP(hipMemcpy(ctx->last_hist.data(), ctx->hist, cnt * sizeof(double), hipMemcpyDeviceToHost));
        res.nhist = cnt;
    }
    return res;
}

// FlexibleCG on the device (source/Solver/CG.cpp:33-83), working precision T.
// A, M already hold T copies.  x holds x0.  One v = A d per iteration: the
// reference's repeated product at CG.cpp:68 returns the same vector.
template <class T>
PcgResult pcg_device(pp_ctx *ctx, pp_csr *A, pp_precond *M, bool flexible, const T *b, T *x, double tol,
                     uint32_t maxit, uint32_t check_every, bool hist_on, int wbase, bool use_graph) {
    const uint32_t n = A->nrow;
    const int DT = dt<T>();
    T *r = (T *)ws_get(ctx, wbase + 0, (size_t)n * sizeof(T));
    T *w = (T *)ws_get(ctx, wbase + 1, (size_t)n * sizeof(T));
    T *d = (T *)ws_get(ctx, wbase + 2, (size_t)n * sizeof(T));
    T *v = (T *)ws_get(ctx, wbase + 3, (size_t)n * sizeof(T));
    if (hist_on) ensure_hist(ctx, maxit + 2);
    ctl_init(ctx, 1, maxit <= 1 ? 1 : 0, maxit);
    set_slot(ctx, S_TOL, tol);
    // r = b - A x
    PP_HIP(hipMemcpyAsync(r, b, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
    launch_spmv(ctx, A, DT, -1.0, x, 1.0, r, -1, nullptr, false);
    launch_dot(ctx, DT, n, b, b, S_BB);
    // w = M^{-1} r ; d = w ; v = A d ; x1, r1
    if (M) precond_apply(ctx, M, DT, r, w);
    else PP_HIP(hipMemcpyAsync(w, r, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
    pcg_d_update<T>(ctx, n, true, w, d, r);
    launch_spmv(ctx, A, DT, 1.0, d, 0.0, v, S_DV, d, false);
    pcg_xr_update<T>(ctx, n, x, r, d, v, hist_on, false);

    int ctl[C_COUNT];
    const uint32_t every = std::max<uint32_t>(1, check_every);
    auto body = [&] {
        StopScope scope(ctx);  // a body past convergence returns at once, SpMV included
        for (uint32_t k = 0; k < every; ++k) {
            if (flexible && M) precond_apply(ctx, M, DT, r, w);
            else PP_HIP(hipMemcpyAsync(w, r, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
            pcg_wv<T>(ctx, n, w, v);
            pcg_d_update<T>(ctx, n, false, w, d, r);
            launch_spmv(ctx, A, DT, 1.0, d, 0.0, v, S_DV, d, false);
            // the body's last check publishes the control words to the host
            pcg_xr_update<T>(ctx, n, x, r, d, v, hist_on, k + 1 == every);
        }
    };
    const std::vector<uint64_t> key = {1,          sizeof(T),  A->uid,     (uint64_t)A->v32,
                                       M ? M->uid : 0, M ? (uint64_t)M->io_permuted : 0, M ? (uint64_t)M->have32 : 0,
                                       (uint64_t)flexible, every, (uint64_t)hist_on, (uint64_t)ctx->hist,
                                       ctx->hist_cap, (uint64_t)r, (uint64_t)w, (uint64_t)d, (uint64_t)v,
                                       (uint64_t)x, (uint64_t)b, variant_key(A, M)};
    double norm;
    drive_bodies(ctx, body, use_graph, (uint64_t)maxit / every + 2, ctl, true, key, &norm);
    return collect_pcg(ctx, ctl, norm, maxit, hist_on);
}

// Single-reduction PCG (k_cgsr_finish / k_cgsr_update): per iteration the
// preconditioner sweeps, one SpMV carrying the three inner products, a
// one-workgroup finish (test, alpha, beta) and one fused vector update.
// Used as the FP32 inner solver of the mixed-precision refinement, and on
// request in FP64.  x holds x0.
template <class T>
PcgResult pcg_sr_device(pp_ctx *ctx, pp_csr *A, pp_precond *M, const T *b, T *x, double tol, uint32_t maxit,
                        uint32_t check_every, bool hist_on, bool use_graph) {
    const uint32_t n = A->nrow;
    const int DT = dt<T>();
    const int base = sizeof(T) == 8 ? W_SR64 : W_SR32;
    T *r = (T *)ws_get(ctx, base + 0, (size_t)n * sizeof(T));
    T *u = (T *)ws_get(ctx, base + 1, (size_t)n * sizeof(T));
    T *w = (T *)ws_get(ctx, base + 2, (size_t)n * sizeof(T));
    T *p = (T *)ws_get(ctx, base + 3, (size_t)n * sizeof(T));
    T *s = (T *)ws_get(ctx, base + 4, (size_t)n * sizeof(T));
    if (hist_on) ensure_hist(ctx, maxit + 2);
    ctl_init(ctx, 1, maxit <= 1 ? 1 : 0, maxit);
    set_slot(ctx, S_TOL, tol);
    // r = b - A x, p = s = 0, first-iteration flag
    PP_HIP(hipMemcpyAsync(r, b, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
    launch_spmv(ctx, A, DT, -1.0, x, 1.0, r, -1, nullptr, false);
    launch_dot(ctx, DT, n, b, b, S_BB);
    launch_fill_zero(ctx, p, (size_t)n * sizeof(T));
    launch_fill_zero(ctx, s, (size_t)n * sizeof(T));
    set_slot(ctx, S_CG_FIRST, 1.0);
    const unsigned np = spmv_grid(A);
    const uint32_t every = std::max<uint32_t>(1, check_every);
    // the update kernel also produces the next apply's dependency-free first
    // forward level (one launch less per iteration)
    const bool l0 = precond_level0_fusable(M);
    uint32_t l0_lo = 0, l0_hi = 0;
    const T *diag0 = nullptr;
    if (l0) {
        l0_lo = M->lvl0_lo;
        l0_hi = M->lvl0_hi;
        diag0 = sizeof(T) == 8 ? (const T *)M->low64.diag : (const T *)M->low32.diag;
        precond_level0(ctx, M, DT, r, u);
    }
    auto body = [&] {
        StopScope scope(ctx);
        for (uint32_t k = 0; k < every; ++k) {
            if (l0) precond_apply_after_level0(ctx, M, DT, r, u);
            else if (M) precond_apply(ctx, M, DT, r, u);
            else launch_convert(ctx, n, DT, r, DT, u);
            launch_spmv(ctx, A, DT, 1.0, u, 0.0, w, S_CG_DELTA, u, false, r);
            cgsr_finish(ctx, np, hist_on, k + 1 == every);
            if (l0) cgsr_update<T>(ctx, n, u, w, p, s, x, r, l0_lo, l0_hi, M->hlow.mode, diag0);
            else cgsr_update<T>(ctx, n, u, w, p, s, x, r);
        }
    };
    int ctl[C_COUNT];
    const std::vector<uint64_t> key = {2,          sizeof(T),  A->uid,     (uint64_t)A->v32,
                                       M ? M->uid : 0, M ? (uint64_t)M->io_permuted : 0, M ? (uint64_t)M->have32 : 0,
                                       (uint64_t)l0, every, (uint64_t)hist_on, (uint64_t)ctx->hist, ctx->hist_cap,
                                       (uint64_t)r, (uint64_t)u, (uint64_t)w, (uint64_t)p, (uint64_t)s,
                                       (uint64_t)x, (uint64_t)b, variant_key(A, M)};
    double norm;
    drive_bodies(ctx, body, use_graph, (uint64_t)maxit / every + 2, ctl, false, key, &norm);
    return collect_pcg(ctx, ctl, norm, maxit, hist_on);
}

// CG on the reduced system of a two-colour IC(0) (ReducedForm): solves
// A e = rin approximately, to the relative reduction tol of the reduced
// residual.  Per iteration y0 = KA z, w = KB z carrying the three inner
// products of the single-reduction recurrence (the preconditioned residual
// is the residual: the IC(0) preconditioner is folded into the operator), a
// one-workgroup finish and one vector update, all on the second colour's n1
// unknowns.  Before: the reduced right-hand side r = KF rin; after:
// e = [KE [rin0; e1]; e1], e1 = s1 * x.  Vectors of the second colour live
// in z[n0..n), where KA reads them.
//
// resume: continue the previous call's iteration instead of starting anew
// (a refinement step that keeps the search directions: the caller replaced
// the residual by the true one, rescaled by `scale` relative to the previous
// right-hand side, so p and s are rescaled by it and gamma_old by its square;
// alpha is scale-free).  Without it every refinement step would restart the
// Krylov iteration and lose its convergence.
//
// rin_norm2: ||rin||^2 when the caller knows it (a normalised residual), else
// negative and it is computed.
template <class T>
PcgResult pcg_reduced_device(pp_ctx *ctx, ReducedForm *R, const T *rin, T *e, double tol, uint32_t maxit,
                             uint32_t check_every, bool use_graph, bool resume = false, double scale = 1.0,
                             double rin_norm2 = -1.0) {
    const uint32_t n0 = R->n0, n1 = R->n1, n = n0 + n1;
    const int DT = dt<T>();
    const int base = sizeof(T) == 8 ? W_RD64 : W_RD32;
    T *z = (T *)ws_get(ctx, base + 0, (size_t)n * sizeof(T));
    T *w = (T *)ws_get(ctx, base + 1, (size_t)n1 * sizeof(T));
    T *p = (T *)ws_get(ctx, base + 2, (size_t)n1 * sizeof(T));
    T *s = (T *)ws_get(ctx, base + 3, (size_t)n1 * sizeof(T));
    T *x = (T *)ws_get(ctx, base + 4, (size_t)n1 * sizeof(T));
    T *r = z + n0;
    const T *s1 = sizeof(T) == 8 ? (const T *)R->s1_64 : (const T *)R->s1_32;
    const T *wt = sizeof(T) == 8 ? (const T *)R->dt_64 : (const T *)R->dt_32;
    launch_spmv(ctx, R->KF, DT, 1.0, rin, 0.0, r, -1, nullptr, false);
    // the test is on the inner system's own residual: ||rin - A e|| is
    // ||Dt^1/2 r|| (the first colour's part vanishes), relative to ||rin||
    if (rin_norm2 < 0.0) launch_dot(ctx, DT, n, rin, rin, S_BB);
    red_begin(ctx, maxit, tol, rin_norm2, resume, scale, DT, n1, x, p, s);
    // KA and KB in sliced-column form where they were built so, else CSR
    const bool ell = R->ka.ok && R->kb.ok;
    const unsigned np = ell ? ell_grid(R->kb, true) : spmv_grid(R->KB);
    // two launches per iteration where KA and KB allow it: the first
    // iteration's products here, then each body iteration's k_red_ka checks
    // the previous one and forms the next KA product with the update folded
    // in, k_red_kb applies the update and forms KB with its inner products
    const bool fused = ell && red_fused_ok(R->ka, R->kb, DT);
    if (fused) {
        launch_ell(ctx, R->ka, DT, 1.0, z, 0.0, z, -1, nullptr, nullptr, nullptr);
        launch_ell(ctx, R->kb, DT, 1.0, z, 0.0, w, S_CG_DELTA, r, r, wt);
    }
    // the check alternates its state blocks: an even number of iterations
    // per body (the device still tests every one)
    const uint32_t every = std::max<uint32_t>(2, check_every + (check_every & 1));
    auto body = [&] {
        StopScope scope(ctx);
        for (uint32_t k = 0; k < every; ++k) {
            if (fused) {
                red_ka<T>(ctx, R->ka, z, z, n0, w, s, np, (int)(k & 1), k + 1 == every);
                red_kb<T>(ctx, R->kb, z, n0, w, p, s, x, wt, np);
                continue;
            }
            if (ell) {
                launch_ell(ctx, R->ka, DT, 1.0, z, 0.0, z, -1, nullptr, nullptr, nullptr);  // z[0, n0) = B r
                launch_ell(ctx, R->kb, DT, 1.0, z, 0.0, w, S_CG_DELTA, r, r, wt);  // w = Sb r, (r,w), (r,r)
            } else {
                launch_spmv(ctx, R->KA, DT, 1.0, z, 0.0, z, -1, nullptr, false);
                launch_spmv(ctx, R->KB, DT, 1.0, z, 0.0, w, S_CG_DELTA, r, false, r, wt);
            }
            red_step<T>(ctx, np, (int)(k & 1), k + 1 == every, n1, w, p, s, x, r);
        }
    };
    int ctl[C_COUNT];
    const std::vector<uint64_t> key = {3,
                                       sizeof(T),
                                       R->KA->uid,
                                       R->KB->uid,
                                       (uint64_t)(sizeof(T) == 8 ? (const void *)R->KA->v64 : (const void *)R->KA->v32),
                                       (uint64_t)(sizeof(T) == 8 ? (const void *)R->KB->v64 : (const void *)R->KB->v32),
                                       every,
                                       (uint64_t)z,
                                       (uint64_t)w,
                                       (uint64_t)p,
                                       (uint64_t)s,
                                       (uint64_t)x,
                                       variant_key(R->KA, nullptr) << 16 | variant_key(R->KB, nullptr),
                                       (uint64_t)fused};
    double norm;
    drive_bodies(ctx, body, use_graph, (uint64_t)maxit / every + 2, ctl, false, key, &norm);
    PcgResult res = collect_pcg(ctx, ctl, norm, maxit, false);
    red_post<T>(ctx, n0, n1, x, s1, rin, z, e);
    launch_spmv(ctx, R->KE, DT, 1.0, z, 0.0, e, -1, nullptr, false);
    return res;
}

double true_relres(pp_ctx *ctx, pp_csr *A, const double *b, const double *x) {
    const uint32_t n = A->nrow;
    double *t = (double *)ws_get(ctx, W_T, (size_t)n * sizeof(double));
    PP_HIP(hipMemcpyAsync(t, b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    launch_spmv(ctx, A, PP_F64, -1.0, x, 1.0, t, -1, nullptr, false);
    launch_dot(ctx, PP_F64, n, t, t, S_TMP2);
    launch_dot(ctx, PP_F64, n, b, b, S_TMP3);
    double out[2];
    read_slots(ctx, S_TMP2, 2, out);
    return std::sqrt(out[0]) / std::sqrt(out[1]);
}

// Mixed precision: FP64 residual r = b - A x and update x += ||r|| e on the
// device, inner FP32 FlexibleCG on A e = r / ||r|| (FP32 factor, SpMV, sweeps
// and accumulation) to a relative reduction max(inner_tol, 0.5 tol ||b||/||r||).
// Iterative refinement: FP64 residual r = b - A x and update x += ||r|| e,
// with e an inner solve of A e = r / ||r|| in precision Tin -- FP32 for the
// mixed-precision solve (tolerance max(inner_tol, what the outer tolerance
// still needs)), FP64 for the reduced CG run as a plain FP64 solve (the
// outer loop then only confirms the true residual).  Inner solver: the
// reduced two-colour CG when R is given, else the single-reduction PCG or
// (PP_CG_REFERENCE) FlexibleCG.
template <class Tin>
PcgResult pcg_refine(pp_ctx *ctx, pp_csr *A, pp_precond *M, ReducedForm *R, const pp_solve_opts &o, const double *b,
                     double *x, int32_t &outer, int32_t &inner_total, int32_t &steps64, double &relres) {
    constexpr bool f32 = sizeof(Tin) == 4;
    const uint32_t n = A->nrow;
    if (f32) {
        if (R) reduced_need_f32(ctx, R);
        else {
            csr_need_f32(ctx, A);
            if (M) precond_need_f32(M);
        }
    }
    double *r = (double *)ws_get(ctx, W_T, (size_t)n * sizeof(double));
    Tin *rin = (Tin *)ws_get(ctx, f32 ? W_B32 : W_RIN64, (size_t)n * sizeof(Tin));
    Tin *e = (Tin *)ws_get(ctx, f32 ? W_E32 : W_E64, (size_t)n * sizeof(Tin));
    PcgResult res;
    launch_dot(ctx, PP_F64, n, b, b, S_TMP3);
    double bnorm = -1.0, prev_rnorm = 0.0;  // ||b|| is read with the first residual norm
    // Reduced FP32 steps continue one Krylov iteration across refinement
    // steps (resume), which needs the replaced residual to agree with the
    // FP32 recurrence's in direction.  FP32 rounding separates the two by
    // about eps32 ||A|| ||e|| per unit residual, e the step's correction --
    // large when the residual is rich in the slowly converging modes.  Each
    // step asks a reduction `step` at least kMargin times that gap (from
    // kResumeStep up to kMaxStep), and the next step resumes only where the
    // finished step met that margin; where the margin would need steps
    // larger than kMaxStep, or a resumed step fell well short of its own
    // test on the true residual, the remaining steps run in FP64.
    constexpr double kResumeStep = 1e-4, kMaxStep = 1e-2, kMargin = 100.0, kEps32 = 5.96e-8;
    bool resume = false;
    bool inner64 = !f32;
    double step = kResumeStep;
    double step_claim = 0.0;  // the reduction the last inner solve was asked for
    outer = 0;
    inner_total = 0;
    steps64 = 0;
    const double inner_tol = f32 ? (o.inner_tol > 0 ? o.inner_tol : 1e-6) : 0.0;
    const uint32_t max_outer = o.max_outer ? o.max_outer : 20;
    const bool graph = o.launch != PP_LAUNCH_STREAM;
    bool have_enorm = false;
    relres = 1.0;
    for (;;) {
        PP_HIP(hipMemcpyAsync(r, b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        launch_spmv(ctx, A, PP_F64, -1.0, x, 1.0, r, -1, nullptr, false);
        launch_dot(ctx, PP_F64, n, r, r, S_TMP2);
        // S_TMP3 holds ||b||^2 on the first read, then ||e||^2 of the last FP32 reduced step
        double v[2];
        read_slots(ctx, S_TMP2, 2, v);
        const double rnorm = std::sqrt(v[0]);
        if (bnorm < 0) bnorm = std::sqrt(v[1]);
        relres = bnorm > 0 ? rnorm / bnorm : rnorm;
        if (!(relres == relres)) {
            res.status = PP_ERR_NOT_FINITE;
            break;
        }
        if (relres < o.tol) {
            res.conv = 1;
            break;
        }
        if (f32 && R && have_enorm && !inner64) {
            const double gap = kEps32 * R->anorm * std::sqrt(v[1]);
            if (resume && rnorm > 10.0 * step_claim * prev_rnorm) {
                inner64 = true;  // the resumed step fell short on the true residual
            } else if (kMargin * gap > step) {
                // the finished step's residual is not the recurrence's: restart,
                // with steps the gap allows, or in FP64 past kMaxStep
                resume = false;
                step = kMargin * gap;
                if (step > kMaxStep) inner64 = true;
            }
            if (inner64) resume = false;
        }
        have_enorm = false;
        if ((uint32_t)outer >= max_outer || (uint32_t)inner_total >= o.maxit) break;
        uint32_t inner_max = std::max<uint32_t>(2, o.maxit - inner_total);
        PcgResult in;
        if (f32 && R && inner64) {
            double *rin64 = (double *)ws_get(ctx, W_RIN64, (size_t)n * sizeof(double));
            double *e64 = (double *)ws_get(ctx, W_E64, (size_t)n * sizeof(double));
            launch_axpby(ctx, PP_F64, n, 1.0 / rnorm, r, 0.0, r, rin64);
            in = pcg_reduced_device<double>(ctx, R, rin64, e64, o.tol * bnorm / rnorm, inner_max, o.check_every,
                                            graph, false, 1.0, 1.0);
            inner_total += in.iters;
            ++outer;
            ++steps64;
            prev_rnorm = rnorm;
            if (in.status != PP_OK && in.status != PP_ERR_NOT_FINITE) {
                res.status = in.status;
                break;
            }
            launch_axpby(ctx, PP_F64, n, rnorm, e64, 1.0, x, x);
            continue;
        }
        if (f32) launch_scale_convert(ctx, n, r, 1.0 / rnorm, (float *)rin);
        else launch_axpby(ctx, PP_F64, n, 1.0 / rnorm, r, 0.0, r, rin);
        if (!R) launch_fill_zero(ctx, e, (size_t)n * sizeof(Tin));  // x0 = 0 (the reduced solve writes all of e)
        // FP32: a margin for the drift between the recurrence and the true
        // residual; FP64 inner solves stop where a plain FP64 solve would
        const double floor_tol = R ? std::max(inner_tol, step) : inner_tol;
        const double want = f32 ? std::max(floor_tol, 0.5 * o.tol * bnorm / rnorm) : o.tol * bnorm / rnorm;
        // a resumed step that has not converged within this many iterations
        // is ended, and the next step restarts
        if (resume) inner_max = std::min<uint32_t>(inner_max, std::max<uint32_t>(512, 4 * inner_total));
        if (R)
            in = pcg_reduced_device<Tin>(ctx, R, rin, e, want, inner_max, o.check_every, graph, resume,
                                         resume ? prev_rnorm / rnorm : 1.0, 1.0);  // rin = r / ||r||
        else if (o.cg_variant == PP_CG_REFERENCE)
            in = pcg_device<Tin>(ctx, A, M, o.flexible != 0, rin, e, want, inner_max, o.check_every, false,
                                 f32 ? W_R32 : W_R, graph);
        else in = pcg_sr_device<Tin>(ctx, A, M, rin, e, want, inner_max, o.check_every, false, graph);
        inner_total += in.iters;
        ++outer;
        if (!f32) ++steps64;
        prev_rnorm = rnorm;
        resume = R && in.conv;  // a step that stopped on its test is continued, any other restarts
        step_claim = want;
        if (in.status != PP_OK && in.status != PP_ERR_NOT_FINITE) {
            res.status = in.status;
            break;
        }
        if (f32 && R) {
            launch_dot(ctx, PP_F32, n, e, e, S_TMP3);
            have_enorm = true;
        }
        if (f32) launch_axpy_f32_to_f64(ctx, n, rnorm, (const float *)e, x);
        else launch_axpby(ctx, PP_F64, n, rnorm, e, 1.0, x, x);
    }
    res.iters = inner_total;
    res.final_error = relres;
    return res;
}

// ------------------------------------------------------------------ GMRES
struct Givens {
    std::vector<double> H;  // (m+1) x m column-major
    std::vector<double> c, s, g;
    uint32_t ld = 0;
    double &h(uint32_t i, uint32_t j) { return H[(size_t)j * ld + i]; }
};

// GMRES::givensTransform (source/Solver/GMRES.cpp:72-101).
void givens_step(Givens &G, uint32_t j) {
    for (uint32_t i = 0; i < j; ++i) {
        const double h1 = G.h(i, j), h2 = G.h(i + 1, j);
        G.h(i + 1, j) = G.c[i] * h2 - G.s[i] * h1;
        G.h(i, j) = G.c[i] * h1 + G.s[i] * h2;
    }
    const double h1 = G.h(j, j), h2 = G.h(j + 1, j);
    if (std::fabs(h1) > std::fabs(h2)) {
        const double t = h2 / h1;
        G.c[j] = 1.0 / std::sqrt(std::fabs(1.0 + t * t));
        G.s[j] = G.c[j] * t;
    } else {
        const double t = h1 / h2;
        G.s[j] = 1.0 / std::sqrt(std::fabs(1.0 + t * t));
        G.c[j] = G.s[j] * t;
    }
    G.h(j, j) = G.c[j] * h1 + G.s[j] * h2;
    G.h(j + 1, j) = 0.0;
    G.g[j + 1] = -G.s[j] * G.g[j];
    G.g[j] = G.c[j] * G.g[j];
}

// Back substitution R y = g over the leading dim x dim block (GMRES.cpp:103-119).
void back_solve(Givens &G, uint32_t dim) {
    for (int i = (int)dim - 1; i >= 0; --i) {
        double t = G.g[i];
        for (uint32_t j = i + 1; j < dim; ++j) t -= G.h(i, j) * G.g[j];
        G.g[i] = t / G.h(i, i);
    }
    G.g[dim] = 0.0;
}

// FlexibleGMRES(m) (source/Solver/GMRES.cpp:122-221) on the device: Z holds
// the preconditioned basis (FP32 when mixed), V the orthonormal basis (FP64),
// classical Gram-Schmidt applied twice (two GEMV^T + GEMV passes) in FP64;
// Givens rotations and the small least-squares problem on the host.
PcgResult gmres_device(pp_ctx *ctx, pp_csr *A, pp_precond *M, const pp_solve_opts &o, const double *b, double *x) {
    const uint32_t n = A->nrow;
    const bool mixed = o.precision == PP_MIXED;
    const bool flexible = o.flexible != 0;
    uint32_t mm = o.restart ? o.restart : 30;
    const int32_t maxcyc = (int32_t)((o.maxit ? o.maxit : 500) / mm);  // GMRES.cpp:63
    if (mm >= n) mm = n - 1;
    PP_CHECK(mm >= 1, PP_ERR_INVALID, "GMRES needs n >= 2");
    PP_CHECK(mm <= 256, PP_ERR_UNSUPPORTED, "restart > 256");
    const size_t ld = n;
    // mixed: the basis V, the preconditioned vectors Z and w in FP32 (half
    // the bytes of the orthogonalisation, which dominates a cycle), the CGS2
    // sums, the Hessenberg column and the least squares in FP64; every
    // restart forms the true residual in FP64, which makes it the
    // refinement step
    double *V = nullptr, *Z = nullptr;
    float *V32 = nullptr, *Z32 = nullptr;
    if (mixed) {
        csr_need_f32(ctx, A);
        if (M) precond_need_f32(M);
        V32 = (float *)ws_get(ctx, W_VB, ld * (mm + 1) * sizeof(float));
        launch_fill_zero(ctx, V32, ld * (mm + 1) * sizeof(float));
        Z32 = (float *)ws_get(ctx, W_ZB32, ld * (mm + 1) * sizeof(float));
        launch_fill_zero(ctx, Z32, ld * (mm + 1) * sizeof(float));
    } else {
        V = (double *)ws_get(ctx, W_VB, ld * (mm + 1) * sizeof(double));
        launch_fill_zero(ctx, V, ld * (mm + 1) * sizeof(double));
        if (flexible) {
            Z = (double *)ws_get(ctx, W_ZB, ld * (mm + 1) * sizeof(double));
            launch_fill_zero(ctx, Z, ld * (mm + 1) * sizeof(double));
        }
    }
    double *w = (double *)ws_get(ctx, W_W, (size_t)n * sizeof(double));
    double *r = (double *)ws_get(ctx, W_R, (size_t)n * sizeof(double));
    float *w32 = mixed ? (float *)ws_get(ctx, W_W32, (size_t)n * sizeof(float)) : nullptr;
    double *hdev = gmres_coeff_buffer(ctx);

    const GmresDev st = gmres_state(ctx, mm);
    const bool hist_on = o.record_history != 0;
    const uint32_t hist_cap = (uint32_t)(maxcyc + 1) * mm + 2;
    if (hist_on) ensure_hist(ctx, hist_cap);
    uint32_t nrec = 0;
    ctx->last_hist.clear();
    ctl_init(ctx, 1, 0);  // the sweeps honour the stop word: clear what a previous solve left

    launch_dot(ctx, PP_F64, n, b, b, S_TMP3);
    const double bnorm = std::sqrt(read_slot(ctx, S_TMP3));
    // r = b - A x ; beta = ||r||
    PP_HIP(hipMemcpyAsync(r, b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    launch_spmv(ctx, A, PP_F64, -1.0, x, 1.0, r, -1, nullptr, false);
    launch_dot(ctx, PP_F64, n, r, r, S_TMP2);
    double beta = std::sqrt(read_slot(ctx, S_TMP2));
    double rnorm = 0.0, relerr = 0.0;
    int conv = 0;
    int32_t k = -1;
    uint32_t m = mm;
    PcgResult res;
    if (!(beta == beta)) {
        res.status = PP_ERR_NOT_FINITE;
        return res;
    }
    if (!o.reference_quirks && beta / (bnorm > 0 ? bnorm : 1.0) < o.tol) {
        res.conv = 1;
        res.iters = 0;
        res.final_error = beta / (bnorm > 0 ? bnorm : 1.0);
        return res;
    }
    const int quirks = o.reference_quirks ? 1 : 0;
    const bool graph = o.launch != PP_LAUNCH_STREAM;
    while (!conv && k < maxcyc) {
        ++k;
        // One Arnoldi cycle, issued without host round trips: the Hessenberg
        // column, Givens rotations and convergence test run on the device
        // (k_gmres_step) and the stop word ends the cycle's remaining kernels.
        ctl_init(ctx, 1, 0);
        launch_gmres_cycle_init(ctx, st, beta, bnorm, o.tol, nrec);
        if (mixed) launch_scale_convert(ctx, n, r, 1.0 / beta, V32);  // v0 = r / beta
        else launch_axpby(ctx, PP_F64, n, 1.0 / beta, r, 0.0, r, V);
        auto cycle = [&] {
            StopScope scope(ctx);
            for (uint32_t j = 0; j < m; ++j) {
                if (mixed) {
                    // z_j = M^{-1} v_j ; w = A z_j ; CGS2 and the next basis vector on the FP32 basis
                    float *vj = V32 + ld * j, *zj = Z32 + ld * j;
                    if (flexible && M) precond_apply(ctx, M, PP_F32, vj, zj);
                    launch_spmv(ctx, A, PP_F32, 1.0, flexible && M ? zj : vj, 0.0, w32, -1, nullptr, false);
                    launch_gemv_t(ctx, n, j + 1, V32, ld, w32, st.h1);
                    launch_gemv_sub_t(ctx, n, j + 1, V32, ld, st.h1, w32, st.h2);
                    launch_gemv_sub(ctx, n, j + 1, V32, ld, st.h2, w32);
                    launch_dot(ctx, PP_F32, n, w32, w32, S_TMP2);
                    launch_gmres_step(ctx, st, j, quirks, hist_on ? ctx->hist : nullptr, j, hist_cap);
                    launch_gmres_next_basis(ctx, st, n, w32, V32 + ld * (j + 1));
                    continue;
                }
                double *vj = V + ld * j;
                // z_j = M^{-1} v_j ; w = A z_j
                if (flexible) {
                    double *zj = Z + ld * j;
                    if (M) precond_apply(ctx, M, PP_F64, vj, zj);
                    else launch_convert(ctx, n, PP_F64, vj, PP_F64, zj);
                    launch_spmv(ctx, A, PP_F64, 1.0, zj, 0.0, w, -1, nullptr, false);
                } else {
                    launch_spmv(ctx, A, PP_F64, 1.0, vj, 0.0, w, -1, nullptr, false);
                }
                // CGS2: h1 = V^T w, w -= V h1, h2 = V^T w, w -= V h2
                launch_gemv_t(ctx, n, j + 1, V, ld, w, st.h1);
                launch_gemv_sub_t(ctx, n, j + 1, V, ld, st.h1, w, st.h2);  // w -= V h1; h2 = V^T w
                launch_gemv_sub(ctx, n, j + 1, V, ld, st.h2, w);
                launch_dot(ctx, PP_F64, n, w, w, S_TMP2);
                launch_gmres_step(ctx, st, j, quirks, hist_on ? ctx->hist : nullptr, j, hist_cap);
                launch_gmres_next_basis(ctx, st, n, w, V + ld * (j + 1));
            }
        };
        // the cycle's launches are the same every cycle of the same length:
        // captured once into a graph and replayed (no host launch gaps)
        if (graph) {
            const std::vector<uint64_t> key = {4, m, (uint64_t)mixed, (uint64_t)flexible, (uint64_t)quirks, A->uid,
                                               (uint64_t)A->v32, (uint64_t)A->spmv_variant << 8 | A->spmv_variant32,
                                               M ? M->uid : 0, M ? (uint64_t)M->io_permuted : 0,
                                               M ? (uint64_t)M->have32 : 0, (uint64_t)V, (uint64_t)V32,
                                               (uint64_t)Z, (uint64_t)Z32,
                                               (uint64_t)w, (uint64_t)w32, (uint64_t)st.misc,
                                               (uint64_t)(hist_on ? ctx->hist : nullptr), hist_cap, ld,
                                               variant_key(A, M)};
            run_graph(ctx, key, cycle);
        } else {
            cycle();
        }
        double misc[kGmMisc];
        PP_HIP(hipMemcpyAsync(ctx->pinned, st.misc, kGmMisc * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        PP_HIP(hipStreamSynchronize(ctx->stream));
        std::copy(ctx->pinned, ctx->pinned + kGmMisc, misc);
        {
            int ctl[C_COUNT];
            read_ctl(ctx, ctl);
            PP_CHECK(!ctl[C_SPIN], PP_ERR_HIP, "sync-free triangular sweep timed out waiting for a dependency");
        }
        nrec += (uint32_t)misc[GM_NREC];
        m = (uint32_t)misc[GM_M];
        rnorm = misc[GM_RNORM];
        relerr = misc[GM_RELERR];
        if ((int)misc[GM_STATUS] != PP_OK) {
            res.status = (int)misc[GM_STATUS];
            break;
        }
        // x += Z y over all mm+1 columns (GMRES.cpp:204-208); in the
        // reference's converged cycle the entry past the back-solve is the
        // residual estimate, left in place
        double *y = hdev + 320;
        launch_gmres_backsolve(ctx, st, quirks, y);
        launch_gemv_add(ctx, n, mm + 1, mixed ? (flexible && M ? Z32 : V32) : nullptr,
                        mixed ? nullptr : (flexible ? Z : V), ld, y, x);
        conv = relerr < o.tol;
        PP_HIP(hipMemcpyAsync(r, b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        launch_spmv(ctx, A, PP_F64, -1.0, x, 1.0, r, -1, nullptr, false);
        if (o.reference_quirks) {
            beta = rnorm;  // GMRES.cpp:214 restarts from the estimate
        } else {
            launch_dot(ctx, PP_F64, n, r, r, S_TMP2);
            beta = std::sqrt(read_slot(ctx, S_TMP2));
            relerr = beta / bnorm;
            conv = relerr < o.tol;
        }
        if (!conv && k < maxcyc) {
            if (!o.reference_quirks) m = mm;
        }
    }
    if (hist_on && nrec) {
        const uint32_t cnt = std::min(nrec, ctx->hist_cap);
        ctx->last_hist.assign(cnt, 0.0);
        PP_HIP(hipMemcpy(ctx->last_hist.data(), ctx->hist, cnt * sizeof(double), hipMemcpyDeviceToHost));
    }
    {
        int ctl[C_COUNT];
        read_ctl(ctx, ctl);
        PP_CHECK(!ctl[C_SPIN], PP_ERR_HIP, "sync-free triangular sweep timed out waiting for a dependency");
    }
    res.conv = conv;
    res.iters = k * (int32_t)mm + (int32_t)m;  // GMRES.cpp:218
    if (!o.reference_quirks && conv && m > 0) res.iters = k * (int32_t)mm + (int32_t)m - 1;
    res.final_error = relerr;
    res.nhist = (int)ctx->last_hist.size();
    return res;
}

}  // namespace
}  // namespace pp

using namespace pp;

template <class F>
static int guarded_solve(pp_ctx *ctx, F &&f) {
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

static int solve_impl(pp_ctx *ctx, pp_csr *A, pp_precond *M, const pp_solve_opts *opts, const double *b, double *x,
                      pp_solve_report *rep) {
    return guarded_solve(ctx, [&] {
        PP_CHECK(A && opts && b && x, PP_ERR_INVALID, "NULL argument");
        PP_CHECK(A->nrow == A->ncol, PP_ERR_INVALID, "square matrix required");
        if (M) PP_CHECK(M->ready, PP_ERR_NOT_READY, "preconditioner used before setup()");
        pp_solve_report R;
        std::memset(&R, 0, sizeof(R));
        const pp_solve_opts &o = *opts;
        PP_CHECK(o.tol > 0, PP_ERR_INVALID, "tol must be positive");
        PP_CHECK(o.maxit >= 1, PP_ERR_INVALID, "maxit must be >= 1");
        PP_CHECK(o.solver == PP_SOLVER_CG || o.solver == PP_SOLVER_GMRES, PP_ERR_INVALID, "unknown solver");
        // A factor of P A P^T (multicolour order): run the whole solve on
        // P A P^T with P b and P x0, so no apply gathers or scatters; x is
        // scattered back once at the end.
        const uint32_t n = A->nrow;
        pp_csr *As = A;
        const double *bs = b;
        double *xs = x;
        const bool permuted = M && !M->order.empty() && !M->multilevel && M->A == A;
        if (permuted) {
            As = precond_permuted_matrix(ctx, M, A);
            double *bp = (double *)ws_get(ctx, W_PB, (size_t)n * sizeof(double));
            double *xp = (double *)ws_get(ctx, W_PX, (size_t)n * sizeof(double));
            launch_permute(ctx, n, M->dorder, PP_F64, b, bp, true);
            launch_permute(ctx, n, M->dorder, PP_F64, x, xp, true);
            bs = bp;
            xs = xp;
        }
        struct Unpermute {
            pp_precond *M;
            ~Unpermute() {
                if (M) M->io_permuted = false;
            }
        } unpermute{permuted ? M : nullptr};
        if (permuted) M->io_permuted = true;
        PP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        PcgResult res;
        // the refinement loop ends on an FP64 true residual of the final x
        // (on the colour-permuted system when permuted: the same vector)
        double refined_relres = -1.0;
        if (o.solver == PP_SOLVER_CG) {
            ReducedForm *red = nullptr;
            if (permuted && (o.cg_variant == PP_CG_AUTO || o.cg_variant == PP_CG_REDUCED) &&
                (o.precision == PP_MIXED || o.cg_variant == PP_CG_REDUCED))
                red = reduced_form(ctx, M, As);
            PP_CHECK(o.cg_variant != PP_CG_REDUCED || red, PP_ERR_UNSUPPORTED,
                     "reduced CG needs IC(0) in a two-colour (multicolour) ordering of a matrix whose colour "
                     "blocks are diagonal");
            if (o.precision == PP_MIXED) {
                double &rel = refined_relres;
                res = pcg_refine<float>(ctx, As, M, red, o, bs, xs, R.outer_iters, R.inner_iters, R.inner_fp64_steps,
                                        rel);
                R.cg_variant = red ? PP_CG_REDUCED
                                   : (o.cg_variant == PP_CG_REFERENCE ? PP_CG_REFERENCE : PP_CG_SINGLE_REDUCTION);
            } else if (red) {
                double &rel = refined_relres;
                res = pcg_refine<double>(ctx, As, M, red, o, bs, xs, R.outer_iters, R.inner_iters, R.inner_fp64_steps,
                                         rel);
                R.cg_variant = PP_CG_REDUCED;
            } else {
                R.cg_variant = o.cg_variant == PP_CG_SINGLE_REDUCTION ? PP_CG_SINGLE_REDUCTION : PP_CG_REFERENCE;
                if (o.cg_variant == PP_CG_SINGLE_REDUCTION)
                    res = pcg_sr_device<double>(ctx, As, M, bs, xs, o.tol, o.maxit, o.check_every ? o.check_every : 1,
                                                o.record_history != 0, o.launch != PP_LAUNCH_STREAM);
                else
                    res = pcg_device<double>(ctx, As, M, o.flexible != 0, bs, xs, o.tol, o.maxit,
                                             o.check_every ? o.check_every : 1, o.record_history != 0, W_R,
                                             o.launch != PP_LAUNCH_STREAM);
            }
        } else {
            res = gmres_device(ctx, As, M, o, bs, xs);
        }
        if (permuted) launch_permute(ctx, n, M->dorder, PP_F64, xs, x, false);
        PP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        PP_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0;
        PP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        R.iters = res.iters;
        R.converged = res.conv;
        R.final_error = res.final_error;
        R.status = res.status;
        R.nhist = res.nhist;
        R.solve_ms = ms;
        R.true_relres = refined_relres >= 0 ? refined_relres : true_relres(ctx, A, b, x);
        if (rep) *rep = R;
        if (res.status != PP_OK) {
            throw Error(res.status, res.status == PP_ERR_BREAKDOWN ? "Krylov breakdown ((d, A d) = 0)"
                                                                   : "non-finite value in the iteration");
        }
    });
}

extern "C" {

void pp_solve_opts_default(pp_solve_opts *o, int solver) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->solver = solver;
    o->flexible = 1;
    o->precision = PP_FP64;
    o->tol = 1e-8;
    o->maxit = solver == PP_SOLVER_GMRES ? 500 : 1000;
    o->restart = 30;
    o->reference_quirks = 0;
    o->check_every = 4;
    o->launch = PP_LAUNCH_STREAM;
    o->inner_tol = 1e-6;
    o->max_outer = 20;
    o->record_history = 1;
}

int pp_solve(pp_ctx *ctx, pp_csr *A, pp_precond *M, const pp_solve_opts *opts, const double *b, double *x,
             pp_solve_report *rep) {
    if (!ctx) return PP_ERR_INVALID;
    return solve_impl(ctx, A, M, opts, b, x, rep);
}

int pp_solve_host(pp_ctx *ctx, pp_csr *A, pp_precond *M, const pp_solve_opts *opts, const double *b_host,
                  double *x_host, pp_solve_report *rep) {
    if (!ctx) return PP_ERR_INVALID;
    double *db = nullptr, *dx = nullptr;
    int st = guarded_solve(ctx, [&] {
        PP_CHECK(A && b_host && x_host, PP_ERR_INVALID, "NULL argument");
        const uint32_t n = A->nrow;
        db = dalloc<double>(n);
        dx = dalloc<double>(n);
        PP_HIP(hipMemcpyAsync(db, b_host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        PP_HIP(hipMemcpyAsync(dx, x_host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    });
    if (st == PP_OK) st = solve_impl(ctx, A, M, opts, db, dx, rep);
    if (st == PP_OK || st == PP_ERR_BREAKDOWN || st == PP_ERR_NOT_FINITE) {
        int st2 = guarded_solve(ctx, [&] {
            PP_HIP(hipMemcpyAsync(x_host, dx, A->nrow * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            PP_HIP(hipStreamSynchronize(ctx->stream));
        });
        if (st == PP_OK) st = st2;
    }
    dfree(db);
    dfree(dx);
    return st;
}

// The FP64 CG solve of pp_solve with every work vector carved from the
// fixed scratch array of pp_noalloc.hip: no device or host allocation, no
// graph instantiation (the iteration kernels are launched on the stream).
int pp_solve_noalloc(pp_ctx *ctx, pp_csr *A, pp_precond *M, const pp_solve_opts *opts, const double *b, double *x,
                     pp_solve_report *rep) {
    if (!ctx) return PP_ERR_INVALID;
    pp_solve_opts o;
    unsigned char *base = nullptr;
    int st = guarded_solve(ctx, [&] {
        PP_CHECK(A && opts && b && x, PP_ERR_INVALID, "NULL argument");
        PP_CHECK(A->nrow == A->ncol, PP_ERR_INVALID, "square matrix required");
        o = *opts;
        PP_CHECK(o.solver == PP_SOLVER_CG && o.precision == PP_FP64, PP_ERR_UNSUPPORTED,
                 "the zero-allocation solve runs FP64 CG only");
        PP_CHECK(o.cg_variant == PP_CG_AUTO || o.cg_variant == PP_CG_REFERENCE ||
                     o.cg_variant == PP_CG_SINGLE_REDUCTION,
                 PP_ERR_UNSUPPORTED, "the zero-allocation solve runs FlexibleCG / NormalCG or the single-reduction CG");
        bool permuted = false;
        if (M) {
            PP_CHECK(M->ready, PP_ERR_NOT_READY, "preconditioner used before setup()");
            PP_CHECK(!M->multilevel && M->opts.kind != PP_PC_AMSED && M->opts.kind != PP_PC_GMSLR, PP_ERR_UNSUPPORTED,
                     "the zero-allocation solve takes triangular preconditioners only");
            permuted = !M->order.empty() && M->A == A;
            // the colour-ordered copy of A exists from setup and is current
            PP_CHECK(!permuted || (M->Aperm && M->Aperm_src_uid == A->uid && M->Aperm_version == A->version),
                     PP_ERR_NOT_READY,
                     "the matrix values changed since the preconditioner's setup: set it up again first");
        }
        base = (unsigned char *)noalloc_scratch(ctx->device);
        PP_CHECK(base, PP_ERR_HIP, "the scratch array has no address on this device");
        // carved slots: the iteration's vectors (four, five for the single
        // reduction), the final residual, and b and x in colour order
        const size_t vec = ((size_t)A->nrow * sizeof(double) + 255) & ~(size_t)255;
        const size_t nvec = (o.cg_variant == PP_CG_SINGLE_REDUCTION ? 5 : 4) + 1 + (permuted ? 2 : 0);
        if (vec * nvec > kNoallocScratchBytes) {
            char msg[200];
            std::snprintf(msg, sizeof msg,
                          "n = %u needs %zu bytes of scratch; the zero-allocation solve holds at most %zu "
                          "(n <= %zu here)",
                          A->nrow, vec * nvec, kNoallocScratchBytes, kNoallocScratchBytes / nvec / sizeof(double));
            throw Error(PP_ERR_INVALID, msg);
        }
        o.launch = PP_LAUNCH_STREAM;
        o.record_history = 0;
    });
    if (st != PP_OK) return st;
    struct Arena {
        pp_ctx *c;
        ~Arena() {
            c->arena = nullptr;
            c->arena_cap = c->arena_used = 0;
            for (int i = 0; i < 48; ++i) c->arena_off[i] = c->arena_len[i] = 0;
        }
    } arena{ctx};
    ctx->arena = base;
    ctx->arena_cap = kNoallocScratchBytes;
    ctx->arena_used = 0;
    for (int i = 0; i < 48; ++i) ctx->arena_off[i] = ctx->arena_len[i] = 0;
    return solve_impl(ctx, A, M, &o, b, x, rep);
}

int pp_solve_history(pp_ctx *ctx, double *out, uint32_t cap, uint32_t *count) {
    if (!ctx || !count) return PP_ERR_INVALID;
    uint32_t c = (uint32_t)ctx->last_hist.size();
    *count = c;
    if (out) std::memcpy(out, ctx->last_hist.data(), std::min(c, cap) * sizeof(double));
    return PP_OK;
}

// Adjoint: A^T lambda = g with a preconditioner built on A^T.
int pp_adjoint_solve(pp_ctx *ctx, pp_csr *A, const pp_precond_opts *popts, const pp_solve_opts *opts,
                     const double *g, double *lambda, pp_solve_report *rep) {
    if (!ctx) return PP_ERR_INVALID;
    pp_csr *T = nullptr;
    int st = guarded_solve(ctx, [&] { T = csr_transpose(ctx, A); });
    if (st != PP_OK) return st;
    pp_precond *M = nullptr;
    if (popts && popts->kind != PP_PC_NONE) {
        st = pp_precond_create(ctx, T, popts, &M);
        if (st == PP_OK) st = pp_precond_setup(ctx, M);
        if (st != PP_OK) {
            pp_precond_destroy(M);
            return st;
        }
    }
    st = solve_impl(ctx, T, M, opts, g, lambda, rep);
    pp_precond_destroy(M);
    return st;
}

}  // extern "C"
