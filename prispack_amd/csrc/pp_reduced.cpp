// pp_reduced.cpp -- IC(0) of a two-colour matrix in Schur-complement form
// (the operators of the reduced CG; see ReducedForm in pp_internal.h).
// Host-side construction from the colour-ordered matrix, once per matrix
// version; the device work is SpMV (k_spmv) on the four operators.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "pp_internal.h"

namespace pp {
namespace {

// Colour split of a colour-ordered matrix: n0 = the first row with an
// entry left of its diagonal.  Valid when every first-colour row couples
// only to the second colour and vice versa, and every row has one diagonal
// entry.
bool split_two_colours(uint32_t n, const std::vector<uint32_t> &rp, const std::vector<uint32_t> &ci, uint32_t &n0) {
    n0 = n;
    for (uint32_t i = 0; i < n && n0 == n; ++i)
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] < i) {
                n0 = i;
                break;
            }
    if (n0 == 0 || n0 >= n) return false;
    for (uint32_t i = 0; i < n; ++i) {
        bool diag = false;
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const uint32_t j = ci[k];
            if (j == i) {
                if (diag) return false;  // a repeated diagonal entry
                diag = true;
                continue;
            }
            if ((i < n0) == (j < n0)) return false;  // coupling inside a colour
        }
        if (!diag) return false;
    }
    return true;
}

}  // namespace

bool reduced_host(uint32_t n, const std::vector<uint32_t> &rp, const std::vector<uint32_t> &ci,
                  const std::vector<double> &av, ReducedHost &H) {
    uint32_t n0 = 0;
    if (!split_two_colours(n, rp, ci, n0)) return false;
    std::vector<double> d(n);
    std::atomic<bool> good{true};
    parallel_for(n, 1u << 15, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) {
            d[i] = 0.0;
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] == i) {
                    d[i] = av[k];
                    break;
                }
            if (!(d[i] > 0.0)) good = false;
        }
    });
    if (!good) return false;
    // Dt_i = a_ii - sum_j a_ij^2 / a_jj over the first-colour neighbours
    // (the IC(0) pivot of a second-colour row)
    std::vector<double> &dt = H.dt;
    dt.assign(n - n0, 0.0);
    parallel_for(n - n0, 1u << 15, [&](size_t q0, size_t q1) {
        for (size_t q = q0; q < q1; ++q) {
            const uint32_t i = n0 + (uint32_t)q;
            double t = d[i];
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] < n0) t -= av[k] * av[k] / d[ci[k]];
            dt[q] = t;
            if (!(t > 0.0 && std::isfinite(t))) good = false;
        }
    });
    if (!good) return false;
    const uint32_t n1 = n - n0;
    H.n0 = n0;
    H.n1 = n1;
    H.anorm = 0.0;
    {
        std::mutex mu;
        parallel_for(n, 1u << 15, [&](size_t i0, size_t i1) {
            double m = 0.0;
            for (size_t i = i0; i < i1; ++i) {
                double t = 0.0;
                for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) t += std::fabs(av[k]);
                m = std::max(m, t);
            }
            std::lock_guard<std::mutex> g(mu);
            H.anorm = std::max(H.anorm, m);
        });
    }
    // every row holds one diagonal: row lengths of the four operators follow
    // from A's (KA: the off-diagonals; KE, KB, KF: the off-diagonals and one more)
    auto rows_of = [&](uint32_t r0, uint32_t r1, uint32_t extra) {
        std::vector<uint32_t> p(r1 - r0 + 1, 0);
        for (uint32_t i = r0; i < r1; ++i) p[i - r0 + 1] = p[i - r0] + (rp[i + 1] - rp[i] - 1) + extra;
        return p;
    };
    H.ka_rp = rows_of(0, n0, 0);
    H.ke_rp = rows_of(0, n0, 1);
    H.kb_rp = rows_of(n0, n, 1);
    H.kf_rp = rows_of(n0, n, 1);
    H.ka_ci.resize(H.ka_rp.back());
    H.ke_ci.resize(H.ke_rp.back());
    H.kb_ci.resize(H.kb_rp.back());
    H.kf_ci.resize(H.kf_rp.back());
    H.ka_v.resize(H.ka_rp.back());
    H.ke_v.resize(H.ke_rp.back());
    H.kb_v.resize(H.kb_rp.back());
    H.kf_v.resize(H.kf_rp.back());
    H.s1.resize(n1);
    parallel_for(n0, 1u << 15, [&](size_t i0, size_t i1) {
        for (uint32_t i = (uint32_t)i0; i < (uint32_t)i1; ++i) {
            uint32_t pa = H.ka_rp[i], pe = H.ke_rp[i];
            H.ke_ci[pe] = i;
            H.ke_v[pe++] = 1.0 / d[i];
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const uint32_t j = ci[k];
                if (j == i) continue;
                const double a = av[k];
                H.ka_ci[pa] = j;
                H.ka_v[pa++] = a / std::sqrt(d[i] * dt[j - n0]);
                H.ke_ci[pe] = j;
                H.ke_v[pe++] = -a / d[i];
            }
        }
    });
    parallel_for(n1, 1u << 15, [&](size_t q0, size_t q1) {
        for (uint32_t q = (uint32_t)q0; q < (uint32_t)q1; ++q) {
            const uint32_t i = n0 + q;
            const double t = dt[q], rs = 1.0 / std::sqrt(t);
            uint32_t pb = H.kb_rp[q], pf = H.kf_rp[q];
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const uint32_t j = ci[k];
                if (j == i) continue;
                const double a = av[k];
                H.kb_ci[pb] = j;
                H.kb_v[pb++] = -a / std::sqrt(d[j] * t);
                H.kf_ci[pf] = j;
                H.kf_v[pf++] = -a / d[j] * rs;
            }
            H.kb_ci[pb] = i;  // the diagonal column i > every first-colour column
            H.kb_v[pb] = d[i] / t;
            H.kf_ci[pf] = i;
            H.kf_v[pf] = rs;
            H.s1[q] = rs;
        }
    });
    return true;
}

void EllOp::free_all() {
    dfree(off);
    dfree(pid);
    dfree(ptab);
    pid = nullptr;
    ptab = nullptr;
    npat = 0;
    dfree(v64);
    dfree(v32);
    dfree(dg64);
    dfree(dg32);
    off = nullptr;
    v64 = dg64 = nullptr;
    v32 = dg32 = nullptr;
    ok = false;
    n = width = 0;
    csplit = UINT32_MAX;
    dsplit = 0;
    cbase1 = 0;
}

namespace {

// The offsets' pattern table and the upload of a sliced-column form whose
// slots (off, val: entry j of row i at j n + i) and separate diagonal (dg,
// empty for none) are built: E's geometry (n, width, bases, splits) is set
// by the caller.
void ell_finish(EllOp &E, uint32_t n, uint32_t width, const std::vector<int16_t> &off,
                const std::vector<double> &val, const std::vector<double> &dg) {
    // Row patterns: when the rows' offset tuples take at most 256 distinct
    // values (a stencil: interior rows share one, the boundary rows a few),
    // a row stores a one-byte pattern number instead of width 16-bit
    // offsets, and the kernels read the row's offsets from a 16-byte table
    // entry (cached): 2 width - 1 bytes less per row.  For the widths the
    // kernels are instantiated at exactly (4 to 7).
    std::vector<uint8_t> pid;
    std::vector<int16_t> ptab;
    if (width >= 4 && width <= 7) {
        // Numbered in order of first appearance, found in parallel: each
        // chunk of rows lists its own patterns (consecutive rows mostly
        // repeat one: a 16-byte compare), the lists are merged in chunk
        // order (which keeps the first-appearance order of a serial pass),
        // and each row's chunk-local number is mapped to the merged one.
        using Key = std::array<int16_t, 8>;
        auto key_of = [&](uint32_t i) {
            Key k{};
            for (uint32_t j = 0; j < width; ++j) k[j] = off[(size_t)j * n + i];
            return k;
        };
        const uint32_t cs = 1u << 16, nchunk = (n + cs - 1) / cs;
        std::vector<std::vector<Key>> local(nchunk);
        pid.resize(n);
        std::atomic<bool> ok{true};
        parallel_for(nchunk, 1, [&](size_t c0, size_t c1) {
            for (size_t c = c0; c < c1 && ok; ++c) {
                std::vector<Key> &pats = local[c];
                int last = -1;
                for (uint32_t i = (uint32_t)c * cs, e = std::min(n, (uint32_t)(c + 1) * cs); i < e; ++i) {
                    const Key k = key_of(i);
                    if (last >= 0 && k == pats[last]) {
                        pid[i] = (uint8_t)last;
                        continue;
                    }
                    last = -1;
                    for (size_t q = 0; q < pats.size(); ++q)
                        if (pats[q] == k) {
                            last = (int)q;
                            break;
                        }
                    if (last < 0) {
                        if (pats.size() == 256) {
                            ok = false;
                            return;
                        }
                        last = (int)pats.size();
                        pats.push_back(k);
                    }
                    pid[i] = (uint8_t)last;
                }
            }
        });
        std::vector<std::vector<uint8_t>> remap(nchunk);
        std::vector<Key> all;
        for (uint32_t c = 0; c < nchunk && ok; ++c)
            for (const Key &k : local[c]) {
                size_t q = std::find(all.begin(), all.end(), k) - all.begin();
                if (q == all.size()) {
                    if (all.size() == 256) {
                        ok = false;
                        break;
                    }
                    all.push_back(k);
                }
                remap[c].push_back((uint8_t)q);
            }
        if (ok) {
            parallel_for(n, cs, [&](size_t i0, size_t i1) {
                for (size_t i = i0; i < i1; ++i) pid[i] = remap[i / cs][pid[i]];
            });
            for (const Key &k : all) ptab.insert(ptab.end(), k.begin(), k.end());
        } else {
            pid.clear();
        }
    }
    if (!pid.empty()) {
        E.npat = (uint32_t)(ptab.size() / 8);
        E.pid = dalloc<uint8_t>(n + 1);  // one more byte: a two-row load at the last even row stays inside
        E.ptab = dalloc<int16_t>(ptab.size());
        PP_HIP(hipMemcpy(E.pid, pid.data(), n, hipMemcpyHostToDevice));
        PP_HIP(hipMemcpy(E.ptab, ptab.data(), ptab.size() * sizeof(int16_t), hipMemcpyHostToDevice));
    } else {
        E.off = dalloc<int16_t>(off.size());
        PP_HIP(hipMemcpy(E.off, off.data(), off.size() * sizeof(int16_t), hipMemcpyHostToDevice));
    }
    E.v64 = dalloc<double>(val.size());
    PP_HIP(hipMemcpy(E.v64, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!dg.empty()) {
        E.dg64 = dalloc<double>(n);
        PP_HIP(hipMemcpy(E.dg64, dg.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    E.ok = true;
}

}  // namespace



// diag_last: the last entry of every row is its diagonal term, at column
// dbase + row; it goes to dg instead of the sliced columns.
// max_fill: the largest padded size width * n allowed, relative to the
// entries kept.
bool ell_build(const std::vector<uint32_t> &rp, const std::vector<uint32_t> &ci, const std::vector<double> &v,
               int64_t cbase, bool diag_last, int64_t dbase, EllOp &E, double max_fill) {
    E.free_all();
    const uint32_t n = (uint32_t)rp.size() - 1;
    const uint32_t tail = diag_last ? 1 : 0;
    uint32_t width = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t len = rp[i + 1] - rp[i];
        if (len <= tail) return false;  // a row with no sliced entry
        width = std::max(width, len - tail);
    }
    if (n == 0 || width > 16 || (uint64_t)width * n >= (1ull << 32)) return false;
    if ((double)width * n > max_fill * (double)(rp[n] - (uint64_t)tail * n)) return false;
    std::vector<int16_t> off((size_t)width * n);
    std::vector<double> val((size_t)width * n), dg(diag_last ? n : 0);
    std::atomic<bool> fits{true};
    parallel_for(n, 1u << 15, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) {
            const uint32_t k0 = rp[i], k1 = rp[i + 1] - tail;
            int16_t last = 0;
            for (uint32_t j = 0; j < width; ++j) {
                const size_t e = (size_t)j * n + i;
                if (k0 + j < k1) {
                    const int64_t o = (int64_t)ci[k0 + j] - (int64_t)i - cbase;
                    if (o < INT16_MIN || o > INT16_MAX) {
                        fits = false;
                        return;
                    }
                    last = (int16_t)o;
                    off[e] = last;
                    val[e] = v[k0 + j];
                } else {  // padding: a zero on the row's last column
                    off[e] = last;
                    val[e] = 0.0;
                }
            }
            if (diag_last) {
                if ((int64_t)ci[k1] != dbase + (int64_t)i) fits = false;
                dg[i] = v[k1];
            }
        }
    });
    if (!fits) return false;
    E.n = n;
    E.width = width;
    E.cbase = cbase;
    E.dbase = dbase;
    ell_finish(E, n, width, off, val, dg);
    return true;
}

// A two-colour matrix (rows [0, split) couple only to [split, n) and back,
// each row holding its diagonal): rows below split take their columns from
// row + split (cbase), the others from row - split (cbase1); the diagonal
// goes to dg, first in its row's sum below split (where it is the row's
// first entry) and last from split on, as the CSR rows order them.
bool ell_build_two_colour(const std::vector<uint32_t> &rp, const std::vector<uint32_t> &ci,
                          const std::vector<double> &v, uint32_t split, EllOp &E, double max_fill) {
    E.free_all();
    const uint32_t n = (uint32_t)rp.size() - 1;
    if (split == 0 || split >= n) return false;
    uint32_t width = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t len = rp[i + 1] - rp[i];
        if (len < 2) return false;  // a row with no sliced entry
        width = std::max(width, len - 1);
    }
    if (width > 16 || (uint64_t)width * n >= (1ull << 32)) return false;
    if ((double)width * n > max_fill * (double)(rp[n] - (uint64_t)n)) return false;
    const int64_t cb0 = (int64_t)split, cb1 = -(int64_t)split;
    std::vector<int16_t> off((size_t)width * n);
    std::vector<double> val((size_t)width * n), dg(n);
    std::atomic<bool> fits{true};
    parallel_for(n, 1u << 15, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1 && fits; ++i) {
            const bool lo = i < split;
            // the diagonal: first entry below split, last from split on
            const uint32_t kd = lo ? rp[i] : rp[i + 1] - 1;
            if (ci[kd] != i) {
                fits = false;
                return;
            }
            dg[i] = v[kd];
            const uint32_t k0 = lo ? rp[i] + 1 : rp[i], k1 = lo ? rp[i + 1] : rp[i + 1] - 1;
            const int64_t cb = lo ? cb0 : cb1;
            int16_t last = 0;
            for (uint32_t j = 0; j < width; ++j) {
                const size_t e = (size_t)j * n + i;
                if (k0 + j < k1) {
                    const uint32_t c = ci[k0 + j];
                    const int64_t o = (int64_t)c - (int64_t)i - cb;
                    if ((c < split) == lo || o < INT16_MIN || o > INT16_MAX) {  // a coupling inside a colour
                        fits = false;
                        return;
                    }
                    last = (int16_t)o;
                    off[e] = last;
                    val[e] = v[k0 + j];
                } else {  // padding: a zero on the row's last column
                    off[e] = last;
                    val[e] = 0.0;
                }
            }
        }
    });
    if (!fits) return false;
    E.n = n;
    E.width = width;
    E.cbase = cb0;
    E.cbase1 = cb1;
    E.csplit = split;
    E.dsplit = split;
    E.dbase = 0;
    ell_finish(E, n, width, off, val, dg);
    return true;
}

ReducedForm *reduced_form(pp_ctx *ctx, pp_precond *M, pp_csr *As) {
    if (!M || M->opts.kind != PP_PC_IC0 || M->order.empty() || !M->io_permuted || M->multilevel || !As) return nullptr;
    // keyed on the matrix's uid (never reused, unlike an address) and its value version
    if (M->reduced && M->reduced->src_uid == As->uid && M->reduced->src_version == As->version)
        return M->reduced->ok ? M->reduced : nullptr;
    // the same matrix with new values on its pattern (a re-factorization):
    // the operators' values recomputed in place on the device
    const auto t0 = std::chrono::steady_clock::now();
    auto ms = [&] {
        return (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    if (M->reduced && M->reduced->ok && M->reduced->src_uid == As->uid && reduced_refresh(ctx, M->reduced, As)) {
        M->reduced->src_version = As->version;
        M->reduced->built = 3;
        M->reduced->build_ms = ms();
        return M->reduced;
    }
    reduced_free(M);
    PoolScope pool(ctx);
    std::unique_ptr<ReducedForm> R(new ReducedForm);
    R->src_uid = As->uid;
    R->src_version = As->version;
    // on the device from the matrix's device arrays when its sliced-column
    // copy serves the operators too, else on the host (always on the host
    // with PP_REDUCED_HOST_BUILD=1 in the environment: the reference
    // construction the device one is tested against)
    const char *host_only = std::getenv("PP_REDUCED_HOST_BUILD");
    try {
        if (!(host_only && host_only[0] == '1') && reduced_build_device(ctx, M, As, R.get()) >= 0) {
            R->built = 2;
            R->build_ms = ms();
            M->reduced = R.release();
            return M->reduced->ok ? M->reduced : nullptr;
        }
    } catch (...) {
        M->reduced = R.release();
        reduced_free(M);
        throw;
    }
    ReducedHost H;
    csr_ensure_host(As);  // a matrix built on the device: its host copy first
    // the factor's first level must be exactly the first colour
    const bool ok = As->nrow == As->ncol && reduced_host(As->nrow, As->h_rp, As->h_ci, As->h_v, H) &&
                    M->low64.nlev == 2 && M->up64.nlev == 2 && M->lvl0_contiguous && M->lvl0_lo == 0 &&
                    M->lvl0_hi == H.n0;
    if (!ok) {
        M->reduced = R.release();
        return nullptr;
    }
    const uint32_t n = As->nrow, n0 = H.n0, n1 = H.n1;
    R->anorm = H.anorm;
    try {
        R->n0 = n0;
        R->n1 = n1;
        // KA and KB run every iteration, whose operands (the two matrices and
        // five vectors of the second colour) stay cache-resident between
        // launches; KF and KE run once per refinement step
        // KA rows (first colour j) reach second-colour columns near n0 + j, KB
        // rows (second colour q) first-colour columns near q and their own
        // diagonal at n0 + q
        if (ell_build(H.ka_rp, H.ka_ci, H.ka_v, n0, false, 0, R->ka))
            if (!ell_build(H.kb_rp, H.kb_ci, H.kb_v, 0, true, n0, R->kb)) R->ka.free_all();
        // the CSR forms are timed only when the iteration runs them
        const int tune = R->ka.ok ? 0 : 2;
        R->KA = csr_from_vectors(ctx, n0, n, std::move(H.ka_rp), std::move(H.ka_ci), std::move(H.ka_v), tune);
        R->KB = csr_from_vectors(ctx, n1, n, std::move(H.kb_rp), std::move(H.kb_ci), std::move(H.kb_v), tune);
        R->KF = csr_from_vectors(ctx, n1, n, std::move(H.kf_rp), std::move(H.kf_ci), std::move(H.kf_v), 0);
        R->KE = csr_from_vectors(ctx, n0, n, std::move(H.ke_rp), std::move(H.ke_ci), std::move(H.ke_v), 0);
        R->s1_64 = dalloc<double>(n1);
        PP_HIP(hipMemcpy(R->s1_64, H.s1.data(), n1 * sizeof(double), hipMemcpyHostToDevice));
        R->dt_64 = dalloc<double>(n1);
        PP_HIP(hipMemcpy(R->dt_64, H.dt.data(), n1 * sizeof(double), hipMemcpyHostToDevice));
        R->ok = true;
        R->built = 1;
        R->build_ms = ms();
    } catch (...) {
        M->reduced = R.release();
        reduced_free(M);
        throw;
    }
    M->reduced = R.release();
    return M->reduced;
}

void reduced_need_f32(pp_ctx *ctx, ReducedForm *R) {
    csr_need_f32(ctx, R->KA);
    csr_need_f32(ctx, R->KB);
    csr_need_f32(ctx, R->KF);
    csr_need_f32(ctx, R->KE);
    auto to32 = [&](const double *src, float *&dst) {  // rounded to nearest, as a host cast
        if (dst) return;
        dst = dalloc<float>(R->n1);
        launch_convert(ctx, R->n1, PP_F64, src, PP_F32, dst);
    };
    to32(R->s1_64, R->s1_32);
    to32(R->dt_64, R->dt_32);
    for (EllOp *E : {&R->ka, &R->kb}) {
        if (!E->ok || E->v32) continue;
        E->v32 = dalloc<float>((size_t)E->width * E->n);
        launch_convert(ctx, E->width * E->n, PP_F64, E->v64, PP_F32, E->v32);
        if (E->dg64) {
            E->dg32 = dalloc<float>(E->n);
            launch_convert(ctx, E->n, PP_F64, E->dg64, PP_F32, E->dg32);
        }
    }
    PP_HIP(hipStreamSynchronize(ctx->stream));
}

void reduced_free(pp_precond *M) {
    if (!M || !M->reduced) return;
    ReducedForm *R = M->reduced;
    csr_free(R->KA);
    csr_free(R->KB);
    csr_free(R->KF);
    csr_free(R->KE);
    dfree(R->s1_64);
    dfree(R->s1_32);
    dfree(R->dt_64);
    dfree(R->dt_32);
    R->ka.free_all();
    R->kb.free_all();
    dfree(R->dfail);
    dfree(R->dwork);
    delete R;
    M->reduced = nullptr;
}

}  // namespace pp
