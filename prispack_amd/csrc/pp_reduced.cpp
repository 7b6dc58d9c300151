// pp_reduced.cpp -- IC(0) of a two-colour matrix in Schur-complement form
// (the operators of the reduced CG; see ReducedForm in pp_internal.h).
// Host-side construction from the colour-ordered matrix, once per matrix
// version; the device work is SpMV (k_spmv) on the four operators.
#include <algorithm>
#include <cmath>
#include <memory>

#include "pp_internal.h"

namespace pp {
pp_csr *csr_new(pp_ctx *ctx, uint32_t nrow, uint32_t ncol, const uint32_t *rp, const uint32_t *ci, const double *v);
void csr_free(pp_csr *A);
void csr_need_f32(pp_ctx *ctx, pp_csr *A);

namespace {

struct Rows {
    std::vector<uint32_t> rp{0}, ci;
    std::vector<double> v;
    void put(uint32_t c, double x) {
        ci.push_back(c);
        v.push_back(x);
    }
    void end_row() { rp.push_back((uint32_t)ci.size()); }
};

// Colour split of a colour-ordered matrix: n0 = the first row with an
// entry left of its diagonal.  Valid when every first-colour row couples
// only to the second colour and vice versa, and every row has a diagonal.
bool split_two_colours(const pp_csr *A, uint32_t &n0) {
    const uint32_t n = A->nrow;
    const auto &rp = A->h_rp;
    const auto &ci = A->h_ci;
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
                diag = true;
                continue;
            }
            if ((i < n0) == (j < n0)) return false;  // coupling inside a colour
        }
        if (!diag) return false;
    }
    return true;
}

double diag_of(const pp_csr *A, uint32_t i) {
    for (uint32_t k = A->h_rp[i]; k < A->h_rp[i + 1]; ++k)
        if (A->h_ci[k] == i) return A->h_v[k];
    return 0.0;
}

}  // namespace

ReducedForm *reduced_form(pp_ctx *ctx, pp_precond *M, pp_csr *As) {
    if (!M || M->opts.kind != PP_PC_IC0 || M->order.empty() || !M->io_permuted || M->multilevel || !As) return nullptr;
    if (M->reduced && M->reduced->src == As && M->reduced->src_version == As->version)
        return M->reduced->ok ? M->reduced : nullptr;
    reduced_free(M);
    std::unique_ptr<ReducedForm> R(new ReducedForm);
    R->src = As;
    R->src_version = As->version;
    uint32_t n0 = 0;
    const uint32_t n = As->nrow;
    // the factor's first level must be exactly the first colour
    bool ok = As->nrow == As->ncol && split_two_colours(As, n0) && M->low64.nlev == 2 && M->up64.nlev == 2 &&
              M->lvl0_contiguous && M->lvl0_lo == 0 && M->lvl0_hi == n0;
    std::vector<double> d(n), dt;
    if (ok) {
        for (uint32_t i = 0; i < n && ok; ++i) {
            d[i] = diag_of(As, i);
            ok = d[i] > 0.0;
        }
    }
    if (ok) {
        // Dt_i = a_ii - sum_j a_ij^2 / a_jj over the first-colour neighbours
        // (the IC(0) pivot of a second-colour row)
        dt.assign(n - n0, 0.0);
        for (uint32_t i = n0; i < n && ok; ++i) {
            double t = d[i];
            for (uint32_t k = As->h_rp[i]; k < As->h_rp[i + 1]; ++k) {
                const uint32_t j = As->h_ci[k];
                if (j < n0) t -= As->h_v[k] * As->h_v[k] / d[j];
            }
            dt[i - n0] = t;
            ok = t > 0.0 && std::isfinite(t);
        }
    }
    if (!ok) {
        M->reduced = R.release();
        return nullptr;
    }
    const uint32_t n1 = n - n0;
    for (uint32_t i = 0; i < n; ++i) {
        double t = 0.0;
        for (uint32_t k = As->h_rp[i]; k < As->h_rp[i + 1]; ++k) t += std::fabs(As->h_v[k]);
        R->anorm = std::max(R->anorm, t);
    }
    Rows KA, KB, KF, KE;
    for (uint32_t i = 0; i < n0; ++i) {
        KE.put(i, 1.0 / d[i]);
        for (uint32_t k = As->h_rp[i]; k < As->h_rp[i + 1]; ++k) {
            const uint32_t j = As->h_ci[k];
            if (j == i) continue;
            const double a = As->h_v[k];
            KA.put(j, a / std::sqrt(d[i] * dt[j - n0]));
            KE.put(j, -a / d[i]);
        }
        KA.end_row();
        KE.end_row();
    }
    std::vector<double> s1(n1);
    for (uint32_t i = n0; i < n; ++i) {
        const double t = dt[i - n0], rs = 1.0 / std::sqrt(t);
        for (uint32_t k = As->h_rp[i]; k < As->h_rp[i + 1]; ++k) {
            const uint32_t j = As->h_ci[k];
            if (j == i) continue;
            const double a = As->h_v[k];
            KB.put(j, -a / std::sqrt(d[j] * t));
            KF.put(j, -a / d[j] * rs);
        }
        KB.put(i, d[i] / t);  // the diagonal column i > every first-colour column
        KF.put(i, rs);
        KB.end_row();
        KF.end_row();
        s1[i - n0] = rs;
    }
    try {
        R->n0 = n0;
        R->n1 = n1;
        R->KA = csr_new(ctx, n0, n, KA.rp.data(), KA.ci.data(), KA.v.data());
        R->KB = csr_new(ctx, n1, n, KB.rp.data(), KB.ci.data(), KB.v.data());
        R->KF = csr_new(ctx, n1, n, KF.rp.data(), KF.ci.data(), KF.v.data());
        R->KE = csr_new(ctx, n0, n, KE.rp.data(), KE.ci.data(), KE.v.data());
        R->s1_64 = dalloc<double>(n1);
        PP_HIP(hipMemcpy(R->s1_64, s1.data(), n1 * sizeof(double), hipMemcpyHostToDevice));
        R->dt_64 = dalloc<double>(n1);
        PP_HIP(hipMemcpy(R->dt_64, dt.data(), n1 * sizeof(double), hipMemcpyHostToDevice));
        R->ok = true;
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
    auto to32 = [&](const double *src, float *&dst) {
        if (dst) return;
        std::vector<double> h(R->n1);
        std::vector<float> f(R->n1);
        PP_HIP(hipMemcpy(h.data(), src, R->n1 * sizeof(double), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < R->n1; ++i) f[i] = (float)h[i];
        dst = dalloc<float>(R->n1);
        PP_HIP(hipMemcpy(dst, f.data(), R->n1 * sizeof(float), hipMemcpyHostToDevice));
    };
    to32(R->s1_64, R->s1_32);
    to32(R->dt_64, R->dt_32);
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
    delete R;
    M->reduced = nullptr;
}

}  // namespace pp
