// pp_factor.cpp -- host-side preconditioner setup (the factorization stays on
// the host, as in the reference; the device only applies the factor).
//
// The threshold factorizations must produce the reference's factor bit for bit
// (iteration-count parity depends on it), so they take the same steps in the
// same order: the kept pattern depends on every drop decision and the values
// on the order of the floating-point updates.  The zero-fill IC(0)/ILU(0)
// have no reference counterpart.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>

#include "pp_internal.h"

namespace pp {

static constexpr uint32_t kNone = 0xFFFFFFFFu;

// T = transpose of the n x ncol CSR (rp, ci, v), each row of T starting
// with `front` reserved slots.  Row j of T lists the entries of column j in
// the order of their positions in A (rows ascending).  Large matrices:
// column counts and placement by atomics over row ranges, then each row of
// T sorted by source position, which restores that order.
static void transpose_into(uint32_t n, uint32_t ncol, const uint32_t *rp, const uint32_t *ci, const double *v,
                           uint32_t front, HostCsr &T) {
    const uint32_t nnz = rp[n];
    T.n = ncol;
    T.rp.assign((size_t)ncol + 1, 0);
    T.ci.resize((size_t)nnz + (size_t)front * ncol);
    T.v.resize(T.ci.size());
    if (nnz < (1u << 20)) {
        for (uint32_t k = 0; k < nnz; ++k) T.rp[ci[k] + 1]++;
        for (uint32_t j = 0; j < ncol; ++j) T.rp[j + 1] += T.rp[j] + front;
        std::vector<uint32_t> next(ncol);
        for (uint32_t j = 0; j < ncol; ++j) next[j] = T.rp[j] + front;
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const uint32_t p = next[ci[k]]++;
                T.ci[p] = i;
                T.v[p] = v[k];
            }
        return;
    }
    uint32_t *cnt = T.rp.data() + 1;
    parallel_for(n, 1u << 14, [&](size_t i0, size_t i1) {
        for (uint32_t k = rp[i0]; k < rp[i1]; ++k) __atomic_fetch_add(&cnt[ci[k]], 1u, __ATOMIC_RELAXED);
    });
    for (uint32_t j = 0; j < ncol; ++j) T.rp[j + 1] += T.rp[j] + front;
    std::vector<uint32_t> next(ncol);
    for (uint32_t j = 0; j < ncol; ++j) next[j] = T.rp[j] + front;
    std::vector<uint64_t> key(T.ci.size());  // source position << 32 | source row
    parallel_for(n, 1u << 14, [&](size_t i0, size_t i1) {
        for (uint32_t i = (uint32_t)i0; i < (uint32_t)i1; ++i)
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
                const uint32_t p = __atomic_fetch_add(&next[ci[k]], 1u, __ATOMIC_RELAXED);
                key[p] = (uint64_t)k << 32 | i;
            }
    });
    parallel_for(ncol, 1u << 14, [&](size_t j0, size_t j1) {
        for (size_t j = j0; j < j1; ++j) {
            const uint32_t b = T.rp[j] + front, e = T.rp[j + 1];
            std::sort(key.begin() + b, key.begin() + e);
            for (uint32_t p = b; p < e; ++p) {
                T.ci[p] = (uint32_t)key[p];
                T.v[p] = v[key[p] >> 32];
            }
        }
    });
}

HostCsr transpose(const HostCsr &A) {
    HostCsr T;
    transpose_into(A.n, A.n, A.rp.data(), A.ci.data(), A.v.data(), 0, T);
    return T;
}

// Threshold IC / LDL^T, left-looking by column
// (IncompleteCholesky.cpp:96-246, IncompleteLDLT.cpp:40-186).
//
// Column j accumulates the lower part of row j of A and subtracts the
// contributions of every factored column k whose next unconsumed entry lies in
// row j.  Those columns sit in per-row singly-linked lists that share one
// array: `chain[r]` is the head of row r's list until row r is factored and
// column r's successor afterwards.  Entries with |value| < tau * sum_{i>=j}
// |a_ij| (measured before scaling) are dropped.  The result is L^T by rows:
// IC stores sqrt(pivot) first and divides by it; LDL^T stores no diagonal
// (unit), keeps the pivot in d and divides by it.
int factor_ict(const HostCsr &A, double tau, bool ldlt, HostCsr &lt, std::vector<double> &d, uint32_t &fail_row) {
    const uint32_t n = A.n;
    std::vector<uint8_t> active(n, 0);
    std::vector<uint32_t> rows;
    rows.reserve(64);
    std::vector<double> acc(n, 0.0);
    std::vector<uint32_t> chain(n, kNone), cursor(n, 0);
    lt.n = n;
    lt.rp.assign(n + 1, 0);
    lt.ci.clear();
    lt.v.clear();
    lt.ci.reserve((size_t)A.rp[n] * 2);
    lt.v.reserve((size_t)A.rp[n] * 2);
    if (ldlt) d.assign(n, 0.0);
    int status = PP_OK;

    for (uint32_t j = 0; j < n; ++j) {
        rows.clear();
        double colsum = 0.0;
        for (uint32_t k = A.rp[j]; k < A.rp[j + 1]; ++k) {
            const uint32_t i = A.ci[k];
            if (i < j) continue;
            colsum += std::fabs(A.v[k]);
            if (!active[i]) {
                active[i] = 1;
                acc[i] = 0.0;
                rows.push_back(i);
            }
            acc[i] += A.v[k];
        }
        const double thr = colsum * tau;

        for (uint32_t k = chain[j]; k != kNone;) {
            uint32_t pos = cursor[k];
            const double ljk = lt.v[pos];
            const double dk = ldlt ? d[k] : 1.0;
            const uint32_t end = lt.rp[k + 1];
            for (uint32_t q = pos; q < end; ++q) {
                const uint32_t i = lt.ci[q];
                if (!active[i]) {
                    active[i] = 1;
                    acc[i] = 0.0;
                    rows.push_back(i);
                }
                if (ldlt) acc[i] -= lt.v[q] * ljk * dk;
                else acc[i] -= lt.v[q] * ljk;
            }
            const uint32_t after = chain[k];
            if (++pos < end) {  // move column k to the list of its next row
                cursor[k] = pos;
                const uint32_t r = lt.ci[pos];
                chain[k] = chain[r];
                chain[r] = k;
            }
            k = after;
        }

        double piv = acc[j];
        if (!ldlt) {
            if (!(piv > 0.0)) {
                status = PP_ERR_NOT_SPD;
                fail_row = j;
                for (uint32_t r = j; r < n; ++r) lt.rp[r + 1] = (uint32_t)lt.ci.size();
                break;
            }
            piv = std::sqrt(piv);
            acc[j] = piv;
            lt.ci.push_back(j);
            lt.v.push_back(piv);
        } else {
            if (std::fabs(piv) < 1e-15) {
                status = PP_ERR_ZERO_PIVOT;
                fail_row = j;
                for (uint32_t r = j; r < n; ++r) lt.rp[r + 1] = (uint32_t)lt.ci.size();
                break;
            }
            d[j] = piv;
        }
        cursor[j] = (uint32_t)lt.ci.size();
        std::sort(rows.begin(), rows.end());
        bool linked = false;
        for (uint32_t i : rows) {
            active[i] = 0;
            if (i == j) continue;
            const double lij = acc[i];
            if (!(std::fabs(lij) >= thr)) continue;
            if (!linked) {  // column j enters the list of its first off-diagonal row
                linked = true;
                chain[j] = chain[i];
                chain[i] = j;
            }
            lt.ci.push_back(i);
            lt.v.push_back(lij / piv);
        }
        lt.rp[j + 1] = (uint32_t)lt.ci.size();
    }
    return status;
}

// Partial selection that brings the `keep` largest |a| to the front and fixes
// the stored order of an ILUT row; same pivoting steps as qsplit
// (include/VectorClass/VectorTools.h:159-201), whose output order is
// observable in the factor.
static void select_largest(double *a, int *ind, int n, int keep) {
    const int cut = keep - 1;
    int lo = 0, hi = n - 1;
    if (cut < lo || cut > hi) return;
    for (;;) {
        int mid = lo;
        const double key = std::fabs(a[mid]);
        for (int j = lo + 1; j <= hi; ++j)
            if (std::fabs(a[j]) > key) {
                ++mid;
                std::swap(a[mid], a[j]);
                std::swap(ind[mid], ind[j]);
            }
        std::swap(a[mid], a[lo]);
        std::swap(ind[mid], ind[lo]);
        if (mid == cut) return;
        if (mid > cut) hi = mid - 1;
        else lo = mid + 1;
    }
}

// ILUT(lfil, tau), row-wise IKJ elimination (IncompleteLU.cpp:38-266).
// Work row layout: lower entries at [0, nl), diagonal at i, upper entries at
// i+1 .. i+nu; where[col] maps a column to its slot (-1 when absent).
int factor_ilut(const HostCsr &A, uint32_t lfil, double tau, HostCsr &L, HostCsr &U, std::vector<double> &D,
                uint32_t &fail_row) {
    const int n = (int)A.n;
    std::vector<int> where(n, -1), col(n, 0), sel(n, 0);
    std::vector<double> w(n, 0.0), mag(n, 0.0);
    D.assign(n, 0.0);
    L.n = U.n = (uint32_t)n;
    L.rp.assign(n + 1, 0);
    U.rp.assign(n + 1, 0);
    L.ci.clear(); L.v.clear(); U.ci.clear(); U.v.clear();
    int status = PP_OK;
    for (int i = 0; i < n; ++i) {
        const uint32_t b = A.rp[i], e = A.rp[i + 1];
        double rownorm = 0.0;
        for (uint32_t k = b; k < e; ++k) rownorm += std::fabs(A.v[k]);
        if (rownorm < 1e-15) {
            status = PP_ERR_ZERO_ROW;
            fail_row = (uint32_t)i;
            break;
        }
        rownorm /= (double)(e - b);
        const double droptol = tau * rownorm;
        int nl = 0, nu = 0;
        col[i] = i;
        w[i] = 0.0;
        where[i] = i;
        for (uint32_t k = b; k < e; ++k) {
            const int c = (int)A.ci[k];
            if (c < i) {
                where[c] = nl; col[nl] = c; w[nl] = A.v[k]; ++nl;
            } else if (c == i) {
                w[i] = A.v[k];
            } else {
                ++nu;
                where[c] = i + nu; col[i + nu] = c; w[i + nu] = A.v[k];
            }
        }
        for (int j = 0; j < nl; ++j) {
            // bring the smallest remaining column to slot j
            int piv_row = col[j], piv_slot = j;
            for (int k = j + 1; k < nl; ++k)
                if (col[k] < piv_row) { piv_row = col[k]; piv_slot = k; }
            if (piv_slot != j) {
                const int c = col[j];
                col[j] = col[piv_slot]; col[piv_slot] = c;
                where[piv_row] = j; where[c] = piv_slot;
                std::swap(w[j], w[piv_slot]);
            }
            const double mult = w[j] * D[piv_row];
            w[j] = mult;
            where[piv_row] = -1;
            for (uint32_t k = U.rp[piv_row]; k < U.rp[piv_row + 1]; ++k) {
                const int c = (int)U.ci[k];
                const int slot = where[c];
                const double upd = -mult * U.v[k];
                if (slot == -1 && std::fabs(upd) < droptol) continue;
                if (c < i) {
                    if (slot == -1) { col[nl] = c; where[c] = nl; w[nl] = upd; ++nl; }
                    else w[slot] += upd;
                } else if (slot != -1) {
                    w[slot] += upd;
                } else if (std::fabs(upd) > droptol) {
                    ++nu;
                    col[i + nu] = c; where[c] = i + nu; w[i + nu] = upd;
                }
                // |upd| == droptol for a new upper entry: dropped (undefined in the reference)
            }
        }
        where[i] = -1;
        for (int j = 0; j < nu; ++j) where[col[i + j + 1]] = -1;
        if (std::fabs(w[i]) < 1e-15) {
            status = PP_ERR_ZERO_PIVOT;
            fail_row = (uint32_t)i;
            break;
        }
        D[i] = 1.0 / w[i];
        int keep = std::min(nl, (int)lfil);
        for (int j = 0; j < nl; ++j) { mag[j] = std::fabs(w[j]); sel[j] = j; }
        select_largest(mag.data(), sel.data(), nl, keep);
        for (int j = 0; j < keep; ++j) { L.ci.push_back((uint32_t)col[sel[j]]); L.v.push_back(w[sel[j]]); }
        L.rp[i + 1] = (uint32_t)L.ci.size();
        keep = std::min(nu, (int)lfil);
        for (int j = 0; j < nu; ++j) { mag[j] = std::fabs(w[i + j + 1]); sel[j] = i + j + 1; }
        select_largest(mag.data(), sel.data(), nu, keep);
        for (int j = 0; j < keep; ++j) { U.ci.push_back((uint32_t)col[sel[j]]); U.v.push_back(w[sel[j]]); }
        U.rp[i + 1] = (uint32_t)U.ci.size();
    }
    if (status != PP_OK)
        for (int r = (int)fail_row; r < n; ++r) {
            L.rp[r + 1] = (uint32_t)L.ci.size();
            U.rp[r + 1] = (uint32_t)U.ci.size();
        }
    return status;
}

// IC(0): zero-fill incomplete Cholesky on the pattern of tril(A), row by row:
//   L_ik = (a_ik - sum_{j<k in rows i and k} L_ij L_kj) / L_kk,  L_ii = sqrt(a_ii - sum_k L_ik^2).
// Returned as L^T with the diagonal first (the threshold-IC layout).
int factor_ic0(const HostCsr &A, HostCsr &lt, uint32_t &fail_row, HostSweep *lower) {
    const uint32_t n = A.n;
    std::vector<uint32_t> rp(n + 1, 0);
    parallel_for(n, 1u << 15, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) {
            uint32_t c = 0;
            for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) c += A.ci[k] < i;
            rp[i + 1] = c;
        }
    });
    for (uint32_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
    std::vector<uint32_t> ci(rp[n]);
    std::vector<double> lv(rp[n]), diag(n), aii(n, 0.0);
    // the strict lower pattern (in the order of A's rows)
    parallel_for(n, 1u << 15, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1; ++i) {
            uint32_t q = rp[i];
            for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const uint32_t c = A.ci[k];
                if (c < i) {
                    ci[q] = c;
                    lv[q] = A.v[k];
                    ++q;
                } else if (c == i) {
                    aii[i] = A.v[k];
                }
            }
        }
    });
    // the dependency level of every row: rows of one level factor independently
    std::vector<uint32_t> lev(n, 0);
    uint32_t nlev = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t l = 0;
        for (uint32_t q = rp[i]; q < rp[i + 1]; ++q) l = std::max(l, lev[ci[q]] + 1);
        lev[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    std::vector<uint32_t> lptr(nlev + 1, 0), rows(n);
    for (uint32_t i = 0; i < n; ++i) lptr[lev[i] + 1]++;
    for (uint32_t l = 0; l < nlev; ++l) lptr[l + 1] += lptr[l];
    {
        std::vector<uint32_t> f(lptr.begin(), lptr.end() - 1);
        for (uint32_t i = 0; i < n; ++i) rows[f[lev[i]]++] = i;
    }
    // row i: l_ij = (a_ij - sum_{k<j} l_ik l_jk) / l_jj with the sum taken
    // in the order of row j's stored entries, l_ii = sqrt(a_ii - sum l_ij^2)
    std::atomic<int64_t> first_fail{-1};
    auto pivot = [&](uint32_t i) {
        double s = aii[i];
        for (uint32_t t = rp[i]; t < rp[i + 1]; ++t) s -= lv[t] * lv[t];
        if (!(s > 0.0)) {
            int64_t expect = first_fail.load();
            while ((expect < 0 || (int64_t)i < expect) && !first_fail.compare_exchange_weak(expect, i)) {
            }
            s = 1.0;
        }
        diag[i] = std::sqrt(s);
    };
    bool ascending = true;
    for (uint32_t i = 0; i < n && ascending; ++i)
        for (uint32_t t = rp[i]; t + 1 < rp[i + 1]; ++t)
            if (!(ci[t] < ci[t + 1])) {
                ascending = false;
                break;
            }
    if (ascending) {
        // rows of one level in parallel; row i's entries before t meet row
        // k's entries in a merge of the two ascending column lists
        for (uint32_t l = 0; l < nlev; ++l) {
            const uint32_t r0 = lptr[l], r1 = lptr[l + 1];
            parallel_for(r1 - r0, 4096, [&](size_t a, size_t b) {
                for (size_t q = a; q < b; ++q) {
                    const uint32_t i = rows[r0 + q];
                    const uint32_t i0 = rp[i], i1 = rp[i + 1];
                    for (uint32_t t = i0; t < i1; ++t) {
                        const uint32_t k = ci[t];
                        double s = lv[t];
                        uint32_t p = i0;
                        for (uint32_t u = rp[k]; u < rp[k + 1]; ++u) {
                            const uint32_t c = ci[u];
                            while (p < t && ci[p] < c) ++p;
                            if (p < t && ci[p] == c) s -= lv[p] * lv[u];
                        }
                        lv[t] = s / diag[k];
                    }
                    pivot(i);
                }
            });
        }
    } else {
        // unsorted or repeated columns: serial, row i's entries found by
        // column through a slot table (the last of repeated columns)
        std::vector<int64_t> slot(n, -1);
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t t = rp[i]; t < rp[i + 1]; ++t) slot[ci[t]] = t;
            for (uint32_t t = rp[i]; t < rp[i + 1]; ++t) {
                const uint32_t k = ci[t];
                double s = lv[t];
                for (uint32_t u = rp[k]; u < rp[k + 1]; ++u) {
                    const int64_t p = slot[ci[u]];
                    if (p >= 0) s -= lv[p] * lv[u];
                }
                lv[t] = s / diag[k];
            }
            pivot(i);
            for (uint32_t t = rp[i]; t < rp[i + 1]; ++t) slot[ci[t]] = -1;
        }
    }
    int status = PP_OK;
    if (first_fail.load() >= 0) {
        status = PP_ERR_NOT_SPD;
        fail_row = (uint32_t)first_fail.load();
    }
    // L^T by rows: sqrt(pivot) first, then column j of the strict lower part
    // (rows ascending).  With sorted rows and a symmetric pattern, column j
    // of L is the upper part of A's row j and each value sits in row i of L
    // at column j: found by binary search, row-parallel, no scatter.
    bool sym = ascending;
    if (sym) {
        lt.n = n;
        lt.rp.assign(n + 1, 0);
        std::atomic<bool> ok{true};
        parallel_for(n, 1u << 15, [&](size_t i0, size_t i1) {
            for (size_t i = i0; i < i1; ++i) {
                uint32_t c = 1, prev = (uint32_t)i;
                for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
                    if (A.ci[k] > i) {
                        if (A.ci[k] <= prev) ok = false;  // upper part not strictly ascending
                        prev = A.ci[k];
                        ++c;
                    }
                lt.rp[i + 1] = c;
            }
        });
        for (uint32_t i = 0; i < n; ++i) lt.rp[i + 1] += lt.rp[i];
        sym = ok && lt.rp[n] == rp[n] + n;
        if (sym) {
            lt.ci.resize(lt.rp[n]);
            lt.v.resize(lt.rp[n]);
            parallel_for(n, 1u << 15, [&](size_t j0, size_t j1) {
                for (uint32_t j = (uint32_t)j0; j < (uint32_t)j1; ++j) {
                    uint32_t q = lt.rp[j];
                    lt.ci[q] = j;
                    lt.v[q++] = diag[j];
                    for (uint32_t k = A.rp[j]; k < A.rp[j + 1]; ++k) {
                        const uint32_t i = A.ci[k];
                        if (i <= j) continue;
                        const auto b = ci.begin() + rp[i], e = ci.begin() + rp[i + 1];
                        const auto it = std::lower_bound(b, e, j);
                        if (it == e || *it != j) {
                            ok = false;
                            return;
                        }
                        lt.ci[q] = i;
                        lt.v[q++] = lv[it - ci.begin()];
                    }
                }
            });
            sym = ok;
        }
    }
    if (!sym) {
        transpose_into(n, n, rp.data(), ci.data(), lv.data(), 1, lt);
        parallel_for(n, 1u << 16, [&](size_t j0, size_t j1) {
            for (size_t j = j0; j < j1; ++j) {
                lt.ci[lt.rp[j]] = (uint32_t)j;
                lt.v[lt.rp[j]] = diag[j];
            }
        });
    }
    // the forward sweep directly from the rows of L (sorted rows: the same
    // rows the transpose of L^T gives)
    if (lower) {
        *lower = HostSweep();
        if (ascending) {
            lower->n = n;
            lower->lower = true;
            lower->mode = HostSweep::DIVIDE;
            lower->rp = std::move(rp);
            lower->ci = std::move(ci);
            lower->val = std::move(lv);
            lower->diag = std::move(diag);
        }
    }
    return status;
}

// ILU(0): zero-fill ILU in IKJ form on the pattern of A; L strict lower (unit),
// U strict upper, D = 1/u_ii.
int factor_ilu0(const HostCsr &A, HostCsr &L, HostCsr &U, std::vector<double> &D, uint32_t &fail_row) {
    const uint32_t n = A.n;
    std::vector<double> w(A.v);
    std::vector<uint32_t> dpos(n, kNone);
    std::vector<int64_t> slot(n, -1);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] == i) dpos[i] = k;
    int status = PP_OK;
    for (uint32_t i = 0; i < n && status == PP_OK; ++i) {
        if (dpos[i] == kNone) { status = PP_ERR_ZERO_PIVOT; fail_row = i; break; }
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) slot[A.ci[k]] = k;
        for (uint32_t t = A.rp[i]; t < A.rp[i + 1]; ++t) {
            const uint32_t k = A.ci[t];
            if (k >= i) break;
            w[t] /= w[dpos[k]];
            for (uint32_t u = dpos[k] + 1; u < A.rp[k + 1]; ++u) {
                const int64_t p = slot[A.ci[u]];
                if (p >= 0) w[p] -= w[t] * w[u];
            }
        }
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) slot[A.ci[k]] = -1;
        if (std::fabs(w[dpos[i]]) < 1e-15) { status = PP_ERR_ZERO_PIVOT; fail_row = i; }
    }
    L.n = U.n = n;
    L.rp.assign(n + 1, 0);
    U.rp.assign(n + 1, 0);
    L.ci.clear(); L.v.clear(); U.ci.clear(); U.v.clear();
    D.assign(n, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const uint32_t c = A.ci[k];
            if (c < i) { L.ci.push_back(c); L.v.push_back(w[k]); }
            else if (c > i) { U.ci.push_back(c); U.v.push_back(w[k]); }
        }
        L.rp[i + 1] = (uint32_t)L.ci.size();
        U.rp[i + 1] = (uint32_t)U.ci.size();
        D[i] = dpos[i] != kNone ? 1.0 / w[dpos[i]] : 0.0;
    }
    return status;
}

}  // namespace pp

namespace pp {

// Greedy distance-1 colouring of the symmetrised pattern in natural order;
// rows sorted by (colour, index).  Every dependency of a zero-fill factor row
// then lies in an earlier colour, so the sweeps have <= ncolors levels.
// Colours in one pass while they fit a 64-bit mask: row i takes the
// smallest colour not in forbid[i] (the colours of earlier rows with an
// entry in column i) nor among its own earlier columns, then forbids it to
// its later columns.  Exactly the greedy colouring below; false when a
// 65th colour would be needed.
static bool greedy_colours_masked(const HostCsr &A, std::vector<uint32_t> &color, uint32_t &ncolors) {
    const uint32_t n = A.n;
    std::vector<uint64_t> forbid(n, 0);
    color.assign(n, 0);
    ncolors = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t m = forbid[i];
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] < i) m |= 1ull << color[A.ci[k]];
        if (m == ~0ull) return false;
        const uint32_t c = (uint32_t)__builtin_ctzll(~m);
        color[i] = c;
        ncolors = std::max(ncolors, c + 1);
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] > i) forbid[A.ci[k]] |= 1ull << c;
    }
    return true;
}

std::vector<uint32_t> multicolor_order(const HostCsr &A, uint32_t &ncolors) {
    const uint32_t n = A.n;
    std::vector<uint32_t> color;
    if (greedy_colours_masked(A, color, ncolors)) {
        std::vector<uint32_t> order(n), start(ncolors + 1, 0);
        for (uint32_t i = 0; i < n; ++i) start[color[i] + 1]++;
        for (uint32_t c = 0; c < ncolors; ++c) start[c + 1] += start[c];
        for (uint32_t i = 0; i < n; ++i) order[start[color[i]]++] = i;
        return order;
    }
    // symmetrised adjacency without the diagonal, as CSR (duplicates harmless)
    std::vector<uint32_t> deg(n + 1, 0);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] != i) {
                deg[i + 1]++;
                deg[A.ci[k] + 1]++;
            }
    for (uint32_t i = 0; i < n; ++i) deg[i + 1] += deg[i];
    std::vector<uint32_t> adj(deg[n]), fill(deg.begin(), deg.end() - 1);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] != i) {
                adj[fill[i]++] = A.ci[k];
                adj[fill[A.ci[k]]++] = i;
            }
    // greedy: each row takes the smallest colour none of its coloured
    // neighbours has
    color.assign(n, kNone);
    std::vector<uint32_t> mark;
    ncolors = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (mark.size() < ncolors + 1) mark.resize(ncolors + 1, kNone);
        for (uint32_t q = deg[i]; q < deg[i + 1]; ++q) {
            const uint32_t j = adj[q];
            if (color[j] != kNone) mark[color[j]] = i;
        }
        uint32_t c = 0;
        while (c < ncolors && mark[c] == i) ++c;
        color[i] = c;
        ncolors = std::max(ncolors, c + 1);
    }
    std::vector<uint32_t> order(n);
    std::vector<uint32_t> start(ncolors + 1, 0);
    for (uint32_t i = 0; i < n; ++i) start[color[i] + 1]++;
    for (uint32_t c = 0; c < ncolors; ++c) start[c + 1] += start[c];
    for (uint32_t i = 0; i < n; ++i) order[start[color[i]]++] = i;
    return order;
}

// P A P^T: row k is original row order[k]; columns renumbered and sorted.
HostCsr permute_symmetric(const HostCsr &A, const std::vector<uint32_t> &order) {
    const uint32_t n = A.n;
    std::vector<uint32_t> inv(n);
    for (uint32_t k = 0; k < n; ++k) inv[order[k]] = k;
    HostCsr B;
    B.n = n;
    B.rp.assign(n + 1, 0);
    for (uint32_t k = 0; k < n; ++k) B.rp[k + 1] = B.rp[k] + (A.rp[order[k] + 1] - A.rp[order[k]]);
    B.ci.resize(B.rp[n]);
    B.v.resize(B.rp[n]);
    parallel_for(n, 1u << 14, [&](size_t k0, size_t k1) {
        std::vector<std::pair<uint32_t, double>> row;
        for (size_t k = k0; k < k1; ++k) {
            const uint32_t i = order[k];
            row.clear();
            for (uint32_t q = A.rp[i]; q < A.rp[i + 1]; ++q) row.emplace_back(inv[A.ci[q]], A.v[q]);
            std::sort(row.begin(), row.end());
            uint32_t p = B.rp[k];
            for (auto &e : row) {
                B.ci[p] = e.first;
                B.v[p] = e.second;
                ++p;
            }
        }
    });
    return B;
}

}  // namespace pp
