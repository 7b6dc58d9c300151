/*
 * ppo.c -- the oracle: plain-C, single-threaded CPU restatement of the
 * PrisPack preconditioned-Krylov hot path (see ppo.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker / the CPU baseline.  The shipped
 * library never links or calls it.
 *
 * Parity pins (checked in tests/test_oracle.py):
 *   - spmv, ict, ilut, ildlt, the three applies, cg and gmres against golden
 *     vectors written by the compiled reference (oracle/_ref, built from
 *     /root/reference sources by oracle/Makefile) and committed under
 *     tests/golden/;
 *   - spmv against the closed-form Laplacian eigenvectors;
 *   - ic0 / ilu0 (no reference counterpart) by their defining invariant
 *     (L L^T)_ij = a_ij, (L U)_ij = a_ij on pattern(A), and by equality with
 *     ict(tau=0) / ilut(lfil=n, tau=0) on matrices whose exact factor has no
 *     fill (tridiagonal);
 *   - ict(tau=0) = complete Cholesky (dense numpy Cholesky on small n);
 *   - cg with M = A converging in one step; gmres with m >= n and an exact
 *     factor converging in one step.
 */
#include "ppo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NIL UINT32_MAX

/* ------------------------------------------------------------------ */
/* CSR storage                                                          */
/* ------------------------------------------------------------------ */
ppo_csr *ppo_csr_new(uint32_t nrow, uint32_t ncol, uint32_t nnz) {
    ppo_csr *a = (ppo_csr *)calloc(1, sizeof(ppo_csr));
    a->nrow = nrow;
    a->ncol = ncol;
    a->nnz = nnz;
    a->rp = (uint32_t *)calloc((size_t)nrow + 1, sizeof(uint32_t));
    a->ci = (uint32_t *)calloc(nnz ? nnz : 1, sizeof(uint32_t));
    a->v = (double *)calloc(nnz ? nnz : 1, sizeof(double));
    return a;
}

void ppo_csr_free(ppo_csr *a) {
    if (!a) return;
    free(a->rp);
    free(a->ci);
    free(a->v);
    free(a);
}

ppo_csr *ppo_csr_from(uint32_t nrow, uint32_t ncol, const uint32_t *rp,
                      const uint32_t *ci, const double *v) {
    uint32_t nnz = rp[nrow];
    ppo_csr *a = ppo_csr_new(nrow, ncol, nnz);
    memcpy(a->rp, rp, ((size_t)nrow + 1) * sizeof(uint32_t));
    memcpy(a->ci, ci, (size_t)nnz * sizeof(uint32_t));
    memcpy(a->v, v, (size_t)nnz * sizeof(double));
    return a;
}

/* Explicit transpose (columns become rows, entries sorted by column index
 * because rows are visited in ascending order). */
ppo_csr *ppo_csr_transpose(const ppo_csr *a) {
    ppo_csr *t = ppo_csr_new(a->ncol, a->nrow, a->nnz);
    for (uint32_t k = 0; k < a->nnz; ++k) t->rp[a->ci[k] + 1]++;
    for (uint32_t i = 0; i < a->ncol; ++i) t->rp[i + 1] += t->rp[i];
    uint32_t *fill = (uint32_t *)malloc(((size_t)a->ncol + 1) * sizeof(uint32_t));
    memcpy(fill, t->rp, ((size_t)a->ncol + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < a->nrow; ++i)
        for (uint32_t k = a->rp[i]; k < a->rp[i + 1]; ++k) {
            uint32_t p = fill[a->ci[k]]++;
            t->ci[p] = i;
            t->v[p] = a->v[k];
        }
    free(fill);
    return t;
}

/* ------------------------------------------------------------------ */
/* Seeded inputs: SplitMix64 (Steele, Lea, Flood 2014)                  */
/* ------------------------------------------------------------------ */
uint64_t ppo_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* x_i = 2 u - 1 with u = (z >> 11) / 2^53 in [0, 1). */
void ppo_fill_uniform(double *x, uint64_t n, uint64_t seed) {
    uint64_t s = seed;
    const double inv53 = 1.0 / 9007199254740992.0;
    for (uint64_t i = 0; i < n; ++i) {
        double u = (double)(ppo_splitmix64(&s) >> 11) * inv53;
        x[i] = 2.0 * u - 1.0;
    }
}

/* Box-Muller on pairs: u1 in (0, 1], u2 in [0, 1);
 * x_2k = r cos(2 pi u2), x_2k+1 = r sin(2 pi u2), r = sqrt(-2 ln u1). */
void ppo_fill_normal(double *x, uint64_t n, uint64_t seed) {
    uint64_t s = seed;
    const double inv53 = 1.0 / 9007199254740992.0;
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t i = 0; i < n; i += 2) {
        double u1 = (double)((ppo_splitmix64(&s) >> 11) + 1) * inv53;
        double u2 = (double)(ppo_splitmix64(&s) >> 11) * inv53;
        double r = sqrt(-2.0 * log(u1));
        double th = two_pi * u2;
        x[i] = r * cos(th);
        if (i + 1 < n) x[i + 1] = r * sin(th);
    }
}

/* ------------------------------------------------------------------ */
/* Model problems                                                       */
/* ------------------------------------------------------------------ */

/* Restates generateLaplacianCOO + transCOO2CSR
 * (source/MatrixClass/ModelProblem.cpp:68-133, include/MatrixClass/MatrixTools.h:16-62):
 * lexicographic order with x fastest; centred convection alpha; the COO rows
 * are insertion-sorted by column, so each CSR row ends up in ascending column
 * order.  2D when any dimension is 1 (diagonal 4), else 3D (diagonal 6). */
ppo_csr *ppo_laplacian(int nx, int ny, int nz, double ax, double ay, double az,
                       double shift) {
    int n = nx * ny * nz;
    int ones = (nx == 1) + (ny == 1) + (nz == 1);
    int is2d = ones > 0;
    double v = -1.0;
    double vxp = v - ax, vxn = v + ax, vyp = v - ay, vyn = v + ay, vzp = v - az, vzn = v + az;
    double vd = (is2d ? 4.0 : 6.0) - shift;
    uint32_t cap = (uint32_t)n * (is2d ? 5u : 7u);
    ppo_csr *a = ppo_csr_new((uint32_t)n, (uint32_t)n, cap);
    uint32_t nnz = 0;
    int nxy = nx * ny;
    for (int p = 0; p < n; ++p) {
        int k = p / nxy;
        int i = (p - k * nxy) / nx;
        int j = p - k * nxy - i * nx;
        /* ascending column order: z-, y-, x-, diag, x+, y+, z+ */
        if (k > 0) { a->ci[nnz] = (uint32_t)(p - nxy); a->v[nnz++] = vzn; }
        if (i > 0) { a->ci[nnz] = (uint32_t)(p - nx); a->v[nnz++] = vyn; }
        if (j > 0) { a->ci[nnz] = (uint32_t)(p - 1); a->v[nnz++] = vxn; }
        a->ci[nnz] = (uint32_t)p; a->v[nnz++] = vd;
        if (j < nx - 1) { a->ci[nnz] = (uint32_t)(p + 1); a->v[nnz++] = vxp; }
        if (i < ny - 1) { a->ci[nnz] = (uint32_t)(p + nx); a->v[nnz++] = vyp; }
        if (k < nz - 1) { a->ci[nnz] = (uint32_t)(p + nxy); a->v[nnz++] = vzp; }
        a->rp[p + 1] = nnz;
    }
    a->nnz = nnz;
    return a;
}

/* First-order upwind convection-diffusion, scaled by h^2 (no reference
 * counterpart: the reference generator is centred).  h = 1/(nx+1).
 * Diagonal 2*dim + h*sum(beta); upstream neighbour -1 - h*beta_k; downstream -1. */
ppo_csr *ppo_upwind(int nx, int ny, int nz, double bx, double by, double bz) {
    int n = nx * ny * nz;
    int is3d = nz > 1;
    double h = 1.0 / (double)(nx + 1);
    double cx = -1.0 - h * bx, cy = -1.0 - h * by, cz = -1.0 - h * bz;
    double vd = is3d ? 6.0 + h * (bx + by + bz) : 4.0 + h * (bx + by);
    ppo_csr *a = ppo_csr_new((uint32_t)n, (uint32_t)n, (uint32_t)n * 7u);
    uint32_t nnz = 0;
    int nxy = nx * ny;
    for (int p = 0; p < n; ++p) {
        int k = p / nxy;
        int i = (p - k * nxy) / nx;
        int j = p - k * nxy - i * nx;
        if (k > 0) { a->ci[nnz] = (uint32_t)(p - nxy); a->v[nnz++] = cz; }
        if (i > 0) { a->ci[nnz] = (uint32_t)(p - nx); a->v[nnz++] = cy; }
        if (j > 0) { a->ci[nnz] = (uint32_t)(p - 1); a->v[nnz++] = cx; }
        a->ci[nnz] = (uint32_t)p; a->v[nnz++] = vd;
        if (j < nx - 1) { a->ci[nnz] = (uint32_t)(p + 1); a->v[nnz++] = -1.0; }
        if (i < ny - 1) { a->ci[nnz] = (uint32_t)(p + nx); a->v[nnz++] = -1.0; }
        if (k < nz - 1) { a->ci[nnz] = (uint32_t)(p + nxy); a->v[nnz++] = -1.0; }
        a->rp[p + 1] = nnz;
    }
    a->nnz = nnz;
    return a;
}

/* 2D variable-coefficient diffusion -div(kappa grad u) on a side x side grid,
 * Dirichlet.  kappa is given per grid node; an interior face takes the
 * harmonic mean 2 k_p k_q / (k_p + k_q); a boundary face takes k_p. */
ppo_csr *ppo_vardiff2d(int side, const double *kappa) {
    int n = side * side;
    ppo_csr *a = ppo_csr_new((uint32_t)n, (uint32_t)n, (uint32_t)n * 5u);
    uint32_t nnz = 0;
    for (int p = 0; p < n; ++p) {
        int j = p / side, i = p - j * side;
        double kp = kappa[p];
        double fs = j > 0 ? 2.0 * kp * kappa[p - side] / (kp + kappa[p - side]) : kp;
        double fw = i > 0 ? 2.0 * kp * kappa[p - 1] / (kp + kappa[p - 1]) : kp;
        double fe = i < side - 1 ? 2.0 * kp * kappa[p + 1] / (kp + kappa[p + 1]) : kp;
        double fn = j < side - 1 ? 2.0 * kp * kappa[p + side] / (kp + kappa[p + side]) : kp;
        if (j > 0) { a->ci[nnz] = (uint32_t)(p - side); a->v[nnz++] = -fs; }
        if (i > 0) { a->ci[nnz] = (uint32_t)(p - 1); a->v[nnz++] = -fw; }
        a->ci[nnz] = (uint32_t)p; a->v[nnz++] = ((fs + fw) + fe) + fn;
        if (i < side - 1) { a->ci[nnz] = (uint32_t)(p + 1); a->v[nnz++] = -fe; }
        if (j < side - 1) { a->ci[nnz] = (uint32_t)(p + side); a->v[nnz++] = -fn; }
        a->rp[p + 1] = nnz;
    }
    a->nnz = nnz;
    return a;
}

/* ------------------------------------------------------------------ */
/* BLAS-1 and SpMV                                                      */
/* ------------------------------------------------------------------ */

/* CSRMatrix::MatPVecWithRawPtr (include/MatrixClass/CSRMatrix.h:360-377):
 * y_i = 0; y_i += x[col] * a, in storage order. */
void ppo_spmv(const ppo_csr *a, const double *x, double *y) {
    for (uint32_t i = 0; i < a->nrow; ++i) {
        double acc = 0.0;
        for (uint32_t k = a->rp[i]; k < a->rp[i + 1]; ++k) acc += x[a->ci[k]] * a->v[k];
        y[i] = acc;
    }
}

/* CSRMatrix::transMatPVec (include/MatrixClass/CSRMatrix.h:453-469): y += A^T x. */
void ppo_spmv_t_acc(const ppo_csr *a, const double *x, double *y) {
    for (uint32_t i = 0; i < a->nrow; ++i)
        for (uint32_t k = a->rp[i]; k < a->rp[i + 1]; ++k) y[a->ci[k]] += a->v[k] * x[i];
}

/* HostVector::innerProduct serial branch (include/VectorClass/BaseVector.h:429-450). */
double ppo_dot(uint32_t n, const double *x, const double *y) {
    double s = 0.0;
    for (uint32_t i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}

/* HostVector::norm_2 serial branch (include/VectorClass/BaseVector.h:508-530). */
double ppo_norm2(uint32_t n, const double *x) {
    double s = 0.0;
    for (uint32_t i = 0; i < n; ++i) s += x[i] * x[i];
    return sqrt(s);
}

/* HostVector::sumKahan(start, len, func) with func = x*x
 * (include/VectorClass/BaseVector.h:635-648). */
static double kahan_sumsq(uint32_t n, const double *x) {
    double sum = 0.0, corr = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double t = x[i] * x[i] + corr;
        double ns = sum + t;
        corr = t - (ns - sum);
        sum = ns;
    }
    return sum + corr;
}

/* ------------------------------------------------------------------ */
/* Preconditioners                                                      */
/* ------------------------------------------------------------------ */
void ppo_prec_free(ppo_prec *p) {
    if (!p) return;
    ppo_csr_free(p->lt);
    ppo_csr_free(p->l);
    ppo_csr_free(p->u);
    free(p->d);
    free(p);
}

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : (x > y);
}

/* Growable CSR row-by-row writer. */
typedef struct {
    uint32_t cap, nnz;
    uint32_t *ci;
    double *v;
} rowbuf;

static void rb_push(rowbuf *b, uint32_t c, double v) {
    if (b->nnz == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->ci = (uint32_t *)realloc(b->ci, (size_t)b->cap * sizeof(uint32_t));
        b->v = (double *)realloc(b->v, (size_t)b->cap * sizeof(double));
    }
    b->ci[b->nnz] = c;
    b->v[b->nnz++] = v;
}

static ppo_csr *rb_finish(rowbuf *b, uint32_t n, uint32_t *rp) {
    ppo_csr *a = (ppo_csr *)calloc(1, sizeof(ppo_csr));
    a->nrow = a->ncol = n;
    a->nnz = b->nnz;
    a->rp = rp;
    a->ci = b->ci ? b->ci : (uint32_t *)calloc(1, sizeof(uint32_t));
    a->v = b->v ? b->v : (double *)calloc(1, sizeof(double));
    return a;
}

/*
 * Threshold incomplete Cholesky / LDL^T, left-looking by column.
 * Restates IncompleteCholesky::setup (source/Preconditioner/IncompleteCholesky.cpp:96-246)
 * and IncompleteLDLT::setup (source/Preconditioner/IncompleteLDLT.cpp:40-186).
 *
 * Column j of L is built in a dense scatter array from the lower part of row
 * j of A (symmetric A), minus L_ik * L_jk (* d_k) for every earlier column k
 * whose next unconsumed entry lies in row j.  Those columns are kept in
 * singly-linked lists, one per row, sharing ONE link array: link[r] is the
 * head of row r's list while r is not yet factored and becomes column r's
 * "next" pointer afterwards (IncompleteCholesky.cpp:158-185, 217-218).  The
 * traversal order fixes the floating-point summation order, so it is kept.
 * An entry is dropped when |value before scaling| < tau * sum_{i>=j} |a_ij|
 * (:139-156, :213).  Output: row j of L^T = column j of L, ascending rows.
 *   ldlt = 0: diagonal sqrt(pivot) stored first; entries divided by it.
 *   ldlt = 1: no diagonal stored (unit); d_j = pivot kept in d[]; entries
 *             divided by the pivot.
 */
static ppo_prec *threshold_cholesky(const ppo_csr *a, double tau, int ldlt) {
    uint32_t n = a->ncol;
    ppo_prec *p = (ppo_prec *)calloc(1, sizeof(ppo_prec));
    p->kind = ldlt ? PPO_PREC_ILDLT : PPO_PREC_IC;
    uint8_t *used = (uint8_t *)calloc(n ? n : 1, 1);
    uint32_t *list = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    double *val = (double *)calloc(n ? n : 1, sizeof(double));
    uint32_t *link = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *cursor = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t *rp = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    double *d = ldlt ? (double *)calloc(n ? n : 1, sizeof(double)) : NULL;
    rowbuf b = {0, 0, NULL, NULL};
    for (uint32_t i = 0; i < n; ++i) link[i] = NIL;

    for (uint32_t j = 0; j < n; ++j) {
        uint32_t cnt = 0;
        double thr = 0.0;
        for (uint32_t k = a->rp[j]; k < a->rp[j + 1]; ++k) {
            uint32_t i = a->ci[k];
            if (i >= j) {
                thr += fabs(a->v[k]);
                if (!used[i]) { val[i] = 0.0; used[i] = 1; list[cnt++] = i; }
                val[i] += a->v[k];
            }
        }
        thr *= tau;
        uint32_t k = link[j];
        while (k != NIL) {
            uint32_t pos = cursor[k];
            double ljk = b.v[pos];
            double dk = ldlt ? d[k] : 0.0;
            for (uint32_t q = pos; q < rp[k + 1]; ++q) {
                uint32_t i = b.ci[q];
                if (!used[i]) { used[i] = 1; val[i] = 0.0; list[cnt++] = i; }
                if (ldlt) val[i] -= b.v[q] * ljk * dk;
                else val[i] -= b.v[q] * ljk;
            }
            pos += 1;
            uint32_t knext = link[k];
            if (pos < rp[k + 1]) {
                cursor[k] = pos;
                uint32_t i = b.ci[pos];
                link[k] = link[i];
                link[i] = k;
            }
            k = knext;
        }
        double piv = val[j];
        if (!ldlt) {
            if (piv <= 0.0) { /* IncompleteCholesky.cpp:189-193: not SPD */
                p->status = 1;
                p->fail_row = j;
                rp[j + 1] = b.nnz;
                for (uint32_t jj = j + 1; jj < n; ++jj) rp[jj + 1] = b.nnz;
                break;
            }
            piv = sqrt(piv);
            val[j] = piv;
            rb_push(&b, j, piv);
        } else {
            if (fabs(piv) < 1e-15) { /* IncompleteLDLT.cpp:131-133: zero pivot */
                p->status = 3;
                p->fail_row = j;
                rp[j + 1] = b.nnz;
                for (uint32_t jj = j + 1; jj < n; ++jj) rp[jj + 1] = b.nnz;
                break;
            }
            d[j] = piv;
        }
        cursor[j] = b.nnz;
        qsort(list, cnt, sizeof(uint32_t), cmp_u32);
        int first = 1;
        for (uint32_t q = 0; q < cnt; ++q) {
            uint32_t i = list[q];
            if (i != j) {
                double lij = val[i];
                if (fabs(lij) >= thr) {
                    if (first) {
                        first = 0;
                        link[j] = link[i];
                        link[i] = j;
                    }
                    rb_push(&b, i, lij / piv);
                }
            }
            used[i] = 0;
        }
        rp[j + 1] = b.nnz;
    }
    p->lt = rb_finish(&b, n, rp);
    p->d = d;
    p->mnnz = p->lt->nnz * 2 + n; /* IncompleteCholesky.cpp:245 / IncompleteLDLT.cpp:185 */
    free(used); free(list); free(val); free(link); free(cursor);
    return p;
}

ppo_prec *ppo_ict(const ppo_csr *a, double tau) { return threshold_cholesky(a, tau, 0); }
ppo_prec *ppo_ildlt(const ppo_csr *a, double tau) { return threshold_cholesky(a, tau, 1); }

/* Partial quick-select that moves the ncut largest |a| to the front; the
 * resulting order is the stored order of ILUT rows.  Restates qsplit
 * (include/VectorClass/VectorTools.h:159-201) step for step, because the
 * exact permutation it leaves is observable in the factor. */
static void quick_split(double *a, int *ind, int n, int ncut_count) {
    int ncut = ncut_count - 1, first = 0, last = n - 1, mid;
    if (ncut < first || ncut > last) return;
    do {
        mid = first;
        double key = fabs(a[mid]);
        for (int j = first + 1; j <= last; ++j) {
            if (fabs(a[j]) > key) {
                mid = mid + 1;
                double t = a[mid]; a[mid] = a[j]; a[j] = t;
                int it = ind[mid]; ind[mid] = ind[j]; ind[j] = it;
            }
        }
        double t = a[mid]; a[mid] = a[first]; a[first] = t;
        int it = ind[mid]; ind[mid] = ind[first]; ind[first] = it;
        if (mid == ncut) break;
        if (mid > ncut) last = mid - 1;
        else first = mid + 1;
    } while (mid != ncut);
}

/*
 * ILUT(lfil, tau), Saad's row-wise IKJ elimination.  Restates
 * IncompleteLU::setup (source/Preconditioner/IncompleteLU.cpp:38-266):
 *  - row tolerance tau * (sum_j |a_ij|) / nnz_row  (:85-91); zero row -> error;
 *  - pivot rows taken in ascending column order by a selection scan (:121-145);
 *  - a new fill entry with |l*u| below the tolerance is dropped (:162, :178);
 *  - the lfil largest of L and of U are kept (quick_split order, :207-259);
 *  - D stores 1/u_ii (:204); zero pivot -> error (:199).
 * The one case the reference leaves undefined (an upper fill entry whose
 * magnitude equals the tolerance exactly, :162/:178) is treated as dropped.
 */
ppo_prec *ppo_ilut(const ppo_csr *a, uint32_t lfil, double tau) {
    int n = (int)a->nrow;
    ppo_prec *p = (ppo_prec *)calloc(1, sizeof(ppo_prec));
    p->kind = PPO_PREC_ILU;
    int *iw = (int *)malloc((n ? n : 1) * sizeof(int));
    int *jbuf = (int *)malloc((n ? n : 1) * sizeof(int));
    double *w = (double *)calloc(n ? n : 1, sizeof(double));
    double *wn = (double *)calloc(n ? n : 1, sizeof(double));
    double *D = (double *)calloc(n ? n : 1, sizeof(double));
    uint32_t *lrp = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    uint32_t *urp = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    rowbuf lb = {0, 0, NULL, NULL}, ub = {0, 0, NULL, NULL};
    for (int i = 0; i < n; ++i) iw[i] = -1;

    for (int i = 0; i < n; ++i) {
        uint32_t beg = a->rp[i], end = a->rp[i + 1];
        double tnorm = 0.0;
        for (uint32_t k = beg; k < end; ++k) tnorm += fabs(a->v[k]);
        if (tnorm < 1e-15) {
            p->status = 2;
            p->fail_row = (uint32_t)i;
            for (int r = i; r < n; ++r) { lrp[r + 1] = lb.nnz; urp[r + 1] = ub.nnz; }
            break;
        }
        tnorm /= (double)(end - beg);
        double tolnorm = tau * tnorm;
        int lenl = 0, lenu = 0;
        jbuf[i] = i;
        w[i] = 0.0;
        iw[i] = i;
        for (uint32_t k = beg; k < end; ++k) {
            int col = (int)a->ci[k];
            double t = a->v[k];
            if (col < i) {
                iw[col] = lenl; jbuf[lenl] = col; w[lenl] = t; lenl++;
            } else if (col == i) {
                w[i] = t;
            } else {
                lenu++;
                int jpos = i + lenu;
                iw[col] = jpos; jbuf[jpos] = col; w[jpos] = t;
            }
        }
        for (int j = 0; j < lenl; ++j) {
            int jrow = jbuf[j], jpos = j;
            for (int k = j + 1; k < lenl; ++k)
                if (jbuf[k] < jrow) { jrow = jbuf[k]; jpos = k; }
            if (jpos != j) {
                int col = jbuf[j];
                jbuf[j] = jbuf[jpos]; jbuf[jpos] = col;
                iw[jrow] = j; iw[col] = jpos;
                double t = w[j]; w[j] = w[jpos]; w[jpos] = t;
            }
            double fact = w[j] * D[jrow];
            w[j] = fact;
            iw[jrow] = -1;
            for (uint32_t k = urp[jrow]; k < urp[jrow + 1]; ++k) {
                int col = (int)ub.ci[k];
                int pos = iw[col];
                double lxu = -fact * ub.v[k];
                if (fabs(lxu) < tolnorm && pos == -1) continue;
                if (col < i) {
                    if (pos == -1) {
                        jbuf[lenl] = col; iw[col] = lenl; w[lenl] = lxu; lenl++;
                    } else {
                        w[pos] += lxu;
                    }
                } else {
                    if (pos == -1) {
                        if (fabs(lxu) > tolnorm) {
                            lenu++;
                            int upos = i + lenu;
                            jbuf[upos] = col; iw[col] = upos; w[upos] = lxu;
                        }
                    } else {
                        w[pos] += lxu;
                    }
                }
            }
        }
        iw[i] = -1;
        for (int j = 0; j < lenu; ++j) iw[jbuf[i + j + 1]] = -1;
        if (fabs(w[i]) < 1e-15) {
            p->status = 3;
            p->fail_row = (uint32_t)i;
            for (int r = i; r < n; ++r) { lrp[r + 1] = lb.nnz; urp[r + 1] = ub.nnz; }
            break;
        }
        D[i] = 1.0 / w[i];
        int len = lenl < (int)lfil ? lenl : (int)lfil;
        for (int j = 0; j < lenl; ++j) { wn[j] = fabs(w[j]); iw[j] = j; }
        quick_split(wn, iw, lenl, len);
        for (int j = 0; j < len; ++j) rb_push(&lb, (uint32_t)jbuf[iw[j]], w[iw[j]]);
        lrp[i + 1] = lb.nnz;
        for (int j = 0; j < lenl; ++j) iw[j] = -1;
        len = lenu < (int)lfil ? lenu : (int)lfil;
        for (int j = 0; j < lenu; ++j) { wn[j] = fabs(w[i + j + 1]); iw[j] = i + j + 1; }
        quick_split(wn, iw, lenu, len);
        for (int j = 0; j < len; ++j) rb_push(&ub, (uint32_t)jbuf[iw[j]], w[iw[j]]);
        urp[i + 1] = ub.nnz;
        for (int j = 0; j < lenu; ++j) iw[j] = -1;
    }
    p->l = rb_finish(&lb, (uint32_t)n, lrp);
    p->u = rb_finish(&ub, (uint32_t)n, urp);
    p->d = D;
    p->mnnz = p->l->nnz + p->u->nnz + (uint32_t)n; /* IncompleteLU.cpp:264 */
    free(iw); free(jbuf); free(w); free(wn);
    return p;
}

/*
 * IC(0): incomplete Cholesky with the sparsity pattern of tril(A), no fill.
 * No reference counterpart.  Row-oriented (up-looking):
 *   L_ik = (a_ik - sum_{j<k, j in row i and row k} L_ij L_kj) / L_kk   (k < i ascending),
 *   L_ii = sqrt(a_ii - sum_{k<i} L_ik^2),
 * with the inner sums taken in ascending j.  Stored as L^T with the diagonal
 * first, like the threshold factor, so the same apply serves both.
 */
ppo_prec *ppo_ic0(const ppo_csr *a) {
    uint32_t n = a->nrow;
    ppo_prec *p = (ppo_prec *)calloc(1, sizeof(ppo_prec));
    p->kind = PPO_PREC_IC;
    /* strict-lower rows of L, then diagonal */
    uint32_t *lrp = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t c = 0;
        for (uint32_t k = a->rp[i]; k < a->rp[i + 1]; ++k) c += a->ci[k] < i;
        lrp[i + 1] = lrp[i] + c;
    }
    uint32_t lnnz = lrp[n];
    uint32_t *lci = (uint32_t *)malloc((lnnz ? lnnz : 1) * sizeof(uint32_t));
    double *lv = (double *)malloc((lnnz ? lnnz : 1) * sizeof(double));
    double *diag = (double *)calloc(n ? n : 1, sizeof(double));
    int64_t *mark = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
    for (uint32_t i = 0; i < n; ++i) mark[i] = -1;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t q = lrp[i];
        double aii = 0.0;
        for (uint32_t k = a->rp[i]; k < a->rp[i + 1]; ++k) {
            uint32_t c = a->ci[k];
            if (c < i) { lci[q] = c; lv[q] = a->v[k]; mark[c] = q; q++; }
            else if (c == i) aii = a->v[k];
        }
        for (uint32_t t = lrp[i]; t < lrp[i + 1]; ++t) {
            uint32_t k = lci[t];
            double s = lv[t];
            for (uint32_t u = lrp[k]; u < lrp[k + 1]; ++u) {
                int64_t pos = mark[lci[u]];
                if (pos >= 0) s -= lv[pos] * lv[u];
            }
            lv[t] = s / diag[k];
        }
        double s = aii;
        for (uint32_t t = lrp[i]; t < lrp[i + 1]; ++t) s -= lv[t] * lv[t];
        for (uint32_t t = lrp[i]; t < lrp[i + 1]; ++t) mark[lci[t]] = -1;
        if (s <= 0.0) {
            p->status = 1;
            p->fail_row = i;
            s = 1.0;
        }
        diag[i] = sqrt(s);
    }
    /* L^T: row j = diagonal then column j of L in ascending row order */
    ppo_csr *lt = ppo_csr_new(n, n, lnnz + n);
    for (uint32_t t = 0; t < lnnz; ++t) lt->rp[lci[t] + 1]++;
    for (uint32_t j = 0; j < n; ++j) lt->rp[j + 1] += lt->rp[j] + 1;
    uint32_t *fill = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
    for (uint32_t j = 0; j < n; ++j) {
        lt->ci[lt->rp[j]] = j;
        lt->v[lt->rp[j]] = diag[j];
        fill[j] = lt->rp[j] + 1;
    }
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t t = lrp[i]; t < lrp[i + 1]; ++t) {
            uint32_t pos = fill[lci[t]]++;
            lt->ci[pos] = i;
            lt->v[pos] = lv[t];
        }
    p->lt = lt;
    p->mnnz = (lnnz + n) * 2 + n; /* L^T entries incl. the diagonal, the IC convention */
    free(fill); free(lrp); free(lci); free(lv); free(diag); free(mark);
    return p;
}

/*
 * ILU(0): incomplete LU with the pattern of A, no fill.  No reference
 * counterpart.  IKJ form: for each row i and each k < i in row i (ascending):
 * a_ik /= u_kk; a_ij -= a_ik * u_kj for j > k in row i and row k.
 * Output in the ILUT layout: L strict lower (unit), U strict upper (ascending
 * columns), D = 1/u_ii.
 */
ppo_prec *ppo_ilu0(const ppo_csr *a) {
    uint32_t n = a->nrow, nnz = a->nnz;
    ppo_prec *p = (ppo_prec *)calloc(1, sizeof(ppo_prec));
    p->kind = PPO_PREC_ILU;
    double *w = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
    memcpy(w, a->v, (size_t)nnz * sizeof(double));
    uint32_t *dpos = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    int64_t *mark = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
    for (uint32_t i = 0; i < n; ++i) { mark[i] = -1; dpos[i] = NIL; }
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = a->rp[i]; k < a->rp[i + 1]; ++k)
            if (a->ci[k] == i) dpos[i] = k;
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t k = a->rp[i]; k < a->rp[i + 1]; ++k) mark[a->ci[k]] = k;
        if (dpos[i] == NIL) { p->status = 3; p->fail_row = i; break; }
        for (uint32_t t = a->rp[i]; t < a->rp[i + 1]; ++t) {
            uint32_t k = a->ci[t];
            if (k >= i) break;
            w[t] /= w[dpos[k]];
            for (uint32_t u = dpos[k] + 1; u < a->rp[k + 1]; ++u) {
                int64_t pos = mark[a->ci[u]];
                if (pos >= 0) w[pos] -= w[t] * w[u];
            }
        }
        for (uint32_t k = a->rp[i]; k < a->rp[i + 1]; ++k) mark[a->ci[k]] = -1;
        if (fabs(w[dpos[i]]) < 1e-15) { p->status = 3; p->fail_row = i; break; }
    }
    uint32_t *lrp = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    uint32_t *urp = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    rowbuf lb = {0, 0, NULL, NULL}, ub = {0, 0, NULL, NULL};
    double *D = (double *)calloc(n ? n : 1, sizeof(double));
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t k = a->rp[i]; k < a->rp[i + 1]; ++k) {
            uint32_t c = a->ci[k];
            if (c < i) rb_push(&lb, c, w[k]);
            else if (c > i) rb_push(&ub, c, w[k]);
        }
        lrp[i + 1] = lb.nnz;
        urp[i + 1] = ub.nnz;
        D[i] = dpos[i] != NIL ? 1.0 / w[dpos[i]] : 0.0;
    }
    p->l = rb_finish(&lb, n, lrp);
    p->u = rb_finish(&ub, n, urp);
    p->d = D;
    p->mnnz = p->l->nnz + p->u->nnz + n;
    free(w); free(dpos); free(mark);
    return p;
}

/* IC forward sweep: column scatter over the rows of L^T
 * (IncompleteCholesky.cpp:250-271); ILU unit-lower gather (IncompleteLU.cpp:285-297);
 * ILDLT unit scatter (IncompleteLDLT.cpp:217-233). */
void ppo_prec_lower(const ppo_prec *p, double *vec) {
    if (p->kind == PPO_PREC_IC) {
        const ppo_csr *t = p->lt;
        for (uint32_t i = 0; i < t->nrow; ++i) {
            uint32_t k = t->rp[i];
            double y = vec[i] / t->v[k];
            vec[i] = y;
            for (k = k + 1; k < t->rp[i + 1]; ++k) vec[t->ci[k]] -= t->v[k] * y;
        }
    } else if (p->kind == PPO_PREC_ILU) {
        const ppo_csr *l = p->l;
        for (uint32_t i = 0; i < l->nrow; ++i)
            for (uint32_t k = l->rp[i]; k < l->rp[i + 1]; ++k) vec[i] -= l->v[k] * vec[l->ci[k]];
    } else if (p->kind == PPO_PREC_ILDLT) {
        const ppo_csr *t = p->lt;
        for (uint32_t i = 0; i < t->nrow; ++i)
            for (uint32_t k = t->rp[i]; k < t->rp[i + 1]; ++k) vec[t->ci[k]] -= t->v[k] * vec[i];
    }
}

/* ILDLT diagonal (IncompleteLDLT.cpp:207-214); identity for IC / ILU. */
void ppo_prec_diag(const ppo_prec *p, double *vec) {
    if (p->kind != PPO_PREC_ILDLT) return;
    for (uint32_t i = 0; i < p->lt->nrow; ++i) vec[i] /= p->d[i];
}

/* IC backward gather (IncompleteCholesky.cpp:275-295); ILU upper gather times
 * 1/u_ii (IncompleteLU.cpp:269-282); ILDLT L^T gather (IncompleteLDLT.cpp:189-204). */
void ppo_prec_upper(const ppo_prec *p, double *vec) {
    if (p->kind == PPO_PREC_IC) {
        const ppo_csr *t = p->lt;
        for (uint32_t i = t->nrow; i-- > 0;) {
            double s = vec[i];
            for (uint32_t k = t->rp[i] + 1; k < t->rp[i + 1]; ++k) s -= t->v[k] * vec[t->ci[k]];
            vec[i] = s / t->v[t->rp[i]];
        }
    } else if (p->kind == PPO_PREC_ILU) {
        const ppo_csr *u = p->u;
        for (uint32_t i = u->nrow; i > 0; --i) {
            for (uint32_t k = u->rp[i - 1]; k < u->rp[i]; ++k) vec[i - 1] -= u->v[k] * vec[u->ci[k]];
            vec[i - 1] *= p->d[i - 1];
        }
    } else if (p->kind == PPO_PREC_ILDLT) {
        const ppo_csr *t = p->lt;
        for (uint32_t i = t->nrow; i-- > 0;)
            for (uint32_t k = t->rp[i]; k < t->rp[i + 1]; ++k) vec[i] -= t->v[k] * vec[t->ci[k]];
    }
}

/* TriangularPrecondition::MInvSolve (include/Preconditioner/BasePreconditon.h:307-316). */
void ppo_prec_apply(const ppo_prec *p, double *vec) {
    if (!p || p->kind == PPO_PREC_NONE) return;
    ppo_prec_lower(p, vec);
    ppo_prec_diag(p, vec);
    ppo_prec_upper(p, vec);
}

/* ------------------------------------------------------------------ */
/* Solvers                                                              */
/* ------------------------------------------------------------------ */

/*
 * CG::solve (source/Solver/CG.cpp:33-83), step for step, including its
 * second product A*d per iteration (:68), its coefficients
 * beta = (w, A d_old)/(d_old, A d_old), alpha = (d, r)/(d, A d), the check of
 * the recursive residual at the top of iterations 1..maxit-1, and the count
 * iters = m + 1 (maxit + 1 when not converged).  flexible = 0 is NormalCG:
 * M^{-1} applied to r_0 only (:48, :65).  hist[i] = ||r_i|| / ||b||.
 */
int ppo_cg(const ppo_csr *a, const ppo_prec *p, int flexible, const double *b,
           double *x, double tol, uint32_t maxit, double *hist, uint32_t hist_cap,
           ppo_report *rep) {
    uint32_t n = a->nrow;
    double *r = (double *)malloc(n * sizeof(double));
    double *v = (double *)malloc(n * sizeof(double));
    double *w = (double *)malloc(n * sizeof(double));
    double *d = (double *)malloc(n * sizeof(double));
    int conv = 0;
    uint32_t m = maxit;
    double normb = ppo_norm2(n, b), norm = -1.0;
    int32_t nh = 0;
    ppo_spmv(a, x, r);
    for (uint32_t i = 0; i < n; ++i) w[i] = 1.0 * b[i] + -1.0 * r[i];
    memcpy(r, w, n * sizeof(double));
    if (p) ppo_prec_apply(p, w);
    memcpy(d, w, n * sizeof(double));
    ppo_spmv(a, d, v);
    double v1 = ppo_dot(n, d, r);
    double v2 = ppo_dot(n, d, v);
    v1 /= v2;
    for (uint32_t i = 0; i < n; ++i) x[i] = 1.0 * x[i] + v1 * d[i];
    for (uint32_t i = 0; i < n; ++i) r[i] = 1.0 * r[i] + -v1 * v[i];
    for (uint32_t it = 1; it < maxit; ++it) {
        norm = ppo_norm2(n, r) / normb;
        if (hist && (uint32_t)nh < hist_cap) hist[nh] = norm;
        nh++;
        if (norm < tol) { conv = 1; m = it; break; }
        memcpy(w, r, n * sizeof(double));
        if (flexible && p) ppo_prec_apply(p, w);
        ppo_spmv(a, d, v);
        v1 = ppo_dot(n, w, v);
        v2 = ppo_dot(n, d, v);
        v1 = v1 / v2;
        for (uint32_t i = 0; i < n; ++i) d[i] = 1.0 * w[i] + -v1 * d[i];
        v1 = ppo_dot(n, d, r);
        ppo_spmv(a, d, v);
        v2 = ppo_dot(n, d, v);
        v1 /= v2;
        for (uint32_t i = 0; i < n; ++i) x[i] = 1.0 * x[i] + v1 * d[i];
        for (uint32_t i = 0; i < n; ++i) r[i] = 1.0 * r[i] + -v1 * v[i];
    }
    rep->iters = (int32_t)(m + 1);
    rep->converged = conv;
    rep->final_error = norm;
    rep->nhist = nh;
    free(r); free(v); free(w); free(d);
    return 0;
}

/* GMRES::givensTransform (source/Solver/GMRES.cpp:72-101); H column-major
 * with leading dimension ld. */
static void givens(double *H, int ld, double *c, double *s, double *ksi, int j) {
#define HH(i, k) H[(size_t)(k) * ld + (i)]
    for (int i = 0; i < j; ++i) {
        double h1 = HH(i, j), h2 = HH(i + 1, j);
        HH(i + 1, j) = c[i] * h2 - s[i] * h1;
        HH(i, j) = c[i] * h1 + s[i] * h2;
    }
    double h1 = HH(j, j), h2 = HH(j + 1, j);
    if (fabs(h1) > fabs(h2)) {
        double tao = h2 / h1;
        c[j] = 1.0 / sqrt(fabs(1.0 + tao * tao));
        s[j] = c[j] * tao;
    } else {
        double tao = h1 / h2;
        s[j] = 1.0 / sqrt(fabs(1.0 + tao * tao));
        c[j] = s[j] * tao;
    }
    HH(j, j) = c[j] * h1 + s[j] * h2;
    HH(j + 1, j) = 0.0;
    ksi[j + 1] = -s[j] * ksi[j];
    ksi[j] = c[j] * ksi[j];
#undef HH
}

/*
 * GMRES::solve with setRestart(true, restart) (source/Solver/GMRES.cpp:59-70,
 * 103-221), step for step:
 *  - cycles allowed: maxit_total / restart + 1 (:63, :155);
 *  - Arnoldi by modified Gram-Schmidt (:174-180), Kahan norms (:150, :181);
 *  - in the converged cycle the back-substitution uses dim = j, dropping
 *    column j (:194-201); ksi[j+1] (the residual estimate) is left in place
 *    and multiplies column j+1 of the persistent, zero-initialised Z (or V),
 *    because the update multiplies all restart+1 columns (:204-208);
 *  - restart beta = the Givens estimate, not ||b - A x|| (:214);
 *  - iters = k * restart + m (:218).
 * flexible = 0 is NormalGMRES: the preconditioner is ignored (:164-171).
 */
int ppo_gmres(const ppo_csr *a, const ppo_prec *p, int flexible, const double *b,
              double *x, double tol, uint32_t restart, uint32_t maxit_total,
              double *hist, uint32_t hist_cap, ppo_report *rep) {
    uint32_t n = a->nrow;
    uint32_t mm = restart;
    int32_t maxcyc = (int32_t)(maxit_total / restart);
    if (mm >= n) mm = n - 1;
    uint32_t m = mm;
    size_t ldv = n;
    double *V = (double *)calloc((size_t)n * (mm + 1), sizeof(double));
    double *Z = (double *)calloc((size_t)n * (mm + 1), sizeof(double));
    double *H = (double *)calloc((size_t)(mm + 1) * mm, sizeof(double));
    double *ksi = (double *)calloc(mm + 1, sizeof(double));
    double *cs = (double *)calloc(mm + 1, sizeof(double));
    double *sn = (double *)calloc(mm + 1, sizeof(double));
    double *vj = (double *)malloc(n * sizeof(double));
    double *wj = (double *)malloc(n * sizeof(double));
    double *r = (double *)malloc(n * sizeof(double));
    int ld = (int)(mm + 1);
    double beta, rnorm = 0.0, relerr = 0.0;
    int conv = 0;
    int32_t nh = 0;
    double bnorm = sqrt(kahan_sumsq(n, b));
    ppo_spmv(a, x, r);
    for (uint32_t i = 0; i < n; ++i) vj[i] = 1.0 * b[i] + -1.0 * r[i];
    beta = sqrt(kahan_sumsq(n, vj));
    int32_t k = -1;
    while (!conv && k < maxcyc) {
        ++k;
        double sc = 1.0 / beta;
        for (uint32_t i = 0; i < n; ++i) vj[i] *= sc;
        memcpy(V, vj, n * sizeof(double));
        for (uint32_t i = 0; i < mm; ++i) ksi[i] = 0.0;
        ksi[0] = beta;
        for (uint32_t j = 0; j < m; ++j) {
            if (flexible) {
                if (p) ppo_prec_apply(p, vj);
                memcpy(Z + ldv * j, vj, n * sizeof(double));
            }
            ppo_spmv(a, vj, wj);
            for (uint32_t i = 0; i < j + 1; ++i) {
                memcpy(vj, V + ldv * i, n * sizeof(double));
                double h = ppo_dot(n, wj, vj);
                H[(size_t)j * ld + i] = h;
                for (uint32_t q = 0; q < n; ++q) vj[q] *= h;
                for (uint32_t q = 0; q < n; ++q) { double t = wj[q]; t += -1.0 * vj[q]; wj[q] = t; }
            }
            double h = sqrt(kahan_sumsq(n, wj));
            H[(size_t)j * ld + j + 1] = h;
            if (fabs(h) == 0) { m = j; break; }
            memcpy(vj, wj, n * sizeof(double));
            sc = 1.0 / h;
            for (uint32_t q = 0; q < n; ++q) vj[q] *= sc;
            memcpy(V + ldv * (j + 1), vj, n * sizeof(double));
            givens(H, ld, cs, sn, ksi, (int)j);
            rnorm = fabs(ksi[j + 1]);
            relerr = rnorm / bnorm;
            if (hist && (uint32_t)nh < hist_cap) hist[nh] = relerr;
            nh++;
            if (relerr < tol) { m = j; break; }
        }
        /* retrospectiveSolver(H, ksi, m) (GMRES.cpp:103-119) */
        for (int i = (int)m - 1; i >= 0; --i) {
            for (int j = i + 1; j < (int)m; ++j) {
                double t = ksi[i];
                t -= H[(size_t)j * ld + i] * ksi[j];
                ksi[i] = t;
            }
            ksi[i] /= H[(size_t)i * ld + i];
        }
        ksi[m] = 0.0;
        /* DenseMatrix::MatVec over all mm+1 columns (DenseMatrix.cpp:208-224) */
        const double *B = flexible ? Z : V;
        for (uint32_t q = 0; q < n; ++q) wj[q] = 0.0;
        for (uint32_t c = 0; c <= mm; ++c)
            for (uint32_t q = 0; q < n; ++q) wj[q] += B[ldv * c + q] * ksi[c];
        for (uint32_t q = 0; q < n; ++q) { double t = x[q]; t += 1.0 * wj[q]; x[q] = t; }
        conv = relerr < tol;
        ppo_spmv(a, x, r);
        for (uint32_t i = 0; i < n; ++i) vj[i] = 1.0 * b[i] + -1.0 * r[i];
        beta = rnorm;
    }
    rep->iters = (int32_t)(k * (int32_t)mm + (int32_t)m);
    rep->converged = conv;
    rep->final_error = relerr;
    rep->nhist = nh;
    free(V); free(Z); free(H); free(ksi); free(cs); free(sn); free(vj); free(wj); free(r);
    return 0;
}
