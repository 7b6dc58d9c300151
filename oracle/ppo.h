/*
 * ppo.h -- the oracle: a plain-C, single-threaded CPU restatement of the
 * PrisPack preconditioned-Krylov hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (prispack_amd/)
 * links, loads or calls this code.  It is used by tests/ (as the checker),
 * by __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline
 * leg (timed beside the GPU, never as the measured product).
 *
 * Every function names the reference file:line it restates.  All arithmetic
 * is IEEE double, evaluated in the same order as the reference's serial code
 * paths, so that with strict compile flags (-ffp-contract=off, no fast-math)
 * the factorizations and sweeps are bit-identical to the reference's.
 */
#ifndef PPO_H
#define PPO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CSR with 32-bit unsigned indices, 0-based, as HOST::CSRMatrix
 * (include/MatrixClass/CSRMatrix.h:19-31). */
typedef struct {
    uint32_t nrow, ncol, nnz;
    uint32_t *rp;   /* nrow+1 */
    uint32_t *ci;   /* nnz    */
    double *v;      /* nnz    */
} ppo_csr;

ppo_csr *ppo_csr_new(uint32_t nrow, uint32_t ncol, uint32_t nnz);
void ppo_csr_free(ppo_csr *a);
ppo_csr *ppo_csr_from(uint32_t nrow, uint32_t ncol, const uint32_t *rp,
                      const uint32_t *ci, const double *v);
ppo_csr *ppo_csr_transpose(const ppo_csr *a);

/* ---------------- seeded inputs ---------------- */
uint64_t ppo_splitmix64(uint64_t *state);
void ppo_fill_uniform(double *x, uint64_t n, uint64_t seed);
void ppo_fill_normal(double *x, uint64_t n, uint64_t seed);

/* ---------------- model problems ---------------- */
ppo_csr *ppo_laplacian(int nx, int ny, int nz, double ax, double ay, double az,
                       double shift);
ppo_csr *ppo_upwind(int nx, int ny, int nz, double bx, double by, double bz);
ppo_csr *ppo_vardiff2d(int side, const double *kappa);

/* ---------------- kernels ---------------- */
void ppo_spmv(const ppo_csr *a, const double *x, double *y);
void ppo_spmv_t_acc(const ppo_csr *a, const double *x, double *y);
double ppo_dot(uint32_t n, const double *x, const double *y);
double ppo_norm2(uint32_t n, const double *x);

/* ---------------- preconditioners ---------------- */
enum { PPO_PREC_NONE = 0, PPO_PREC_IC = 1, PPO_PREC_ILU = 2, PPO_PREC_ILDLT = 3 };

typedef struct {
    int kind;
    int status;       /* 0 ok, 1 non-SPD pivot, 2 zero row, 3 zero pivot */
    uint32_t fail_row;
    ppo_csr *lt;      /* IC: L^T with diagonal first; ILDLT: strict upper (unit) */
    ppo_csr *l;       /* ILU: strict lower, unit diagonal implicit */
    ppo_csr *u;       /* ILU: strict upper */
    double *d;        /* ILU: 1/u_ii;  ILDLT: d_ii */
    uint32_t mnnz;    /* getPreconditionNonZeroCount() */
} ppo_prec;

void ppo_prec_free(ppo_prec *p);
ppo_prec *ppo_ict(const ppo_csr *a, double tau);
ppo_prec *ppo_ilut(const ppo_csr *a, uint32_t lfil, double tau);
ppo_prec *ppo_ildlt(const ppo_csr *a, double tau);
ppo_prec *ppo_ic0(const ppo_csr *a);
ppo_prec *ppo_ilu0(const ppo_csr *a);
void ppo_prec_apply(const ppo_prec *p, double *vec);           /* in place */
void ppo_prec_lower(const ppo_prec *p, double *vec);
void ppo_prec_diag(const ppo_prec *p, double *vec);
void ppo_prec_upper(const ppo_prec *p, double *vec);

/* ---------------- solvers ---------------- */
typedef struct {
    int32_t iters;      /* reference counting convention */
    int32_t converged;
    double final_error; /* reference getError() */
    int32_t nhist;      /* entries written to hist */
} ppo_report;

int ppo_cg(const ppo_csr *a, const ppo_prec *p, int flexible, const double *b,
           double *x, double tol, uint32_t maxit, double *hist, uint32_t hist_cap,
           ppo_report *rep);
int ppo_gmres(const ppo_csr *a, const ppo_prec *p, int flexible, const double *b,
              double *x, double tol, uint32_t restart, uint32_t maxit_total,
              double *hist, uint32_t hist_cap, ppo_report *rep);

#ifdef __cplusplus
}
#endif
#endif
