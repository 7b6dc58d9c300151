// pp_ml.h -- internal types of the multilevel low-rank Schur preconditioners.
#pragma once
#include "pp_internal.h"

namespace pp {

// Hierarchical interface decomposition: new position k holds original row
// perm[k]; level l owns positions [lev_ptr[l], lev_ptr[l+1]); its blocks are
// dom_ptr[l] (absolute positions).
struct Hid {
    uint32_t nlev = 0;
    std::vector<uint32_t> perm, iperm, lev_ptr;
    std::vector<std::vector<uint32_t>> dom_ptr;
};

Hid build_hid(const HostCsr &A, uint32_t ndom, uint32_t nlev);
HostCsr extract_block(const HostCsr &A, const Hid &H, uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1,
                      uint32_t &ncol);
HostCsr transpose_rect(const HostCsr &A, uint32_t ncol);
void tridiag_eig(std::vector<double> &d, std::vector<double> e, std::vector<double> &Z);
bool hessenberg_eigenvalues(std::vector<double> H, int n, std::vector<double> &re, std::vector<double> &im);
bool hessenberg_schur(std::vector<double> &H, int n, std::vector<double> &Z, std::vector<double> &wr,
                      std::vector<double> &wi);
std::vector<double> inverse_iteration(const std::vector<double> &H, int n, double lambda);
// W (n x r, col-major) = V (n x k, ld n) * Y (k x r, ld k), FP64 on the matrix cores
void launch_gemm_tall_f64(pp_ctx *ctx, uint32_t n, uint32_t k, uint32_t r, const double *V, const double *Y,
                          double *W);

// One level l < nlev-1 of A_l = [B E; F C] (F = E^T for AMSED).
struct MlLevel {
    uint32_t off = 0, nb = 0, nrest = 0;  // B rows start at off (new numbering)
    pp_csr *B = nullptr, *E = nullptr, *F = nullptr;
    pp_precond *Bpc = nullptr;
    double *W = nullptr, *Hd = nullptr;   // low-rank term W H W^T (device)
    double *Md = nullptr;                 // nonsymmetric: dense r x r core (W orthonormal)
    uint32_t rank = 0;
    double *save = nullptr, *t = nullptr, *h = nullptr;  // device work vectors
    float *W32 = nullptr, *save32 = nullptr, *t32 = nullptr;  // FP32 apply
    std::vector<double> Hhost, theta;     // host copy of H and the Ritz values
};

struct MlState {
    bool symmetric = true;                // AMSED (true) or GMSLR (false)
    Hid hid;
    std::vector<MlLevel> lev;             // nlev-1 entries
    pp_csr *Clast = nullptr;
    pp_precond *last = nullptr;
    uint32_t *dperm = nullptr;            // device perm
    double *xp = nullptr;                 // permuted work vector
    float *xp32 = nullptr;
    bool have32 = false;                  // FP32 copies of every operand exist
    uint32_t n = 0;
    uint32_t lanczos_steps = 0;
    int lowrank = 0;                      // PP_LOWRANK_* form the low-rank terms were built with
};

}  // namespace pp
