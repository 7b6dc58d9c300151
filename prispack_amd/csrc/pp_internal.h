// pp_internal.h -- shared types of the MI355X hot-path library (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <functional>
#include <map>
#include <vector>

#include "../../include/prispack_amd.h"


namespace pp {

// A failure carried to the C ABI boundary as a status + message.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

#define PP_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::pp::Error(PP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define PP_CHECK(cond, status, msg)                        \
    do {                                                   \
        if (!(cond)) throw ::pp::Error((status), (msg));   \
    } while (0)

constexpr int kBlock = 256;             // threads per workgroup of the streaming kernels
constexpr uint32_t kSpmvNnzBlock = 2048; // nnz staged in LDS per SpMV workgroup
constexpr uint32_t kStreamLevelRows = 4096; // sweep levels this wide use the LDS-staged kernel
constexpr uint32_t kTuneSweepRows = 1u << 20; // factors this large time their sweep variants at setup
                                              // (and keep the sliced-column copies the timing weighs)
constexpr uint32_t kNoBase = 0xFFFFFFFFu;   // row block without 16-bit column offsets
// 16-bit column offsets of a row block: bit 15 selects one of the block's two
// bases, bits 0-14 the distance from it (two windows of 32768 columns each,
// e.g. the two colour halves of a colour-ordered matrix)
constexpr uint16_t kHiBase = 0x8000u, kOffMask = 0x7FFFu;
// Encodes cols[0..cnt) as two-base offsets; false if they do not fit.
bool two_base_offsets(const uint32_t *cols, uint32_t cnt, uint32_t &base_lo, uint32_t &base_hi, uint16_t *out);
// streaming-kernel variants (bit set): non-temporal operand loads, 16-bit
// offsets, batched loads (every load of a thread issued before the first use)
constexpr int kVarNT = 1, kVarIdx16 = 2, kVarBatch = 4;
constexpr int kVarCount = 8;
// the sliced-column form (EllOp) instead of CSR-stream: a ninth variant,
// timed only where the matrix has that form
constexpr int kVarEll = 8;
constexpr int kVarDefault = kVarIdx16 | kVarBatch;
// One row block of a staged kernel: first row (position) r0, first entry k0,
// the two 16-bit offset bases (kNoBase: the block reads 32-bit columns).  An
// array of nblocks + 1 descriptors; the next descriptor's r0 / k0 end a block.
struct alignas(16) BlockDesc {
    uint32_t r0, k0, blo, bhi;
};
constexpr uint32_t kSpmvRowsBlock = 256; // rows per SpMV workgroup (one per thread)
constexpr int kMaxPartials = 4096;       // reduction partials per slot

// Device scalar slots used by the Krylov kernels (FP64 even in mixed mode).
enum Slot : int {
    S_RR = 0,   // (r, r)
    S_DV,       // (d, A d)
    S_DR,       // (d, r)
    S_WV,       // (w, v)
    S_BB,       // (b, b)
    S_TMP0,
    S_TMP1,
    S_TMP2,
    S_TMP3,
    S_CG_DELTA,  // single-reduction PCG: (w, u), (r, u), (r, r) -- consecutive
    S_CG_GAMMA,
    S_CG_RHO,
    S_CG_ALPHA,  // alpha, beta, previous gamma, first-iteration flag
    S_CG_BETA,
    S_CG_GOLD,
    S_CG_FIRST,
    S_TOL,       // convergence tolerance of the running PCG (read by its check)
    // fused finish-and-update of the reduced CG: the iteration state it reads
    // and the one it writes alternate between two blocks (every workgroup
    // reads the state while the first writes the next one), each holding
    // RS_GOLD (previous gamma), RS_AOLD (previous alpha), RS_FIRST, RS_LAST
    // (the next check is the last), RS_CHECK (index of this check)
    S_RED0,
    S_RED1 = S_RED0 + 5,
    S_COUNT = S_RED1 + 5
};
enum RedState : int { RS_GOLD = 0, RS_AOLD, RS_FIRST, RS_LAST, RS_CHECK };
// Integer control words.
enum Ctl : int {
    C_STOP = 0,   // bodies skip when set
    C_CONV,       // converged
    C_M,          // reference loop index at exit
    C_CHECK,      // index of the next convergence check
    C_LAST,       // the next body is the last one
    C_STATUS,     // PP_OK / PP_ERR_BREAKDOWN / PP_ERR_NOT_FINITE
    C_SPIN,       // a sync-free sweep hit its wait deadline
    C_PUB,        // iteration bodies that have published their control words
    C_MAXIT,      // iteration cap of the running PCG (read by its check)
    C_NORM_SET,   // a check of this solve has written its norm (S_TMP0) ...
    C_NORM_LO,    // ... whose bits are these two words, so that the host
    C_NORM_HI,    //     reads it with the control words from the ring
    C_COUNT = 16
};

// Deterministic reduction workspace: one row of per-workgroup partials per slot.
struct RedWs {
    double *partials = nullptr;  // [S_COUNT][kMaxPartials]
};

struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    void reset();
    ~Graph() { reset(); }
};

}  // namespace pp

struct pp_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    int cu_count = 256;
    size_t max_lds = 65536;         // LDS bytes one workgroup may use (device attribute)
    pp::RedWs red;
    double *scal = nullptr;       // device scalars [S_COUNT]
    int *ctl = nullptr;           // device control words [C_COUNT]
    double *hist = nullptr;       // device residual history
    uint32_t hist_cap = 0;
    double *pinned = nullptr;     // pinned host mirror (scalars, 64 doubles)
    void *flush_buf = nullptr;    // cache-flush buffer of the variant timers (allocated on first use)
    // device blocks of finished setup temporaries (size -> block), taken by
    // later allocations inside a PoolScope on this context; freed with it
    std::multimap<size_t, void *> pool;
    size_t pool_bytes = 0;
    bool pool_dirty = false;  // blocks returned since the stream's last synchronization
    int *pinned_ctl = nullptr;    // pinned host mirror of ctl
    int *ring_host = nullptr;     // mapped pinned slots the PCG bodies publish ctl into
    int *ring_dev = nullptr;      // the same memory as the device sees it
    std::vector<double> last_hist;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    static constexpr int kRing = 4;      // publication slots of ring_host
    // iteration bodies kept in flight: one running and one queued hides the
    // host's poll-and-launch latency; every further one only adds bodies
    // that run (as early-exit kernels) after convergence
    static constexpr int kInFlight = 2;
    void *ws[48] = {};             // cached solver workspaces
    size_t ws_bytes[48] = {};
    // while arena is set (pp_solve_noalloc), ws_get carves the workspace
    // slots out of [arena, arena + arena_cap) instead of allocating
    unsigned char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    size_t arena_off[48] = {}, arena_len[48] = {};
    long long spin_deadline = 0;   // wall_clock64 ticks a sync-free wait may last
    // while set (StopScope), SpMV / dot / GEMV / convert kernels return at
    // once when the word is non-zero, like the sweeps and the PCG kernels
    const int *stop_word = nullptr;
    // instantiated iteration-body graphs, keyed by everything their kernels
    // were launched with (object ids, buffers, options); reused across solves
    std::vector<std::pair<std::vector<uint64_t>, std::unique_ptr<pp::Graph>>> graphs;
};

namespace pp {

// Kernels launched inside the scope skip their work once ctl[C_STOP] is set
// (iteration bodies issued ahead of the host's convergence check).
struct StopScope {
    pp_ctx *c;
    const int *prev;
    explicit StopScope(pp_ctx *ctx);
    ~StopScope() { c->stop_word = prev; }
};

// Device-side copy of a gather-form triangular sweep, rows stored in level order.
template <class T>
struct DevSweep {
    uint32_t n = 0, nnz = 0, nlev = 0;
    uint32_t *perm = nullptr;   // [n] row processed at position k
    uint32_t *rp = nullptr;     // [n+1] positions in level order
    uint32_t *ci = nullptr;     // [nnz] dependency rows (original numbering)
    T *val = nullptr;           // [nnz]
    T *diag = nullptr;          // [n] (original numbering) or nullptr
    std::vector<uint32_t> level_ptr;  // host [nlev+1]
    // host [nlev]: perm[k] = k + level_off[l] over level l when the level's rows
    // are contiguous in the original numbering (the sweep then skips perm),
    // else kNoOffset
    std::vector<int64_t> level_off;
    static constexpr int64_t kNoOffset = INT64_MIN;
    // wide levels (>= kStreamLevelRows rows) are swept in row blocks of at
    // most kSpmvRowsBlock rows / kSpmvNnzBlock dependencies (positions
    // blocks[j]..blocks[j+1]); level l owns blocks [level_blk[2l],
    // level_blk[2l+1]) (host), an empty range for a narrow level
    uint32_t *blocks = nullptr;
    BlockDesc *desc = nullptr;  // [nblk + 1] descriptors of those blocks
    std::vector<uint32_t> level_blk;
    // 16-bit dependency offsets of the row blocks (as for the SpMV): block
    // j's dependencies are blk_base[j] + ci16[e] unless blk_base[j] == kNoBase
    uint16_t *ci16 = nullptr;
    uint32_t *blk_base = nullptr;
    uint64_t nnz16 = 0, nblk = 0, perm_rows = 0;  // entries on 16-bit offsets, row blocks, rows read through perm
    int variant = kVarDefault;                     // kVarNT | kVarIdx16 for the staged kernel
    uint32_t max_level_rows = 0;
    // the index arrays (perm, rp, ci, blocks, desc, ci16, blk_base) belong to
    // another sweep of the same rows (the FP32 sweep shares the FP64 one's)
    bool borrowed = false;
    // sliced-column form of the wide, contiguous levels (variant kVarEll):
    // row t of level l is original row row0 + t; its entry j at [j * nrows
    // + t] with dependency column row0 + t + cbase + off (16-bit offset),
    // padded with zeros on the row's last dependency
    struct EllLevel {
        bool ok = false;
        uint32_t nrows = 0, width = 0;
        uint64_t entries = 0;  // the level's dependencies (unpadded)
        int64_t row0 = 0, cbase = 0;
        int16_t *off = nullptr;  // shared with the FP32 sweep (borrowed)
        T *val = nullptr;
    };
    std::vector<EllLevel> ell;
    int ell_fallback = kVarDefault;  // the staged variant of the other levels when variant == kVarEll
    void free_all();
};

// PP_TRSV_BLOCK: the rows split into contiguous chunks that no dependency of
// either sweep crosses, each small enough for its part of the vector to sit
// in LDS.  One workgroup per chunk runs both sweeps over its rows, level by
// level with a workgroup barrier between levels, so a dependency chain of L
// levels costs L barriers instead of L launches, and independent diagonal
// blocks (the subdomains of a multilevel preconditioner) run side by side.
struct ChunkSet {
    std::vector<uint32_t> start;   // host [nchunk+1] first original row of each chunk
    uint32_t *dstart = nullptr;    // the same on the device
    uint32_t nchunk = 0, max_rows = 0;
    void free_all();
};
// Rows are stored by position (chunk-major, level order inside a chunk), the
// entries of consecutive positions contiguous, so the entries of a level are
// one contiguous range: the workgroup loads them coalesced, forms the
// products in LDS, and each row's thread subtracts its products in storage
// order.  Row and column numbers are local to the chunk (16 bits).
template <class T>
struct DevChunkSweep {
    uint32_t n = 0, nnz = 0, max_levels = 0, max_level_nnz = 0;
    uint16_t *row = nullptr;       // [n] position -> row minus the chunk's first row
    uint32_t *rp = nullptr;        // [n+1] entries by position
    uint16_t *ci = nullptr;        // [nnz] dependency row minus the chunk's first row
    T *val = nullptr;              // [nnz]
    T *pdiag = nullptr;            // [n] diagonal by position, or nullptr (unit sweep)
    uint32_t *lp = nullptr;        // level boundaries (positions); chunk c: lp[lev_off[c] .. lev_off[c+1]-1]
    uint32_t *le = nullptr;        // rp at those boundaries (first entry of each level)
    uint32_t *lev_off = nullptr;   // [nchunk+1]
    void free_all();
};

// Host gather-form sweep: row i = s_i -> (b_i - sum val*x[ci]) op diag_i.
// Symbolic state of a numeric re-factorization on the device
// (pp_precond_refactor, IC(0) / ILU(0)): the factor's pattern, its
// dependency levels, and for every value slot of the sweeps (level-ordered
// storage, pivots) the factor entry it holds.
struct DevRefactor {
    uint32_t n = 0, nnz = 0;       // factor entries: IC(0) L's strict lower part; ILU(0) the matrix pattern
    uint32_t *rp = nullptr, *ci = nullptr;
    uint32_t *src = nullptr;       // IC(0): the matrix entry each L entry starts from
    uint32_t *dsrc = nullptr;      // IC(0): the diagonal entry of each row; ILU(0): its position in the row
    double *val = nullptr;         // [nnz + n]: the factor entries, then the pivots (IC(0) l_ii; ILU(0) 1/u_ii)
    uint32_t *rows = nullptr;      // rows in level order
    std::vector<uint32_t> lptr;    // level bounds in rows
    uint32_t *lo_map = nullptr, *lo_dmap = nullptr, *up_map = nullptr, *up_dmap = nullptr;
    uint32_t lo_nnz = 0, up_nnz = 0;
    bool lo_diag = false, up_diag = false;
    uint32_t *aperm_map = nullptr; // entry k of P A P^T <- entry aperm_map[k] of A (ordered factors)
    uint32_t aperm_nnz = 0;
    int *fail = nullptr;
    // host: which factor entry (index into val) each entry of the host copy holds
    std::vector<uint32_t> h_lt_map, h_l_map, h_u_map;
    void free_all();
    ~DevRefactor() { free_all(); }
};

struct HostSweep {
    enum Mode { DIVIDE = 0, UNIT = 1, MULTIPLY = 2, PREDIVIDE = 3 };
    int mode = DIVIDE;
    bool lower = true;
    uint32_t n = 0;
    std::vector<uint32_t> rp, ci;
    std::vector<double> val, diag;
    // dependency levels, computed once: rows in level order, level bounds
    mutable std::vector<uint32_t> lev_perm, lev_ptr;
};

// A row-per-thread sliced-column form of a sparse matrix (the reduced CG's
// KA and KB; the SpMV variant kVarEll of a CSR matrix): entry j of row i at
// [j * n + i] (every load coalesced), its column i + cbase + off (16-bit
// offset), rows padded to `width` with zero values on the row's last
// column; an optional diagonal term dg[i] * x[dbase + i] is added last.
// Built only when every row has an entry, every offset fits 16 bits and the
// padding stays small.
struct EllOp {
    bool ok = false;
    uint32_t n = 0, width = 0;
    int64_t cbase = 0, dbase = 0;
    // two-colour forms (ell_build_two_colour): rows from csplit on take their
    // columns from row + cbase1 instead of row + cbase, and rows below dsplit
    // add their diagonal (dg) first instead of last
    uint32_t csplit = UINT32_MAX, dsplit = 0;
    int64_t cbase1 = 0;
    int16_t *off = nullptr;      // [width * n], or nullptr when the rows use patterns:
    uint8_t *pid = nullptr;      // [n] the row's offset pattern (stencil-like matrices:
    int16_t *ptab = nullptr;     //     at most 256 distinct offset rows) and [npat * 8] the patterns
    uint32_t npat = 0;
    double *v64 = nullptr;       // [width * n]
    float *v32 = nullptr;
    double *dg64 = nullptr;      // [n] or nullptr
    float *dg32 = nullptr;
    void free_all();
};

}  // namespace pp

struct pp_csr {
    pp_ctx *ctx = nullptr;
    uint32_t nrow = 0, ncol = 0, nnz = 0;
    uint32_t *rp = nullptr, *ci = nullptr;
    double *v64 = nullptr;
    float *v32 = nullptr;
    uint32_t *row_blocks = nullptr;  // adaptive SpMV row blocks
    uint32_t nblocks = 0;
    pp::BlockDesc *blk_desc = nullptr;  // [nblocks + 1] descriptors of the row blocks
    // 16-bit column indices relative to a per-block base: block b's columns
    // are blk_base[b] + ci16[k]; blk_base[b] == kNoBase marks a block whose
    // column span does not fit 16 bits (it reads ci).  NULL when no block fits.
    uint16_t *ci16 = nullptr;
    uint32_t *blk_base = nullptr;
    uint32_t nblocks16 = 0;          // blocks that use ci16
    int spmv_variant = pp::kVarDefault;    // FP64 loads, chosen by timing at creation for large matrices
    int spmv_variant32 = pp::kVarDefault;  // FP32 loads, chosen by timing when the FP32 copy is made
    bool tuned32 = false;
    pp::EllOp ell;                   // sliced-column copy (SpMV variant kVarEll) when the rows fit it
    // time the load variants with warm caches: the matrix runs in an
    // iteration whose operands stay cache-resident between launches
    bool tune_warm = false;
    uint64_t nnz16 = 0;              // entries in those blocks
    pp_csr *trans = nullptr;         // explicit transpose, built on demand
    std::vector<uint32_t> h_rp, h_ci;
    std::vector<double> h_v;
    bool f32_dirty = true;
    uint64_t version = 0;            // bumped by every value update
    uint64_t uid = 0;                // unique per object (never reused)
};

namespace pp {
// IC(0) of a two-colour matrix in Schur-complement form.  When the rows
// split into colours [0, n0) and [n0, n) whose diagonal blocks are diagonal
// (a 5- or 7-point stencil in red-black order), the IC(0) factor is
// L = [D0^1/2, 0; A10 D0^-1/2, Dt^1/2] with Dt = diag(A11 - A10 D0^-1 A01),
// and the split-preconditioned operator L^-1 A L^-T is block-diagonal:
// the identity on the first colour and Sb = Dt^-1/2 (A11 - A10 D0^-1 A01)
// Dt^-1/2 on the second (Eisenstat, SIAM J. Sci. Stat. Comput. 2 (1981) 1-4:
// the preconditioner fused with the operator).  CG on Sb therefore runs the
// same Krylov iteration as IC(0)-PCG on A, on half the unknowns and without
// any triangular sweep: one operator application is y0 = B v (first-colour
// rows) and w = c * v - Bt y0 (second-colour rows), two SpMVs over the
// off-diagonal entries of A.  All four operators are stored as CSR over the
// full column range [0, n) so that they read one vector z = [y0 or r0; v]:
//   KA  rows [0, n0):  B_ij  = a_ij / sqrt(a_ii Dt_j)          (j >= n0)
//   KB  rows [n0, n):  -a_ij / sqrt(a_jj Dt_i) (j < n0), c_i = a_ii / Dt_i at j = i
//   KF  rows [n0, n):  -a_ij / (a_jj sqrt(Dt_i)) (j < n0), 1 / sqrt(Dt_i) at j = i
//                      (f = Dt^-1/2 (r1 - A10 D0^-1 r0), the reduced right-hand side)
//   KE  rows [0, n0):  1 / a_ii at j = i, -a_ij / a_ii (j >= n0)
//                      (e0 = D0^-1 (r0 - A01 e1), the eliminated unknowns)
// and s1 = Dt^-1/2 maps the reduced solution back: e1 = s1 * x.
struct ReducedForm {
    bool ok = false;  // false: the matrix / factor does not have this structure
    uint32_t n0 = 0, n1 = 0;
    pp_csr *KA = nullptr, *KB = nullptr, *KF = nullptr, *KE = nullptr;
    double *s1_64 = nullptr;  // Dt^-1/2
    float *s1_32 = nullptr;
    double *dt_64 = nullptr;  // Dt: the true residual of the second colour is Dt^1/2 times the reduced one
    float *dt_32 = nullptr;
    EllOp ka, kb;                 // KA and KB in sliced-column form (when they fit)
    double anorm = 0.0;           // max absolute row sum of the colour-ordered matrix
    uint64_t src_uid = 0;  // built from the matrix of this uid at src_version
    uint64_t src_version = 0;
    int *dfail = nullptr;     // device word of reduced_refresh's pivot test
    double *dwork = nullptr;  // its diagonal of the matrix (n)
    int built = 0;            // 1 host construction, 2 device, 3 refreshed on the device
    float build_ms = 0.f;     // that construction's wall time
};
// Host arrays of the reduced form of one colour-ordered matrix: KA (n0
// rows), KB and KF (n1 rows), KE (n0 rows), columns in the matrix's
// numbering; s1 = Dt^-1/2 and Dt over the second colour.
struct ReducedHost {
    uint32_t n0 = 0, n1 = 0;
    std::vector<uint32_t> ka_rp, ka_ci, kb_rp, kb_ci, kf_rp, kf_ci, ke_rp, ke_ci;
    std::vector<double> ka_v, kb_v, kf_v, ke_v, s1, dt;
    double anorm = 0.0;
};
// false when the matrix has no two-colour split with diagonal colour
// blocks, or a pivot of its IC(0) is not positive
bool reduced_host(uint32_t n, const std::vector<uint32_t> &rp, const std::vector<uint32_t> &ci,
                  const std::vector<double> &av, ReducedHost &H);
// The reduced form of M for the colour-ordered matrix As (built once and
// cached in M), or nullptr when it does not apply.
ReducedForm *reduced_form(pp_ctx *ctx, pp_precond *M, pp_csr *As);
// The reduced form built on the device from As's device arrays (R empty on
// entry): 1 built, 0 the structure is absent or a pivot is not positive
// (R->ok stays false), -1 As has no two-colour sliced-column copy of the
// operators' width (the host construction applies).
int reduced_build_device(pp_ctx *ctx, pp_precond *M, pp_csr *As, ReducedForm *R);
// New values of an existing reduced form from As's current values (same
// pattern as it was built from), on the device, bit-identical to a host
// rebuild; false when a pivot is not positive (pp_reduced_dev.hip).
bool reduced_refresh(pp_ctx *ctx, ReducedForm *R, pp_csr *As);
void reduced_need_f32(pp_ctx *ctx, ReducedForm *R);
void reduced_free(pp_precond *M);
}  // namespace pp

struct pp_precond {
    pp_ctx *ctx = nullptr;
    uint64_t uid = 0;             // unique per object (never reused)
    pp_csr *A = nullptr;
    pp_precond_opts opts{};
    int ready = 0;
    uint32_t n = 0, mnnz = 0, fail_row = 0;
    double setup_ms = 0;
    // host factor (reference layouts)
    std::vector<uint32_t> lt_rp, lt_ci, l_rp, l_ci, u_rp, u_ci;
    std::vector<double> lt_v, l_v, u_v, d;
    // device sweeps
    pp::HostSweep hlow, hup;
    bool has_diag_stage = false;  // ILDLT: D^{-1} between the sweeps (folded into upper)
    pp::DevSweep<double> low64, up64;
    pp::DevSweep<float> low32, up32;
    bool have32 = false;
    int trsv = PP_TRSV_LEVELS;    // resolved SpTRSV schedule
    bool fuse_mid = false;        // last forward level == first backward level (fused launch)
    // the forward sweep's first level is the contiguous original rows
    // [lvl0_lo, lvl0_hi) (no dependencies): a caller that produces the
    // right-hand side can also produce those rows of the result
    bool lvl0_contiguous = false;
    uint32_t lvl0_lo = 0, lvl0_hi = 0;
    int sweep_variant = pp::kVarDefault;  // staged-sweep load variant (timed at setup)
    uint32_t fuse_rows = 0;
    std::vector<uint32_t> order;  // factor row k = original row order[k] (empty: natural)
    uint32_t ncolors = 0;
    uint32_t *dorder = nullptr;
    double *xp64 = nullptr;       // permuted work vectors
    float *xp32 = nullptr;
    double *tmp64 = nullptr;      // out-of-place buffer between the two sweeps
    float *tmp32 = nullptr;
    void *multilevel = nullptr;   // AMSED / GMSLR state
    // P A P^T in the factor's row order, for solves that run in that order
    // (no gather/scatter around each apply); built from the matrix of uid Aperm_src_uid at
    // Aperm_version.  io_permuted: the current solve's vectors are already in
    // the factor's order.
    pp_csr *Aperm = nullptr;
    // numeric re-factorization on the device (built by the first pp_precond_refactor)
    pp::DevRefactor *dref = nullptr;
    uint32_t device_refactors = 0;
    bool host_factor_stale = false;  // the host copy of the factor predates a device re-factorization
    // set up on the device: the host structures (factor, sweeps, P A P^T's
    // columns and values) are built only when a caller needs them
    bool host_pending = false;
    bool dev_setup = false;  // the last setup ran on the device
    uint64_t Aperm_src_uid = 0;
    uint64_t Aperm_version = 0;
    bool io_permuted = false;
    pp::ReducedForm *reduced = nullptr;  // two-colour Schur-complement form (IC(0) only)
    bool sweep_pinned = false;           // pp_precond_set_sweep_variant: the FP32 sweeps are not timed
    float phase_ms[4] = {0, 0, 0, 0};     // setup: ordering, factorization, analysis + upload, schedule
    // PP_TRSV_BLOCK storage (built when the rows chunk, see ChunkSet)
    pp::ChunkSet chunks;
    pp::DevChunkSweep<double> clow64, cup64;
    pp::DevChunkSweep<float> clow32, cup32;
};

namespace pp {

// ---- core helpers (pp_core.cpp) ----
uint64_t next_uid();
void set_device(pp_ctx *ctx);
template <class T>
T *dalloc(size_t n);
void dfree(void *p);
// Inside a PoolScope, dalloc first takes a block of the context's pool that
// is at most a quarter larger than asked (after synchronizing the stream
// once blocks were returned, so queued work is done with them).  pool_put
// returns a block whose last users are queued on ctx->stream; beyond
// kPoolCap bytes it is freed instead.  hipFree costs 0.1-0.3 ms per
// megabyte-sized block, the setup's temporaries number dozens.
struct PoolScope {
    pp_ctx *prev;
    explicit PoolScope(pp_ctx *ctx);
    ~PoolScope();
};
constexpr size_t kPoolCap = size_t(4) << 30;
void pool_put(pp_ctx *ctx, void *p, size_t bytes);
void pool_release(pp_ctx *ctx);
void ensure_hist(pp_ctx *ctx, uint32_t cap);
void *ws_get(pp_ctx *ctx, int idx, size_t bytes);
// ---- zero-allocation solves (pp_noalloc.hip) ----
// size of the fixed device array the work vectors of pp_solve_noalloc live in
constexpr size_t kNoallocScratchBytes = (size_t)128 << 20;
// looks up the array's address on `device` (which must be current); once per device
void noalloc_setup(int device);
// that address, or nullptr before noalloc_setup
void *noalloc_scratch(int device);
// Workspace slots owned by the GMRES / Gram-Schmidt helpers; the solvers'
// own vectors use the slots below kWsGmresState and from kWsFree on.
constexpr int kWsGmresState = 30, kWsGmresCoeff = 31, kWsFree = 32;

// ---- host factorizations (pp_factor.cpp) ----
struct HostCsr {
    uint32_t n = 0;
    std::vector<uint32_t> rp, ci;
    std::vector<double> v;
};
int factor_ict(const HostCsr &A, double tau, bool ldlt, HostCsr &lt, std::vector<double> &d,
               uint32_t &fail_row);
int factor_ilut(const HostCsr &A, uint32_t lfil, double tau, HostCsr &L, HostCsr &U,
                std::vector<double> &D, uint32_t &fail_row);
// lower (optional): the forward sweep of L, filled when A's rows are sorted
int factor_ic0(const HostCsr &A, HostCsr &lt, uint32_t &fail_row, HostSweep *lower = nullptr);
int factor_ilu0(const HostCsr &A, HostCsr &L, HostCsr &U, std::vector<double> &D, uint32_t &fail_row);
HostCsr transpose(const HostCsr &A);
std::vector<uint32_t> multicolor_order(const HostCsr &A, uint32_t &ncolors);
HostCsr permute_symmetric(const HostCsr &A, const std::vector<uint32_t> &order);
pp_csr *precond_permuted_matrix(pp_ctx *ctx, pp_precond *M, pp_csr *A);
void precond_apply(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y);
bool precond_level0_fusable(const pp_precond *M);
void precond_level0(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y);
void precond_apply_after_level0(pp_ctx *ctx, pp_precond *M, int dtype, const void *x, void *y);
void precond_need_f32(pp_precond *M);
void csr_need_f32(pp_ctx *ctx, pp_csr *A);
// device CSR of host arrays (copied / taken over); tune: time the SpMV load
// variants on it, for matrices in an iteration's hot loop (0 no, 1 with
// cold caches, 2 with warm caches: an iteration whose operands stay resident)
pp_csr *csr_new(pp_ctx *ctx, uint32_t nrow, uint32_t ncol, const uint32_t *rp, const uint32_t *ci, const double *v);
pp_csr *csr_from_vectors(pp_ctx *ctx, uint32_t nrow, uint32_t ncol, std::vector<uint32_t> &&rp,
                         std::vector<uint32_t> &&ci, std::vector<double> &&v, int tune);
pp_csr *csr_create_host(pp_ctx *ctx, HostCsr &&H);
void csr_free(pp_csr *A);

// ---- tall-skinny GEMVs and GMRES helpers (pp_kernels_krylov.hip) ----
// device coefficient buffer: [0,320) h, [320,640) y or a copy of h, then partials
double *gmres_coeff_buffer(pp_ctx *ctx);
// dest[c] = sum_r V[r, c] w[r] for c < ncol (dest NULL: the coefficient
// buffer).  V and w in FP64 or FP32 (the mixed GMRES basis); sums in FP64.
template <class TV>
void launch_gemv_t(pp_ctx *ctx, uint32_t n, uint32_t ncol, const TV *V, size_t ld, const TV *w, double *dest);
template <class TV>
void launch_gemv_sub(pp_ctx *ctx, uint32_t n, uint32_t ncol, const TV *V, size_t ld, const double *h, TV *w);
// w -= V h1 followed by dest = V^T w, in one pass over V where it can
template <class TV>
void launch_gemv_sub_t(pp_ctx *ctx, uint32_t n, uint32_t ncol, const TV *V, size_t ld, const double *h1, TV *w,
                       double *dest);
void launch_gemv_add(pp_ctx *ctx, uint32_t n, uint32_t ncol, const float *Z32, const double *Z64, size_t ld,
                     const double *y, double *x);
void read_slots(pp_ctx *ctx, int first, int count, double *out);
// Device-resident Givens least-squares of FlexibleGMRES(m): H ((m+1) x m,
// column-major), c, s, g, the two CGS passes' coefficients h1, h2, and
// scalars misc[kGmMisc]
struct GmresDev {
    double *H = nullptr, *c = nullptr, *s = nullptr, *g = nullptr, *h1 = nullptr, *h2 = nullptr, *misc = nullptr;
    uint32_t mm = 0;
};
// GM_HBASE: the residual-history index of the cycle's first step
enum GmMisc {
    GM_INVH = 0, GM_RELERR, GM_RNORM, GM_M, GM_STATUS, GM_CONVCYC, GM_BNORM, GM_TOL, GM_NREC, GM_HBASE, kGmMisc = 16
};
GmresDev gmres_state(pp_ctx *ctx, uint32_t mm);
void launch_gmres_cycle_init(pp_ctx *ctx, const GmresDev &st, double beta, double bnorm, double tol,
                             uint32_t hist_base);
void launch_gmres_step(pp_ctx *ctx, const GmresDev &st, uint32_t j, int quirks, double *hist, uint32_t hist_idx,
                       uint32_t hist_cap);
template <class TV>
void launch_gmres_next_basis(pp_ctx *ctx, const GmresDev &st, uint32_t n, const TV *w, TV *vnext);
void launch_gmres_backsolve(pp_ctx *ctx, const GmresDev &st, int quirks, double *y);
void launch_permute(pp_ctx *ctx, uint32_t n, const uint32_t *perm, int dtype, const void *src, void *dst,
                    bool gather);

// ---- level analysis + upload (pp_precond.cpp) ----
template <class T>
void upload_sweep(const HostSweep &s, DevSweep<T> &d);

// ---- kernels (pp_kernels.hip) ----
// y = alpha A x + beta y; dot_slot >= 0: also (dot_with, y) into that slot;
// aux != NULL (with dot_slot): also (aux, dot_with) and (aux, aux) into the
// next two slots, left as partials for the caller's finish kernel; wgt !=
// NULL weights the last: sum wgt_i aux_i^2
void launch_spmv(pp_ctx *ctx, const pp_csr *A, int dtype, double alpha, const void *x,
                 double beta, void *y, int dot_slot, const void *dot_with, bool dot_first,
                 const void *aux = nullptr, const void *wgt = nullptr);
void launch_fill_zero(pp_ctx *ctx, void *p, size_t bytes);
// dst = src, `bytes` bytes, 16-byte aligned pointers; groups 0: 64 per CU
void launch_copy_stream(pp_ctx *ctx, const void *src, void *dst, size_t bytes, unsigned groups);
// ---- numeric IC(0) / ILU(0) on the device (pp_devfactor.hip) ----
void launch_ic0_level(pp_ctx *ctx, const uint32_t *rows, uint32_t nrows, const uint32_t *rp, const uint32_t *ci,
                      const uint32_t *src, const uint32_t *dsrc, const double *av, double *lv, double *diag,
                      int *fail);
void launch_ilu0_level(pp_ctx *ctx, const uint32_t *rows, uint32_t nrows, const uint32_t *rp, const uint32_t *ci,
                       const uint32_t *dpos, const double *av, double *w, double *dinv, int *fail);
// out[q] = in[map[q]] (dtype of out)
void launch_gather_map(pp_ctx *ctx, uint32_t n, const uint32_t *map, const double *in, int dtype, void *out);
// out[q] = (uint32_t)in[q]
void launch_to_index(pp_ctx *ctx, uint32_t n, const double *in, uint32_t *out);
// Average duration (ms) of fn()'s launches on ctx->stream with the last-level
// cache cleared before each one (a read of `flush_bytes` of other memory), so
// that operands are streamed from HBM as they are inside a solve.
// shortest time of fn over reps launches after one warm-up (timing noise
// only ever adds); cold: the caches (L2 and the last-level MALL) flushed
// before each
float time_cold_ms(pp_ctx *ctx, const std::function<void()> &fn, int reps, bool cold = true);
// Host setup work split over threads: fn(begin, end) on contiguous ranges of
// [0, n), at most 16 threads and at least `grain` items each (small n runs
// on the calling thread).  fn must only write what its range owns.
void parallel_for(size_t n, size_t grain, const std::function<void(size_t, size_t)> &fn);
// 16-bit two-base offsets for every row block of a CSR in one pass (blocks
// blk[j]..blk[j+1] of positions into rp, columns ci): off16 per entry, base
// 2 per block (kNoBase where a block does not fit); returns the entries on
// 16-bit offsets.
uint64_t two_base_blocks(const std::vector<uint32_t> &blk, const std::vector<uint32_t> &rp, const uint32_t *ci,
                         std::vector<uint16_t> &off16, std::vector<uint32_t> &base);
// Host-built descriptors of row blocks blk[j]..blk[j+1] (positions into rp);
// base[2j], base[2j+1] the 16-bit offset bases or kNoBase.
std::vector<BlockDesc> block_descs(const std::vector<uint32_t> &blk, const std::vector<uint32_t> &rp,
                                   const std::vector<uint32_t> &base);
void launch_convert(pp_ctx *ctx, uint32_t n, int sdt, const void *s, int ddt, void *d);
void launch_axpby(pp_ctx *ctx, int dtype, uint32_t n, double a, const void *x, double b,
                  const void *y, void *z);
void launch_dot(pp_ctx *ctx, int dtype, uint32_t n, const void *x, const void *y, int slot);
// y = E x (EllOp); dot_slot >= 0: also the three sums of the single-reduction
// recurrence into that slot and the next two, as launch_spmv does with dotw,
// aux and wgt.  ell_grid: the workgroups (partials) such a launch uses.
void launch_ell(pp_ctx *ctx, const EllOp &E, int dtype, double alpha, const void *x, double beta, void *y,
                int dot_slot, const void *dotw, const void *aux, const void *wgt, unsigned grid = 0);
unsigned ell_grid(const EllOp &E, bool dot);
template <class T>
void ell_fill_rows(uint32_t n, uint32_t width, const uint32_t *rp, const T *v, T *out);
// E's values (dtype) from the CSR values v of the matrix it was built from
// (a form with a separate diagonal takes each row's last entry from dg, not
// from the slots)
void ell_fill_values(pp_ctx *ctx, const EllOp &E, int dtype, const uint32_t *rp, const void *v);
// host form -> EllOp (FP64 values on the device); false (nothing kept) when
// a row is empty or an offset does not fit 16 bits
bool ell_build(const std::vector<uint32_t> &rp, const std::vector<uint32_t> &ci, const std::vector<double> &v,
               int64_t cbase, bool diag_last, int64_t dbase, EllOp &E, double max_fill = 1.25);
// The same forms built on the device from A's device arrays (A->h_rp on the
// host for the row-length conditions): the natural form (no base, no
// separate diagonal) or, two_colour, the two-colour form split at `split`;
// the result equals the host construction's, pattern numbers included.
bool ell_build_device(pp_ctx *ctx, const pp_csr *A, bool two_colour, uint32_t split, EllOp &E, double max_fill);
// The sliced-column form of a two-colour matrix (rows [0, split) couple only
// to [split, n) and back, every row holding its diagonal; a colour-ordered
// stencil): per-colour column bases, the diagonal in dg (first in the rows'
// sums below split, last from split on, as the CSR rows order it).
bool ell_build_two_colour(const std::vector<uint32_t> &rp, const std::vector<uint32_t> &ci,
                          const std::vector<double> &v, uint32_t split, EllOp &E, double max_fill = 1.1);
void launch_outer_on_pattern(pp_ctx *ctx, const pp_csr *A, double scale, const double *lam,
                             const double *x, double *g);
void build_row_blocks(const std::vector<uint32_t> &rp, std::vector<uint32_t> &blocks);
// The SpMV load variants of A timed now (FP64; FP32 at its first use), as
// for a matrix created with timing.
void csr_tune(pp_ctx *ctx, pp_csr *A);
// The host copies of A's columns and values, downloaded when A was built on
// the device without them.
void csr_ensure_host(pp_csr *A);
// The 16-bit two-base column offsets of nblk row blocks (device arrays:
// block bounds, row pointers, columns) computed on the device, as
// two_base_blocks does on the host: the entries in those blocks, off16
// [nnz] and dbase [2 nblk] on the device (both null when no block fits) and
// the bases on the host.
uint64_t two_base_device(pp_ctx *ctx, uint32_t nblk, const uint32_t *blocks, const uint32_t *rp, const uint32_t *ci,
                         uint32_t nnz, uint16_t *&off16, uint32_t *&dbase, std::vector<uint32_t> &base);
// IC(0) or ILU(0) in multicolour order set up on the device
// (pp_setup_dev.hip): the order, P A P^T, the factor, both FP64 sweeps and
// the re-factorization state, the host's results without its host copies;
// false (M unchanged) when the matrix is outside that path (unsorted or
// unsymmetric pattern, long rows, more than 64 colours, a failing pivot).
bool setup_mc_device(pp_ctx *ctx, pp_precond *M);
// The CSR object of device arrays it takes over (freed with it, also when
// this throws), with the host copy of the row pointers: row blocks and
// their 16-bit column form for the CSR-stream SpMV (computed on the
// device), no variant timing, no host copy of the columns or values.
pp_csr *csr_from_device(pp_ctx *ctx, uint32_t nrow, uint32_t ncol, uint32_t *rp, uint32_t *ci, double *v,
                        std::vector<uint32_t> &&h_rp);

template <class T>
void launch_sweep(pp_ctx *ctx, const DevSweep<T> &s, int mode, const T *b, T *x);
// Both sweeps of an apply in one launch, one workgroup per chunk
// (PP_TRSV_BLOCK; DevChunkSweep).  y = U^{-1} L^{-1} x, y may alias x.
template <class T>
void launch_sweep_pair_chunks(pp_ctx *ctx, const ChunkSet &C, const DevChunkSweep<T> &lo, int mlo,
                              const DevChunkSweep<T> &up, int mup, const T *x, T *y);
// The same sweeps, one wavefront computing the levels while the rest of the
// workgroup stages their data in LDS (pp_sweep_pipe.hip); sweep_pipe_rings:
// the LDS ring sizes, false when the chunks leave too little LDS for them.
template <class T>
void launch_sweep_pair_pipe(pp_ctx *ctx, const ChunkSet &C, const DevChunkSweep<T> &lo, int mlo,
                            const DevChunkSweep<T> &up, int mup, const T *x, T *y);
template <class T>
bool sweep_pipe_rings(pp_ctx *ctx, const ChunkSet &C, uint32_t &re, uint32_t &rr);
template <class T>
void launch_sweep_pair(pp_ctx *ctx, const DevSweep<T> &lo, int mlo, const DevSweep<T> &up, int mup, bool fuse,
                       const T *x, T *y, uint32_t first_level = 0);
// y[i] = x[i] op diag[i] over [lo, hi) (the dependency-free first level)
template <class T>
void launch_level0(pp_ctx *ctx, uint32_t lo, uint32_t hi, int mode, const T *diag, const T *x, T *y);
template <class T>
void launch_sweep_syncfree(pp_ctx *ctx, const DevSweep<T> &s, int mode, const T *b, T *y);

// PCG device steps (templated on the working precision)
template <class T>
void pcg_d_update(pp_ctx *ctx, uint32_t n, bool first, const T *w, T *d, const T *r);
template <class T>
void pcg_xr_update(pp_ctx *ctx, uint32_t n, T *x, T *r, const T *d, const T *v, int hist_on, bool publish);
template <class T>
void pcg_wv(pp_ctx *ctx, uint32_t n, const T *w, const T *v);
// single-reduction PCG: finish of the three SpMV reductions (+ check), update
void cgsr_finish(pp_ctx *ctx, unsigned np, int hist_on, bool publish);
template <class T>
void red_update(pp_ctx *ctx, uint32_t n, const T *w, T *p, T *s, T *x, T *r);
// One launch that finishes the three reductions of the reduced CG (every
// workgroup sums the np partials in the same order, so all hold the same
// alpha and beta), runs the convergence test (the first workgroup keeps the
// control words, the S_CG_* slots and the ring publication when `publish`)
// and applies the update; parity selects the state block read (the other
// is written).  Iteration bodies hold an even number of these.
template <class T>
void red_step(pp_ctx *ctx, unsigned np, int parity, bool publish, uint32_t n, const T *w, T *p, T *s, T *x, T *r);
// The reduced-CG iteration in two launches (k_red_ka, k_red_kb): red_ka =
// the previous iteration's check and step lengths, then y0 = KA r_new with
// the residual update folded into the gathers; red_kb = the update of p, s,
// x, r in place, then w = KB [y0; r_new] with the three inner products (grid:
// the partial count, as the DOT sliced-column launch).  Available when KA
// and KB have exact-width (4-7) sliced-column forms (red_fused_ok).
bool red_fused_ok(const EllOp &ka, const EllOp &kb, int dtype);
template <class T>
void red_ka(pp_ctx *ctx, const EllOp &E, const T *z, T *y, uint32_t rbase, const T *w, const T *s, unsigned np,
            int parity, bool publish);
template <class T>
void red_kb(pp_ctx *ctx, const EllOp &E, T *z, uint32_t rbase, T *w, T *p, T *s, T *x, const T *wt, unsigned grid);
// State block 0 for a fresh reduced solve, or (resume) for one that
// continues from S_CG_GOLD / S_CG_ALPHA / S_CG_FIRST, gamma scaled by gscale.
void red_init(pp_ctx *ctx, bool resume, double gscale, bool last);
// ctl words, tolerance, ||rin||^2 (bb >= 0) and iteration state of a
// reduced solve, then x = 0 and p, s rescaled (resume) or zeroed: two launches
void red_begin(pp_ctx *ctx, uint32_t maxit, double tol, double bb, bool resume, double scale, int dtype,
               uint32_t n1, void *x, void *p, void *s);
template <class T>
void red_post(pp_ctx *ctx, uint32_t n0, uint32_t n1, const T *x, const T *s1, const T *rin, T *z, T *e);
// lvl0_hi > lvl0_lo: also u[i] = r_new[i] op diag0[i] over [lvl0_lo, lvl0_hi)
// (the next preconditioner apply's first level)
template <class T>
void cgsr_update(pp_ctx *ctx, uint32_t n, T *u, const T *w, T *p, T *s, T *x, T *r, uint32_t lvl0_lo = 0,
                 uint32_t lvl0_hi = 0, int mode0 = 1, const T *diag0 = nullptr);
unsigned spmv_grid(const pp_csr *A);
void set_slot(pp_ctx *ctx, int slot, double v);
void scale_slot(pp_ctx *ctx, int slot, double v);  // scal[slot] *= v
void ctl_init(pp_ctx *ctx, int first_check, int last, uint32_t maxit = 0);
void copy_slot(pp_ctx *ctx, int dst, int src);

// Reads slot s to the host (synchronizes the stream).
double read_slot(pp_ctx *ctx, int slot);
void read_ctl(pp_ctx *ctx, int *out);
void write_ctl(pp_ctx *ctx, const int *vals);

}  // namespace pp
