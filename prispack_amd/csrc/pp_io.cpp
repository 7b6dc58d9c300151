// pp_io.cpp -- host-side input and ordering that sit either side of the
// solver path: MatrixMarket files into CSR (HOST::ReadMtxTools,
// source/utils/TestTools/ReadMtxTools.cpp) and the reverse Cuthill-McKee
// ordering (HOST::rcmReorderCSR, include/utils/ExternalTools/MatrixReorderTools.h:59-80,
// which calls the vendored genrcm of external_packages/RCM).
#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "pp_internal.h"

struct pp_mtx {
    uint32_t nrow = 0, ncol = 0;
    std::vector<uint32_t> rp, ci;
    std::vector<double> v;
};

namespace pp {
namespace {

std::string lower(std::string s) {
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

enum class Storage { General, Symmetric, Hermitian, Skew };

struct MtxHeader {
    bool coordinate = true;
    std::string field;
    Storage storage = Storage::General;
    uint32_t nrow = 0, ncol = 0;
    uint64_t nnz = 0;
};

// The banner is the '%' line whose first token is %%MatrixMarket; the first
// line not starting with '%' holds the sizes (ReadMtxTools.cpp:50-121).
MtxHeader read_header(std::istream &in, const std::string &path) {
    MtxHeader h;
    bool banner = false;
    std::string line;
    bool have_size_line = false;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] != '%') {
            if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
            have_size_line = true;
            break;
        }
        std::istringstream is(line);
        std::string tok;
        is >> tok;
        if (tok != "%%MatrixMarket") continue;
        banner = true;
        std::string object, format, field, symmetry;
        is >> object >> format >> field >> symmetry;
        PP_CHECK(lower(object) == "matrix", PP_ERR_UNSUPPORTED, path + ": not a MatrixMarket matrix");
        format = lower(format);
        PP_CHECK(format == "coordinate" || format == "array", PP_ERR_UNSUPPORTED,
                 path + ": unknown MatrixMarket format '" + format + "'");
        h.coordinate = format == "coordinate";
        h.field = lower(field);
        PP_CHECK(h.field == "real" || h.field == "complex" || h.field == "pattern" || h.field == "integer",
                 PP_ERR_UNSUPPORTED, path + ": unknown MatrixMarket field '" + h.field + "'");
        symmetry = lower(symmetry);
        if (symmetry == "general") h.storage = Storage::General;
        else if (symmetry == "symmetric") h.storage = Storage::Symmetric;
        else if (symmetry == "hermitian") h.storage = Storage::Hermitian;
        else if (symmetry == "skew-symmetric") h.storage = Storage::Skew;
        else throw Error(PP_ERR_UNSUPPORTED, path + ": unknown MatrixMarket symmetry '" + symmetry + "'");
    }
    PP_CHECK(banner, PP_ERR_INVALID, path + ": no %%MatrixMarket banner");
    PP_CHECK(have_size_line, PP_ERR_INVALID, path + ": no size line");
    std::istringstream is(line);
    long long r = -1, c = -1, z = -1;
    is >> r >> c;
    PP_CHECK(r >= 0 && c >= 0 && r <= UINT32_MAX && c <= UINT32_MAX, PP_ERR_INVALID, path + ": bad size line");
    h.nrow = (uint32_t)r;
    h.ncol = (uint32_t)c;
    if (h.coordinate) {
        is >> z;
        PP_CHECK(z >= 0, PP_ERR_INVALID, path + ": bad size line");
        h.nnz = (uint64_t)z;
    }
    return h;
}

}  // namespace

// Coordinate real file -> CSR, with the reference reader's rules
// (ReadMtxTools.cpp:187-230, 339-384): entries of magnitude <= 10 DBL_EPSILON
// are dropped; indices are shifted down by `base`; symmetric storage adds the
// mirrored entry and skew-symmetric storage the negated mirrored entry for
// every off-diagonal entry (hermitian storage of real data is taken as
// stored, as the reference does); entries are bucketed by row in file order
// and each row is then sorted by column with a stable insertion order
// (transCOO2CSR, MatrixTools.h:16-62), duplicates kept.
void read_mtx(const char *path, int base, pp_mtx &out) {
    PP_CHECK(path, PP_ERR_INVALID, "NULL path");
    PP_CHECK(base == 0 || base == 1, PP_ERR_INVALID, "base must be 0 or 1");
    std::ifstream in(path, std::ios::binary);
    PP_CHECK(in.is_open(), PP_ERR_INVALID, std::string("cannot open ") + path);
    const MtxHeader h = read_header(in, path);
    PP_CHECK(h.coordinate, PP_ERR_UNSUPPORTED, std::string(path) + ": a dense (array) file is not a sparse matrix");
    PP_CHECK(h.field == "real", PP_ERR_UNSUPPORTED,
             std::string(path) + ": only real matrices are read (field '" + h.field + "')");
    const bool mirror = h.storage == Storage::Symmetric || h.storage == Storage::Skew;
    std::vector<uint32_t> ri, cj;
    std::vector<double> vv;
    ri.reserve(h.nnz * (mirror ? 2 : 1));
    cj.reserve(ri.capacity());
    vv.reserve(ri.capacity());
    std::string line;
    uint64_t seen = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '%') continue;
        const char *p = line.c_str();
        char *end = nullptr;
        const long long r = std::strtoll(p, &end, 10);
        if (end == p) continue;  // a blank line
        p = end;
        const long long c = std::strtoll(p, &end, 10);
        PP_CHECK(end != p, PP_ERR_INVALID, std::string(path) + ": bad entry line '" + line + "'");
        p = end;
        const double v = std::strtod(p, &end);
        PP_CHECK(end != p, PP_ERR_INVALID, std::string(path) + ": bad entry line '" + line + "'");
        ++seen;
        PP_CHECK(seen <= h.nnz, PP_ERR_INVALID, std::string(path) + ": more entries than the size line states");
        if (std::fabs(v) <= DBL_EPSILON * 10) continue;
        const long long r0 = r - base, c0 = c - base;
        PP_CHECK(r0 >= 0 && c0 >= 0 && r0 < (long long)h.nrow && c0 < (long long)h.ncol, PP_ERR_INVALID,
                 std::string(path) + ": index out of range in '" + line + "'");
        ri.push_back((uint32_t)r0);
        cj.push_back((uint32_t)c0);
        vv.push_back(v);
        if (mirror && r0 != c0) {
            PP_CHECK(c0 < (long long)h.nrow && r0 < (long long)h.ncol, PP_ERR_INVALID,
                     std::string(path) + ": mirrored index out of range");
            ri.push_back((uint32_t)c0);
            cj.push_back((uint32_t)r0);
            vv.push_back(h.storage == Storage::Skew ? -1.0 * v : v);
        }
    }
    PP_CHECK(ri.size() <= UINT32_MAX, PP_ERR_INVALID, std::string(path) + ": more than 2^32 - 1 entries");
    const size_t nnz = ri.size();
    out.nrow = h.nrow;
    out.ncol = h.ncol;
    out.rp.assign((size_t)h.nrow + 1, 0);
    for (size_t k = 0; k < nnz; ++k) ++out.rp[ri[k] + 1];
    for (uint32_t i = 0; i < h.nrow; ++i) out.rp[i + 1] += out.rp[i];
    out.ci.resize(nnz);
    out.v.resize(nnz);
    std::vector<uint32_t> next(out.rp.begin(), out.rp.end() - 1);
    for (size_t k = 0; k < nnz; ++k) {
        const uint32_t q = next[ri[k]]++;
        out.ci[q] = cj[k];
        out.v[q] = vv[k];
    }
    std::vector<std::pair<uint32_t, double>> row;
    for (uint32_t i = 0; i < h.nrow; ++i) {
        const uint32_t a = out.rp[i], b = out.rp[i + 1];
        bool sorted = true;
        for (uint32_t q = a + 1; q < b && sorted; ++q) sorted = out.ci[q - 1] <= out.ci[q];
        if (sorted) continue;
        row.clear();
        for (uint32_t q = a; q < b; ++q) row.emplace_back(out.ci[q], out.v[q]);
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<uint32_t, double> &x, const std::pair<uint32_t, double> &y) {
                             return x.first < y.first;
                         });
        for (uint32_t q = a; q < b; ++q) {
            out.ci[q] = row[q - a].first;
            out.v[q] = row[q - a].second;
        }
    }
}

// Dense (array) real file, column-major (ReadMtxTools.cpp:260-290): nrow x ncol values.
void read_mtx_dense(const char *path, uint32_t &nrow, uint32_t &ncol, std::vector<double> *vals) {
    PP_CHECK(path, PP_ERR_INVALID, "NULL path");
    std::ifstream in(path, std::ios::binary);
    PP_CHECK(in.is_open(), PP_ERR_INVALID, std::string("cannot open ") + path);
    const MtxHeader h = read_header(in, path);
    PP_CHECK(!h.coordinate, PP_ERR_UNSUPPORTED, std::string(path) + ": a right-hand side file must be dense (array)");
    PP_CHECK(h.field == "real", PP_ERR_UNSUPPORTED, std::string(path) + ": only real right-hand sides are read");
    nrow = h.nrow;
    ncol = h.ncol;
    if (!vals) return;
    const uint64_t total = (uint64_t)h.nrow * h.ncol;
    vals->resize(total);
    for (uint64_t k = 0; k < total; ++k) {
        in >> (*vals)[k];
        PP_CHECK(!in.fail(), PP_ERR_INVALID, std::string(path) + ": fewer values than the size line states");
    }
}

// Reverse Cuthill-McKee of the graph of a structurally symmetric CSR pattern
// (its stored entries, the diagonal included, are the adjacency lists).
// Follows the GENRCM procedure of George and Liu (Computer Solution of Large
// Sparse Positive Definite Systems, Prentice-Hall 1981, ch. 4), which the
// reference calls through its vendored genrcm (rcmReorderCSR,
// MatrixReorderTools.h:59-80):
//   - components in order of their lowest-numbered node; for each, a
//     pseudo-peripheral root: the rooted level structure of the current root
//     is built, and the node of least degree (first in level order) of its
//     last level replaces the root while the structure gets deeper;
//   - from the root, breadth first: each node's unnumbered neighbours in
//     adjacency order, then that group, after its first node, ordered by
//     increasing degree with a stable insertion sort;
//   - the component's sequence reversed.
// Degrees count the stored entries whose node is not yet numbered, the
// diagonal included.  perm[k] = the node placed at position k.
std::vector<uint32_t> rcm_order(uint32_t n, const uint32_t *rp, const uint32_t *ci) {
    std::vector<uint32_t> perm(n);
    std::vector<uint8_t> mask(n, 1);  // 1: not yet numbered
    std::vector<uint32_t> level_row(n + 1), deg(n);
    // rooted level structure of `root` over unnumbered nodes; returns the number of levels
    auto level_set = [&](uint32_t root, uint32_t *level, uint32_t &size) -> uint32_t {
        mask[root] = 0;
        level[0] = root;
        uint32_t nlev = 0, lvl_end = 0;
        size = 1;
        for (;;) {
            const uint32_t lbegin = lvl_end;
            lvl_end = size;
            level_row[nlev++] = lbegin;
            for (uint32_t i = lbegin; i < lvl_end; ++i) {
                const uint32_t node = level[i];
                for (uint32_t k = rp[node]; k < rp[node + 1]; ++k) {
                    const uint32_t nb = ci[k];
                    if (mask[nb]) {
                        mask[nb] = 0;
                        level[size++] = nb;
                    }
                }
            }
            if (size == lvl_end) break;
        }
        level_row[nlev] = lvl_end;
        for (uint32_t i = 0; i < size; ++i) mask[level[i]] = 1;
        return nlev;
    };
    auto live_degree = [&](uint32_t node) {
        uint32_t d = 0;
        for (uint32_t k = rp[node]; k < rp[node + 1]; ++k) d += mask[ci[k]] ? 1u : 0u;
        return d;
    };
    uint32_t num = 0;
    for (uint32_t start = 0; start < n && num < n; ++start) {
        if (!mask[start]) continue;
        uint32_t *seq = perm.data() + num;
        // pseudo-peripheral root
        uint32_t root = start, size = 0;
        uint32_t nlev = level_set(root, seq, size);
        if (nlev != 1 && nlev != size) {
            for (;;) {
                uint32_t mindeg = size;
                const uint32_t first = level_row[nlev - 1];
                root = seq[first];
                if (first + 1 < size) {
                    for (uint32_t j = first; j < size; ++j) {
                        const uint32_t d = live_degree(seq[j]);
                        if (d < mindeg) {
                            root = seq[j];
                            mindeg = d;
                        }
                    }
                }
                uint32_t size2 = 0;
                const uint32_t nlev2 = level_set(root, seq, size2);
                if (nlev2 <= nlev) break;
                nlev = nlev2;
                if (size <= nlev) break;
            }
        }
        // degrees over the component (unnumbered neighbours, the diagonal included)
        {
            uint32_t cnt = 0;
            seq[cnt++] = root;
            mask[root] = 2;  // 2: reached by this breadth-first pass
            for (uint32_t i = 0; i < cnt; ++i) {
                const uint32_t node = seq[i];
                uint32_t d = 0;
                for (uint32_t k = rp[node]; k < rp[node + 1]; ++k) {
                    const uint32_t nb = ci[k];
                    if (mask[nb]) {
                        ++d;
                        if (mask[nb] == 1) {
                            mask[nb] = 2;
                            seq[cnt++] = nb;
                        }
                    }
                }
                deg[node] = d;
            }
            for (uint32_t i = 0; i < cnt; ++i) mask[seq[i]] = 1;
            size = cnt;
        }
        // Cuthill-McKee from the root
        mask[root] = 0;
        seq[0] = root;
        uint32_t lnbr = 1, lvl_end = 0;
        while (lvl_end < lnbr) {
            const uint32_t lbegin = lvl_end;
            lvl_end = lnbr;
            for (uint32_t i = lbegin; i < lvl_end; ++i) {
                const uint32_t node = seq[i];
                const uint32_t fnbr = lnbr;
                for (uint32_t k = rp[node]; k < rp[node + 1]; ++k) {
                    const uint32_t nb = ci[k];
                    if (mask[nb]) {
                        mask[nb] = 0;
                        seq[lnbr++] = nb;
                    }
                }
                // stable insertion sort of the group by increasing degree, in
                // the form the reference runs (rcm.cpp:3137-3350): the group's
                // first node keeps its place and the others are sorted behind
                // it (George and Liu's Fortran compares it too; the C version
                // the reference vendors stops one short)
                for (uint32_t k = fnbr + 1; k < lnbr; ++k) {
                    const uint32_t nb = seq[k];
                    uint32_t l = k;
                    while (l > fnbr + 1 && deg[seq[l - 1]] > deg[nb]) {
                        seq[l] = seq[l - 1];
                        --l;
                    }
                    seq[l] = nb;
                }
            }
        }
        std::reverse(seq, seq + size);
        num += size;
    }
    return perm;
}

}  // namespace pp

using namespace pp;

void set_host_error(const std::string &m);  // pp_core.cpp: the message pp_global_error() returns

// Runs f, turning exceptions into a status and the global error message.
template <class F>
static int guarded(std::nullptr_t, F &&f) {
    try {
        f();
        return PP_OK;
    } catch (const pp::Error &e) {
        set_host_error(e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        set_host_error("host allocation failed");
        return PP_ERR_NOMEM;
    } catch (const std::exception &e) {
        set_host_error(e.what());
        return PP_ERR_INVALID;
    }
}

extern "C" {

int pp_mtx_read(const char *path, int base, pp_mtx **out) {
    if (!out) return PP_ERR_INVALID;
    *out = nullptr;
    return guarded(nullptr, [&] {
        std::unique_ptr<pp_mtx> m(new pp_mtx);
        read_mtx(path, base, *m);
        *out = m.release();
    });
}

int pp_mtx_dims(const pp_mtx *m, uint32_t *nrow, uint32_t *ncol, uint32_t *nnz) {
    if (!m) return PP_ERR_INVALID;
    if (nrow) *nrow = m->nrow;
    if (ncol) *ncol = m->ncol;
    if (nnz) *nnz = (uint32_t)m->ci.size();
    return PP_OK;
}

int pp_mtx_csr(const pp_mtx *m, uint32_t *rowptr, uint32_t *colidx, double *values) {
    if (!m) return PP_ERR_INVALID;
    if (rowptr) std::memcpy(rowptr, m->rp.data(), m->rp.size() * sizeof(uint32_t));
    if (colidx && !m->ci.empty()) std::memcpy(colidx, m->ci.data(), m->ci.size() * sizeof(uint32_t));
    if (values && !m->v.empty()) std::memcpy(values, m->v.data(), m->v.size() * sizeof(double));
    return PP_OK;
}

void pp_mtx_destroy(pp_mtx *m) { delete m; }

int pp_mtx_read_dense(const char *path, uint32_t *nrow, uint32_t *ncol, double *values) {
    if (!nrow || !ncol) return PP_ERR_INVALID;
    return guarded(nullptr, [&] {
        std::vector<double> v;
        read_mtx_dense(path, *nrow, *ncol, values ? &v : nullptr);
        if (values && !v.empty()) std::memcpy(values, v.data(), v.size() * sizeof(double));
    });
}

int pp_reorder_rcm(uint32_t n, const uint32_t *rowptr, const uint32_t *colidx, uint32_t *perm) {
    if (!rowptr || !perm || (n && rowptr[n] && !colidx)) return PP_ERR_INVALID;
    return guarded(nullptr, [&] {
        for (uint32_t i = 0; i < n; ++i)
            PP_CHECK(rowptr[i] <= rowptr[i + 1], PP_ERR_INVALID, "row pointers must not decrease");
        for (uint32_t k = 0; k < rowptr[n]; ++k) PP_CHECK(colidx[k] < n, PP_ERR_INVALID, "column index out of range");
        const std::vector<uint32_t> p = rcm_order(n, rowptr, colidx);
        std::memcpy(perm, p.data(), (size_t)n * sizeof(uint32_t));
    });
}

}  // extern "C"
