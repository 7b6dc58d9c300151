// pp_ml_setup.cpp -- host-side building blocks of the multilevel low-rank
// Schur preconditioners (AMSED / GMSLR): a nested-dissection partition into a
// hierarchical interface decomposition, level extraction, rectangular CSR
// helpers and a symmetric tridiagonal eigensolver.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>

#include "pp_ml.h"

namespace pp {

// ---------------------------------------------------------------------------- graph partition
namespace {

struct AdjGraph {
    uint32_t n = 0;
    std::vector<uint32_t> rp, adj;  // symmetrised pattern without the diagonal
};

AdjGraph symmetric_graph(const HostCsr &A) {
    AdjGraph G;
    G.n = A.n;
    std::vector<std::vector<uint32_t>> nb(A.n);
    for (uint32_t i = 0; i < A.n; ++i)
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const uint32_t j = A.ci[k];
            if (j == i) continue;
            nb[i].push_back(j);
            nb[j].push_back(i);
        }
    G.rp.assign(A.n + 1, 0);
    for (uint32_t i = 0; i < A.n; ++i) {
        auto &v = nb[i];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        G.rp[i + 1] = G.rp[i] + (uint32_t)v.size();
    }
    G.adj.reserve(G.rp[A.n]);
    for (auto &v : nb) G.adj.insert(G.adj.end(), v.begin(), v.end());
    return G;
}

// BFS inside the vertex set marked `in`; returns the level of each reached
// vertex (UINT32_MAX elsewhere) and the last vertex reached.
uint32_t bfs(const AdjGraph &G, const std::vector<uint32_t> &set, const std::vector<char> &in, uint32_t start,
             std::vector<uint32_t> &lev) {
    for (uint32_t v : set) lev[v] = UINT32_MAX;
    std::queue<uint32_t> q;
    lev[start] = 0;
    q.push(start);
    uint32_t last = start;
    while (!q.empty()) {
        const uint32_t u = q.front();
        q.pop();
        last = u;
        for (uint32_t k = G.rp[u]; k < G.rp[u + 1]; ++k) {
            const uint32_t w = G.adj[k];
            if (in[w] && lev[w] == UINT32_MAX) {
                lev[w] = lev[u] + 1;
                q.push(w);
            }
        }
    }
    return last;
}

// Split `set` into two parts and a vertex separator: BFS level structure from
// a pseudo-peripheral vertex, separator = the level where the reached count
// crosses half.  Vertices BFS does not reach (other components) join the
// smaller part; they touch neither the separator nor the other part.
void bisect(const AdjGraph &G, const std::vector<uint32_t> &set, std::vector<char> &in, std::vector<uint32_t> &lev,
            std::vector<uint32_t> &p0, std::vector<uint32_t> &p1, std::vector<uint32_t> &sep) {
    p0.clear();
    p1.clear();
    sep.clear();
    if (set.empty()) return;
    for (uint32_t v : set) in[v] = 1;
    // pseudo-peripheral start: a few BFS sweeps from the farthest vertex
    uint32_t s = set[0];
    for (int it = 0; it < 3; ++it) s = bfs(G, set, in, s, lev);
    bfs(G, set, in, s, lev);
    uint32_t maxl = 0, reached = 0;
    for (uint32_t v : set)
        if (lev[v] != UINT32_MAX) {
            maxl = std::max(maxl, lev[v]);
            ++reached;
        }
    std::vector<uint32_t> cnt(maxl + 2, 0);
    for (uint32_t v : set)
        if (lev[v] != UINT32_MAX) cnt[lev[v]]++;
    uint32_t cut = 0, acc = 0;
    for (uint32_t l = 0; l <= maxl; ++l) {
        if (acc + cnt[l] / 2 >= reached / 2) {
            cut = l;
            break;
        }
        acc += cnt[l];
    }
    for (uint32_t v : set) {
        const uint32_t l = lev[v];
        if (l == UINT32_MAX) continue;
        if (l < cut) p0.push_back(v);
        else if (l == cut) sep.push_back(v);
        else p1.push_back(v);
    }
    for (uint32_t v : set)
        if (lev[v] == UINT32_MAX) (p0.size() <= p1.size() ? p0 : p1).push_back(v);
    for (uint32_t v : set) in[v] = 0;
}

}  // namespace

// Hierarchical interface decomposition by recursive bisection to `ndom`
// (a power of two) leaves.  Level 0 = the leaves' interiors; level t
// (1 <= t <= nlev-2) = the separators found at bisection depth D-t+1 (they are
// pairwise non-adjacent: each lies inside a different subdomain of depth
// D-t); the last level = all remaining separators.  Every level is a set of
// mutually non-adjacent blocks, so each B_l is block diagonal.
Hid build_hid(const HostCsr &A, uint32_t ndom, uint32_t nlev) {
    PP_CHECK(ndom >= 2 && (ndom & (ndom - 1)) == 0, PP_ERR_INVALID, "ndom must be a power of two >= 2");
    uint32_t depth = 0;
    while ((1u << depth) < ndom) ++depth;
    PP_CHECK(nlev >= 2 && nlev <= depth + 1, PP_ERR_INVALID, "levels must lie in [2, log2(ndom) + 1]");
    const AdjGraph G = symmetric_graph(A);
    std::vector<char> in(A.n, 0);
    std::vector<uint32_t> lev(A.n, UINT32_MAX);
    // sep_depth[v]: 0 = leaf interior, d = separator found at depth d
    std::vector<uint32_t> sep_depth(A.n, 0), leaf(A.n, 0);
    std::vector<std::vector<uint32_t>> parts{std::vector<uint32_t>(A.n)};
    std::iota(parts[0].begin(), parts[0].end(), 0u);
    std::vector<std::vector<std::vector<uint32_t>>> seps(depth + 1);
    for (uint32_t d = 1; d <= depth; ++d) {
        std::vector<std::vector<uint32_t>> next;
        for (auto &P : parts) {
            std::vector<uint32_t> p0, p1, s;
            bisect(G, P, in, lev, p0, p1, s);
            for (uint32_t v : s) sep_depth[v] = d;
            seps[d].push_back(std::move(s));
            next.push_back(std::move(p0));
            next.push_back(std::move(p1));
        }
        parts.swap(next);
    }
    Hid H;
    H.nlev = nlev;
    // level -> list of blocks (each a vertex list)
    std::vector<std::vector<std::vector<uint32_t>>> blocks(nlev);
    blocks[0] = parts;  // ndom leaves
    for (uint32_t t = 1; t + 1 < nlev; ++t) blocks[t] = seps[depth - t + 1];
    std::vector<uint32_t> rest;
    for (uint32_t d = 1; d <= depth - (nlev - 2); ++d)
        for (auto &s : seps[d]) rest.insert(rest.end(), s.begin(), s.end());
    std::sort(rest.begin(), rest.end());
    blocks[nlev - 1] = {rest};
    H.perm.reserve(A.n);
    H.lev_ptr.push_back(0);
    for (uint32_t l = 0; l < nlev; ++l) {
        std::vector<uint32_t> dptr{(uint32_t)H.perm.size()};
        for (auto &b : blocks[l]) {
            std::vector<uint32_t> bb = b;
            std::sort(bb.begin(), bb.end());  // natural order inside a block
            H.perm.insert(H.perm.end(), bb.begin(), bb.end());
            dptr.push_back((uint32_t)H.perm.size());
        }
        H.dom_ptr.push_back(dptr);
        H.lev_ptr.push_back((uint32_t)H.perm.size());
    }
    PP_CHECK(H.perm.size() == A.n, PP_ERR_INVALID, "partition lost vertices");
    H.iperm.assign(A.n, 0);
    for (uint32_t k = 0; k < A.n; ++k) H.iperm[H.perm[k]] = k;
    return H;
}

// Rows [r0, r1) and columns [c0, c1) of P A P^T (new numbering), CSR with
// local indices, columns ascending.
HostCsr extract_block(const HostCsr &A, const Hid &H, uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1,
                      uint32_t &ncol) {
    HostCsr B;
    B.n = r1 - r0;
    ncol = c1 - c0;
    B.rp.assign(B.n + 1, 0);
    std::vector<std::pair<uint32_t, double>> row;
    for (uint32_t i = r0; i < r1; ++i) {
        row.clear();
        const uint32_t oi = H.perm[i];
        for (uint32_t k = A.rp[oi]; k < A.rp[oi + 1]; ++k) {
            const uint32_t j = H.iperm[A.ci[k]];
            if (j >= c0 && j < c1) row.emplace_back(j - c0, A.v[k]);
        }
        std::sort(row.begin(), row.end());
        for (auto &e : row) {
            B.ci.push_back(e.first);
            B.v.push_back(e.second);
        }
        B.rp[i - r0 + 1] = (uint32_t)B.ci.size();
    }
    return B;
}

// Transpose of an nrow x ncol CSR.
HostCsr transpose_rect(const HostCsr &A, uint32_t ncol) {
    HostCsr T;
    T.n = ncol;
    T.rp.assign(ncol + 1, 0);
    const uint32_t nnz = A.rp[A.n];
    T.ci.resize(nnz);
    T.v.resize(nnz);
    for (uint32_t k = 0; k < nnz; ++k) T.rp[A.ci[k] + 1]++;
    for (uint32_t j = 0; j < ncol; ++j) T.rp[j + 1] += T.rp[j];
    std::vector<uint32_t> nx(T.rp.begin(), T.rp.end() - 1);
    for (uint32_t i = 0; i < A.n; ++i)
        for (uint32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const uint32_t p = nx[A.ci[k]]++;
            T.ci[p] = i;
            T.v[p] = A.v[k];
        }
    return T;
}

// ---------------------------------------------------------------------------- tridiagonal eigensolver
// Eigenvalues/vectors of the symmetric tridiagonal matrix (diagonal d[0..n),
// off-diagonal e[0..n-1)) by the implicit QL method with Wilkinson shifts
// (Bowdler, Martin, Reinsch & Wilkinson, "The QR and QL algorithms for
// symmetric matrices", Numer. Math. 11, 1968).  On return d holds the
// eigenvalues in ascending order and Z (column-major n x n) the eigenvectors.
void tridiag_eig(std::vector<double> &d, std::vector<double> e, std::vector<double> &Z) {
    const int n = (int)d.size();
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    if (n <= 1) return;
    e.resize(n, 0.0);
    e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-15 * dd) break;
            }
            if (m != l) {
                PP_CHECK(++iter <= 60, PP_ERR_BREAKDOWN, "tridiagonal eigensolver did not converge");
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i], b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {  // rotate columns i, i+1 of Z
                        double *zi = &Z[(size_t)i * n], *zi1 = &Z[(size_t)(i + 1) * n];
                        f = zi1[k];
                        zi1[k] = s * zi[k] + c * f;
                        zi[k] = c * zi[k] - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    // sort ascending with their vectors
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    std::vector<double> d2(n), Z2((size_t)n * n);
    for (int k = 0; k < n; ++k) {
        d2[k] = d[idx[k]];
        std::copy(Z.begin() + (size_t)idx[k] * n, Z.begin() + (size_t)(idx[k] + 1) * n, Z2.begin() + (size_t)k * n);
    }
    d.swap(d2);
    Z.swap(Z2);
}

}  // namespace pp
