// pp_ml_eig.cpp -- small dense eigen-solvers for the nonsymmetric low-rank
// correction: eigenvalues of an upper Hessenberg matrix by the Francis
// double-shift QR iteration (Golub & Van Loan, "Matrix Computations", 4th ed.,
// Algorithms 7.5.1-7.5.2, with the standard deflation test), and eigenvectors
// of selected real eigenvalues by inverse iteration with a dense LU.
#include <algorithm>
#include <cmath>
#include <complex>

#include "pp_ml.h"

namespace pp {

namespace {

// Householder reflector P = I - 2 v v^T / v^T v with P [x y z]^T = [* 0 0]^T
// (nz = 3) or P [x y]^T = [* 0]^T (nz = 2), applied to rows r0..r0+nz-1 of
// columns [c0, c1] and to columns r0.. of rows [q0, q1].
void reflect(std::vector<double> &H, int n, int r0, int nz, double x, double y, double z, int c0, int c1, int q0,
             int q1) {
    auto at = [&](int i, int j) -> double & { return H[(size_t)j * n + i]; };
    double alpha = std::sqrt(x * x + y * y + (nz == 3 ? z * z : 0.0));
    if (alpha == 0.0) return;
    if (x > 0) alpha = -alpha;
    double v[3] = {x - alpha, y, nz == 3 ? z : 0.0};
    const double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (vv == 0.0) return;
    for (int j = c0; j <= c1; ++j) {  // rows
        double s = 0.0;
        for (int t = 0; t < nz; ++t) s += v[t] * at(r0 + t, j);
        s = 2.0 * s / vv;
        for (int t = 0; t < nz; ++t) at(r0 + t, j) -= s * v[t];
    }
    for (int i = q0; i <= q1; ++i) {  // columns
        double s = 0.0;
        for (int t = 0; t < nz; ++t) s += at(i, r0 + t) * v[t];
        s = 2.0 * s / vv;
        for (int t = 0; t < nz; ++t) at(i, r0 + t) -= s * v[t];
    }
}

}  // namespace

// Eigenvalues (re, im) of the n x n upper Hessenberg matrix H (column-major,
// destroyed).  Returns false if the iteration did not converge.
bool hessenberg_eigenvalues(std::vector<double> H, int n, std::vector<double> &re, std::vector<double> &im) {
    auto at = [&](int i, int j) -> double & { return H[(size_t)j * n + i]; };
    re.assign(n, 0.0);
    im.assign(n, 0.0);
    int p = n - 1, iter = 0;
    while (p >= 0) {
        // find the start l of the unreduced block ending at p
        int l = p;
        while (l > 0) {
            const double s = std::fabs(at(l - 1, l - 1)) + std::fabs(at(l, l));
            if (std::fabs(at(l, l - 1)) <= 1e-15 * (s == 0.0 ? 1.0 : s)) {
                at(l, l - 1) = 0.0;
                break;
            }
            --l;
        }
        if (l == p) {  // 1x1 block
            re[p] = at(p, p);
            im[p] = 0.0;
            --p;
            iter = 0;
            continue;
        }
        if (l == p - 1) {  // 2x2 block
            const double a = at(p - 1, p - 1), b = at(p - 1, p), c = at(p, p - 1), d = at(p, p);
            const double tr = a + d, det = a * d - b * c, disc = 0.25 * tr * tr - det;
            if (disc >= 0.0) {
                const double sq = std::sqrt(disc);
                re[p - 1] = 0.5 * tr + sq;
                re[p] = 0.5 * tr - sq;
                im[p - 1] = im[p] = 0.0;
            } else {
                const double sq = std::sqrt(-disc);
                re[p - 1] = re[p] = 0.5 * tr;
                im[p - 1] = sq;
                im[p] = -sq;
            }
            p -= 2;
            iter = 0;
            continue;
        }
        if (++iter > 60 * n) return false;
        // double shift from the trailing 2x2 (exceptional shift every 10 steps)
        double s = at(p - 1, p - 1) + at(p, p);
        double t = at(p - 1, p - 1) * at(p, p) - at(p - 1, p) * at(p, p - 1);
        if (iter % 10 == 0) {
            const double e = std::fabs(at(p, p - 1)) + std::fabs(at(p - 1, p - 2));
            s = 1.5 * e;
            t = e * e;
        }
        double x = at(l, l) * at(l, l) + at(l, l + 1) * at(l + 1, l) - s * at(l, l) + t;
        double y = at(l + 1, l) * (at(l, l) + at(l + 1, l + 1) - s);
        double z = at(l + 1, l) * at(l + 2, l + 1);
        for (int k = l; k <= p - 2; ++k) {
            reflect(H, n, k, 3, x, y, z, std::max(l, k - 1), p, l, std::min(k + 3, p));
            x = at(k + 1, k);
            y = at(k + 2, k);
            if (k < p - 2) z = at(k + 3, k);
        }
        reflect(H, n, p - 1, 2, x, y, 0.0, p - 2, p, l, p);
    }
    return true;
}

// Real Schur form of the n x n upper Hessenberg matrix H (column-major):
// H = Z T Z^T with Z orthogonal and T quasi upper triangular, by the same
// Francis double-shift iteration with every reflector applied to all of T
// and accumulated into Z (Golub & Van Loan Alg. 7.5.2 with the full update).
// 2 x 2 diagonal blocks with real eigenvalues are split by one rotation, so
// wr[i], wi[i] are the eigenvalues in the order they sit on T's diagonal
// (complex pairs as adjacent entries).  Returns false if the iteration did
// not converge.
bool hessenberg_schur(std::vector<double> &H, int n, std::vector<double> &Z, std::vector<double> &wr,
                      std::vector<double> &wi) {
    auto at = [&](int i, int j) -> double & { return H[(size_t)j * n + i]; };
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    wr.assign(n, 0.0);
    wi.assign(n, 0.0);
    // reflector on rows/columns r0..r0+nz-1: T <- P T P over the whole matrix, Z <- Z P
    auto reflect_all = [&](int r0, int nz, double x, double y, double z) {
        double alpha = std::sqrt(x * x + y * y + (nz == 3 ? z * z : 0.0));
        if (alpha == 0.0) return;
        if (x > 0) alpha = -alpha;
        const double v[3] = {x - alpha, y, nz == 3 ? z : 0.0};
        const double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (vv == 0.0) return;
        for (int j = std::max(0, r0 - 1); j < n; ++j) {
            double s = 0.0;
            for (int t = 0; t < nz; ++t) s += v[t] * at(r0 + t, j);
            s = 2.0 * s / vv;
            for (int t = 0; t < nz; ++t) at(r0 + t, j) -= s * v[t];
        }
        const int q1 = std::min(r0 + 3, n - 1);
        for (int i = 0; i <= q1; ++i) {
            double s = 0.0;
            for (int t = 0; t < nz; ++t) s += at(i, r0 + t) * v[t];
            s = 2.0 * s / vv;
            for (int t = 0; t < nz; ++t) at(i, r0 + t) -= s * v[t];
        }
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int t = 0; t < nz; ++t) s += Z[(size_t)(r0 + t) * n + i] * v[t];
            s = 2.0 * s / vv;
            for (int t = 0; t < nz; ++t) Z[(size_t)(r0 + t) * n + i] -= s * v[t];
        }
    };
    // rotation [c s; -s c] on rows/columns p-1, p of T and columns of Z
    auto rotate_all = [&](int p, double c, double s) {
        for (int j = p - 1; j < n; ++j) {
            const double a = at(p - 1, j), b = at(p, j);
            at(p - 1, j) = c * a + s * b;
            at(p, j) = -s * a + c * b;
        }
        for (int i = 0; i <= p; ++i) {
            const double a = at(i, p - 1), b = at(i, p);
            at(i, p - 1) = c * a + s * b;
            at(i, p) = -s * a + c * b;
        }
        for (int i = 0; i < n; ++i) {
            double *za = &Z[(size_t)(p - 1) * n + i], *zb = &Z[(size_t)p * n + i];
            const double a = *za, b = *zb;
            *za = c * a + s * b;
            *zb = -s * a + c * b;
        }
    };
    int p = n - 1, iter = 0;
    while (p >= 0) {
        int l = p;
        while (l > 0) {
            const double s = std::fabs(at(l - 1, l - 1)) + std::fabs(at(l, l));
            if (std::fabs(at(l, l - 1)) <= 1e-15 * (s == 0.0 ? 1.0 : s)) {
                at(l, l - 1) = 0.0;
                break;
            }
            --l;
        }
        if (l == p) {
            wr[p] = at(p, p);
            wi[p] = 0.0;
            --p;
            iter = 0;
            continue;
        }
        if (l == p - 1) {
            const double a = at(p - 1, p - 1), b = at(p - 1, p), c = at(p, p - 1), d = at(p, p);
            const double tr = a + d, det = a * d - b * c, disc = 0.25 * tr * tr - det;
            if (disc >= 0.0) {
                // split: the rotation taking an eigenvector of the block to e_1
                const double lam = 0.5 * tr + (0.5 * (a - d) >= 0 ? 1.0 : -1.0) * std::sqrt(disc);
                double x = b, y = lam - a;
                if (std::fabs(x) + std::fabs(y) == 0.0) {
                    x = lam - d;
                    y = c;
                }
                const double r = std::hypot(x, y);
                if (r > 0.0) rotate_all(p, x / r, y / r);
                at(p, p - 1) = 0.0;
                wr[p - 1] = at(p - 1, p - 1);
                wr[p] = at(p, p);
                wi[p - 1] = wi[p] = 0.0;
            } else {
                const double sq = std::sqrt(-disc);
                wr[p - 1] = wr[p] = 0.5 * tr;
                wi[p - 1] = sq;
                wi[p] = -sq;
            }
            p -= 2;
            iter = 0;
            continue;
        }
        if (++iter > 60 * n) return false;
        double s = at(p - 1, p - 1) + at(p, p);
        double t = at(p - 1, p - 1) * at(p, p) - at(p - 1, p) * at(p, p - 1);
        if (iter % 10 == 0) {
            const double e = std::fabs(at(p, p - 1)) + std::fabs(at(p - 1, p - 2));
            s = 1.5 * e;
            t = e * e;
        }
        double x = at(l, l) * at(l, l) + at(l, l + 1) * at(l + 1, l) - s * at(l, l) + t;
        double y = at(l + 1, l) * (at(l, l) + at(l + 1, l + 1) - s);
        double z = at(l + 1, l) * at(l + 2, l + 1);
        for (int k = l; k <= p - 2; ++k) {
            reflect_all(k, 3, x, y, z);
            x = at(k + 1, k);
            y = at(k + 2, k);
            if (k < p - 2) z = at(k + 3, k);
        }
        reflect_all(p - 1, 2, x, y, 0.0);
    }
    return true;
}

// Unit eigenvector y of the n x n matrix H (column-major) for the real
// eigenvalue lambda, by inverse iteration on (H - lambda I).
std::vector<double> inverse_iteration(const std::vector<double> &H, int n, double lambda) {
    std::vector<double> M(H);
    const double shift = lambda + 1e-10 * std::max(1.0, std::fabs(lambda));
    for (int i = 0; i < n; ++i) M[(size_t)i * n + i] -= shift;
    // LU with partial pivoting (column-major)
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int pr = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(M[(size_t)k * n + i]) > std::fabs(M[(size_t)k * n + pr])) pr = i;
        piv[k] = pr;
        if (pr != k)
            for (int j = 0; j < n; ++j) std::swap(M[(size_t)j * n + k], M[(size_t)j * n + pr]);
        double d = M[(size_t)k * n + k];
        if (d == 0.0) d = M[(size_t)k * n + k] = 1e-300;
        for (int i = k + 1; i < n; ++i) {
            const double f = (M[(size_t)k * n + i] /= d);
            for (int j = k + 1; j < n; ++j) M[(size_t)j * n + i] -= f * M[(size_t)j * n + k];
        }
    }
    std::vector<double> y(n, 1.0);
    for (int it = 0; it < 3; ++it) {
        for (int k = 0; k < n; ++k) std::swap(y[k], y[piv[k]]);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < i; ++k) y[i] -= M[(size_t)k * n + i] * y[k];
        for (int i = n - 1; i >= 0; --i) {
            for (int k = i + 1; k < n; ++k) y[i] -= M[(size_t)k * n + i] * y[k];
            y[i] /= M[(size_t)i * n + i];
        }
        double nrm = 0.0;
        for (double v : y) nrm += v * v;
        nrm = std::sqrt(nrm);
        if (!(nrm > 0) || !std::isfinite(nrm)) break;
        for (double &v : y) v /= nrm;
    }
    return y;
}

}  // namespace pp
