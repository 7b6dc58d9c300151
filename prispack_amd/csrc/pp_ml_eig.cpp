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
