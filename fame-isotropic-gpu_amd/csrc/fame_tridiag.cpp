/*
 * fame_tridiag.cpp -- small dense pieces of the Lanczos restart, on the host.
 *
 * fame_tridiag_eig replaces LAPACKE_zpteqr (Lanczos_Isotropic.cu:53): all
 * eigenpairs of the symmetric tridiagonal T, eigenvalues in DESCENDING order,
 * eigenvectors as columns of z (n x n, column-major).  Method: cyclic Jacobi
 * rotations on the dense n x n matrix (n <= 128 here), which is unconditionally
 * convergent and accurate to a few ulps of ||T||.
 *
 * fame_tridiag_qr_step is the exact-shift restart step of
 * Lanczos_LockPurge_gpu.cu:7-48 (one explicit shifted QR step T - mu I = QR,
 * T' = RQ + mu I), written from the textbook step (Golub & Van Loan, Matrix
 * Computations, 4th ed., Sec. 8.3.3, explicit form).  The behaviour the restart
 * depends on is the reference's: the shift is exact (an unwanted Ritz value),
 * the rotation that annihilates T(i+1,i) has c = a/r, s = -b/r with
 * r = hypot(a, b) >= 0 (so R has a non-negative diagonal), and the
 * off-diagonal of T' is the mean of its computed sub- and super-diagonal.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fame_hip.h"

extern "C" int fame_tridiag_eig(int n, const double* d, const double* e, double* w, double* z) {
    if (n <= 0 || !d || !w || !z || (n > 1 && !e)) return FAME_ERR_ARG;
    std::vector<double> A((size_t)n * n, 0.0), V((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        A[(size_t)i * n + i] = d[i];
        if (i + 1 < n) { A[(size_t)i * n + i + 1] = e[i]; A[(size_t)(i + 1) * n + i] = e[i]; }
        V[(size_t)i * n + i] = 1.0;
    }
    double fro = 0;
    for (double v : A) fro += v * v;
    fro = std::sqrt(fro);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
        if (std::sqrt(off) <= 1e-17 * fro || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {   /* A <- A J  (columns p, q) */
                    double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {   /* A <- J^T A  (rows p, q) */
                    double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                A[(size_t)p * n + q] = 0.0; A[(size_t)q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {   /* V <- V J (V stored row-major: V[k][p]) */
                    double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return A[(size_t)a * n + a] > A[(size_t)b * n + b]; });
    for (int j = 0; j < n; ++j) {
        w[j] = A[(size_t)idx[j] * n + idx[j]];
        for (int i = 0; i < n; ++i) z[(size_t)j * n + i] = V[(size_t)i * n + idx[j]];
    }
    return FAME_OK;
}

/*
 * One shifted QR step on the symmetric tridiagonal (diag[0..n-1], off[0..n-2]).
 *
 * QR sweep.  G_i acts on rows (i, i+1) as [[c, -s], [s, c]] and zeroes the
 * sub-diagonal entry b = (i+1, i) against the current pivot a = (i, i).  R is
 * upper triangular with three bands, kept as rd (diagonal, in diag), ru (first
 * super-diagonal) and rw (second super-diagonal).
 *
 * RQ sweep.  R G_0^T ... G_{n-2}^T: G_i^T mixes columns (i, i+1).  Before it is
 * applied, column i holds (hi, hd) at rows (i-1, i) from the previous rotation,
 * column i+1 holds R's (rw[i-1], ru[i], rd[i+1]).  After it, column i is final:
 *   T'(i-1, i) = c hi - s rw[i-1],  T'(i, i) = c hd - s ru[i],  T'(i+1, i) = -s rd[i+1],
 * and column i+1 carries s hd + c ru[i] at row i and c rd[i+1] at row i+1 into
 * the next rotation (its entry at row i-1 lies outside the band of the symmetric
 * result and is not formed).
 *
 * ru and rw are caller scratch of n doubles; c and s receive the n-1 rotations.
 */
extern "C" int fame_tridiag_qr_step(double* diag, double* off, double* ru, double* rw, double* c, double* s,
                                    double mu, int n) {
    if (n < 1 || !diag || (n > 1 && (!off || !ru || !rw || !c || !s))) return FAME_ERR_ARG;
    if (n == 1) return FAME_OK;
    for (int i = 0; i < n; ++i) diag[i] -= mu;
    for (int i = 0; i + 1 < n; ++i) ru[i] = off[i];
    rw[n - 1] = rw[n - 2] = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
        const double a = diag[i], b = off[i];
        const double r = std::hypot(a, b);
        const double ci = r > 0.0 ? a / r : 1.0, si = r > 0.0 ? -b / r : 0.0;
        c[i] = ci; s[i] = si;
        diag[i] = r;
        const double up = ru[i], dn = diag[i + 1];
        ru[i] = ci * up - si * dn;
        diag[i + 1] = si * up + ci * dn;
        if (i + 2 < n) {
            const double u1 = ru[i + 1];
            rw[i] = -si * u1;
            ru[i + 1] = ci * u1;
        }
    }
    /* off now receives the sub-diagonal of RQ, ru its super-diagonal */
    double hi = 0.0, hd = diag[0];
    for (int i = 0; i + 1 < n; ++i) {
        const double ci = c[i], si = s[i];
        const double rdn = diag[i + 1], rui = ru[i];
        if (i > 0) ru[i - 1] = ci * hi - si * rw[i - 1];
        diag[i] = ci * hd - si * rui;
        off[i] = -si * rdn;
        hi = si * hd + ci * rui;
        hd = ci * rdn;
    }
    diag[n - 1] = hd;
    ru[n - 2] = hi;
    for (int i = 0; i + 1 < n; ++i) off[i] = 0.5 * (off[i] + ru[i]);
    for (int i = 0; i < n; ++i) diag[i] += mu;
    return FAME_OK;
}
