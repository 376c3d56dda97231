/*
 * fame_tridiag.cpp -- small dense pieces of the restart, on the host.
 *
 * fame_tridiag_eig replaces LAPACKE_zpteqr (Lanczos_Isotropic.cu:53): all
 * eigenpairs of the symmetric tridiagonal T, eigenvalues in DESCENDING order,
 * eigenvectors as columns of z (n x n, column-major).  Method: cyclic Jacobi
 * rotations on the dense n x n matrix (n <= 63 here), which is unconditionally
 * convergent and accurate to a few ulps of ||T||.
 *
 * fame_gvqrrq restates GVqrrq_g / given_rotation (Lanczos_LockPurge_gpu.cu:95-266):
 * one exact-shift QR step on T with the reference's Givens sign convention and
 * the (T1 + T3)/2 symmetrisation.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

int fame_tridiag_eig(int n, const double* d, const double* e, double* w, double* z) {
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
    return 0;
}

static void given_rotation(double* c, double* s, double* r, double a, double b) {
    double t, u;
    if (std::fabs(a) > std::fabs(b)) {
        t = b / a; u = std::fabs(a) / a * std::sqrt(1 + t * t);
        *c = 1.0 / u; *s = *c * t; *r = a * u; *s = -*s;
    } else {
        t = a / b; u = std::fabs(b) / b * std::sqrt(1 + t * t);
        *s = 1.0 / u; *c = *s * t; *r = b * u; *s = -*s;
    }
}

void fame_gvqrrq(double* T0, double* T1, double* T2, double* T3, double* c, double* s, double shift, int n) {
    double a, b, r, R[6];
    std::memcpy(T3, T1, (n - 1) * sizeof(double));
    for (int i = 0; i < n; ++i) T0[i] -= shift;
    for (int i = 0; i < n - 1; ++i) {            /* Q' T */
        a = T0[i]; b = T1[i];
        given_rotation(&c[i], &s[i], &r, a, b);
        T0[i] = r; T1[i] = 0.0;
        a = T3[i]; b = T0[i + 1];
        T3[i] = c[i] * a - s[i] * b;
        T0[i + 1] = s[i] * a + c[i] * b;
        if (i < n - 2) { T2[i] = -s[i] * T3[i + 1]; T3[i + 1] = c[i] * T3[i + 1]; }
    }
    R[0] = T0[0]; R[1] = T3[0]; R[2] = T1[0]; R[3] = T0[1];    /* T Q */
    T0[0] = c[0] * R[0] - s[0] * R[1];
    T3[0] = s[0] * R[0] + c[0] * R[1];
    T1[0] = c[0] * R[2] - s[0] * R[3];
    T0[1] = s[0] * R[2] + c[0] * R[3];
    for (int i = 1; i < n - 1; ++i) {
        R[0] = T3[i - 1]; R[1] = T2[i - 1]; R[2] = T0[i]; R[3] = T3[i]; R[4] = T1[i]; R[5] = T0[i + 1];
        T3[i - 1] = c[i] * R[0] - s[i] * R[1];
        T2[i - 1] = s[i] * R[0] + c[i] * R[1];
        T0[i] = c[i] * R[2] - s[i] * R[3];
        T3[i] = s[i] * R[2] + c[i] * R[3];
        T1[i] = c[i] * R[4] - s[i] * R[5];
        T0[i + 1] = s[i] * R[4] + c[i] * R[5];
    }
    for (int i = 0; i < n - 1; ++i) T1[i] = (T1[i] + T3[i]) / 2.0;
    for (int i = 0; i < n; ++i) T0[i] += shift;
}
