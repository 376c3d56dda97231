/*
 * fame_device.hpp -- device-side building blocks for gfx950:
 * complex arithmetic, the batched-line FFT engine (LDS Stockham autosort with
 * radices 2,3,4,5,7,8,11,13 and Bluestein chirp-z for lengths with larger
 * prime factors) and the on-the-fly Fourier symbols / Pi bases of the
 * null-space-free Maxwell operator.
 *
 * Axes.  Every problem is stored with three axes (a = inner, contiguous;
 * b = middle; c = outer).  For General lattices (a,b,c) = (z,y,x), for Simple
 * lattices (a,b,c) = (x,y,z); in both cases the reference's Fourier index is
 * f = (c*nb + b)*na + a (FAME_Matrix_Lambdas.cu:96-132 Simple, :317-371 General).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fame {

template <typename T> struct cx;
template <> struct cx<double> { typedef double2 t; };
template <> struct cx<float>  { typedef float2 t; };

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) { return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); } /* a*conj(b) */
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) { return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

/* exp(i*pi*x) with the argument in units of pi, accurate for any x */
__device__ __forceinline__ double2 expipi(double x) { double s, c; sincospi(x, &s, &c); return make_double2(c, s); }

/* ------------------------------------------------------------------------- */
/*  FFT plan as seen by a kernel                                              */
/* ------------------------------------------------------------------------- */
#define FAME_MAX_PASSES 10
struct FftPlanDev {
    int n;          /* transform length                                        */
    int M;          /* Stockham length: n (direct) or the Bluestein length     */
    int blue;       /* 1 when Bluestein chirp-z is used                        */
    int nrad;
    int rad[FAME_MAX_PASSES];
    const void* W;      /* M entries exp(-2 pi i m / M), precision of the kernel */
    const void* chirp;  /* n entries exp(-i pi m^2 / n)                          */
    const void* bhat_m; /* M entries: FFT_M of the kernel for the forward sign   */
    const void* bhat_p; /* M entries: same for the inverse sign                  */
};

/* One Stockham radix-R pass over L lines held in LDS (src -> dst).
 * Butterfly j of a line: inputs src[j + t*M/R], twiddle W_{Ns R}^{t k},
 * outputs dst[(j/Ns)*Ns*R + k + t*Ns], k = j mod Ns. sign < 0: forward. */
template <typename T2, int R>
__device__ __forceinline__ void stockham_pass(const T2* __restrict__ src, T2* __restrict__ dst, int L, int pitch,
                                              int M, int Ns, const T2* __restrict__ W, int sign) {
    const int nb = M / R;
    const int tstep = M / (Ns * R);
    const int total = L * nb;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int line = idx / nb;
        const int j = idx - line * nb;
        const T2* a = src + line * pitch;
        T2* b = dst + line * pitch;
        const int k = j % Ns;
        T2 v[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            T2 x = a[j + t * nb];
            if (t > 0 && k > 0) {
                T2 w = W[(t * k * tstep) % M];
                if (sign > 0) w = cconj(w);
                x = cmul(x, w);
            }
            v[t] = x;
        }
        T2 o[R];
        if (R == 2) {
            o[0] = cadd(v[0], v[1]); o[1] = csub(v[0], v[1]);
        } else if (R == 4) {
            T2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]), a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
            T2 j3;
            if (sign < 0) { j3.x = a3.y; j3.y = -a3.x; } else { j3.x = -a3.y; j3.y = a3.x; }
            o[0] = cadd(a0, a2); o[2] = csub(a0, a2); o[1] = cadd(a1, j3); o[3] = csub(a1, j3);
        } else {
            const int step = M / R;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                T2 acc = v[0];
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    T2 w = W[((q * t) % R) * step];
                    if (sign > 0) w = cconj(w);
                    acc = cadd(acc, cmul(v[t], w));
                }
                o[q] = acc;
            }
        }
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) b[base + t * Ns] = o[t];
    }
}

/* Full Stockham transform of L lines of length pl.M held in LDS buffer A (B is
 * scratch of the same size).  Returns the buffer holding the result.          */
template <typename T2>
__device__ T2* stockham_lines(T2* A, T2* B, int L, int pitch, const FftPlanDev& pl, const T2* Wl, int sign) {
    int Ns = 1;
    for (int p = 0; p < pl.nrad; ++p) {
        const int r = pl.rad[p];
        switch (r) {
            case 2: stockham_pass<T2, 2>(A, B, L, pitch, pl.M, Ns, Wl, sign); break;
            case 3: stockham_pass<T2, 3>(A, B, L, pitch, pl.M, Ns, Wl, sign); break;
            case 4: stockham_pass<T2, 4>(A, B, L, pitch, pl.M, Ns, Wl, sign); break;
            case 5: stockham_pass<T2, 5>(A, B, L, pitch, pl.M, Ns, Wl, sign); break;
            case 7: stockham_pass<T2, 7>(A, B, L, pitch, pl.M, Ns, Wl, sign); break;
            case 8: stockham_pass<T2, 8>(A, B, L, pitch, pl.M, Ns, Wl, sign); break;
            case 11: stockham_pass<T2, 11>(A, B, L, pitch, pl.M, Ns, Wl, sign); break;
            case 13: stockham_pass<T2, 13>(A, B, L, pitch, pl.M, Ns, Wl, sign); break;
            default: break;
        }
        __syncthreads();
        T2* t = A; A = B; B = t;
        Ns *= r;
    }
    return A;
}

/* Transform L lines of length pl.n that sit (unpadded) at the front of each
 * LDS line of `pitch` >= pl.M elements.  Handles both direct and Bluestein
 * plans.  Returns the buffer (A or B) holding the n outputs of each line.
 * Caller must have synchronised after filling A.                              */
template <typename T2, typename T>
__device__ T2* fft_lines(T2* A, T2* B, int L, int pitch, const FftPlanDev& pl, const T2* Wl, int sign) {
    if (!pl.blue) return stockham_lines<T2>(A, B, L, pitch, pl, Wl, sign);
    const T2* chirp = reinterpret_cast<const T2*>(pl.chirp);
    const T2* bhat = reinterpret_cast<const T2*>(sign < 0 ? pl.bhat_m : pl.bhat_p);
    const int n = pl.n, M = pl.M;
    /* pre-chirp and zero padding: a[j] = x[j] w[j], w = exp(sign i pi j^2/n) */
    for (int idx = threadIdx.x; idx < L * M; idx += blockDim.x) {
        int line = idx / M, j = idx - line * M;
        T2* a = A + line * pitch;
        if (j < n) {
            T2 w = chirp[j];
            if (sign > 0) w = cconj(w);
            a[j] = cmul(a[j], w);
        } else {
            T2 z; z.x = 0; z.y = 0;
            a[j] = z;
        }
    }
    __syncthreads();
    T2* R = stockham_lines<T2>(A, B, L, pitch, pl, Wl, -1);
    T2* S = (R == A) ? B : A;
    for (int idx = threadIdx.x; idx < L * M; idx += blockDim.x) {
        int line = idx / M, j = idx - line * M;
        R[line * pitch + j] = cmul(R[line * pitch + j], bhat[j]);
    }
    __syncthreads();
    R = stockham_lines<T2>(R, S, L, pitch, pl, Wl, +1);
    const T invM = (T)1 / (T)M;
    for (int idx = threadIdx.x; idx < L * n; idx += blockDim.x) {
        int line = idx / n, j = idx - line * n;
        T2 w = chirp[j];
        if (sign > 0) w = cconj(w);
        R[line * pitch + j] = cscale(cmul(R[line * pitch + j], w), invM);
    }
    __syncthreads();
    return R;
}

/* ------------------------------------------------------------------------- */
/*  Per-problem constants of the operator                                      */
/* ------------------------------------------------------------------------- */
struct ProbDev {
    double t0, t1, t2;   /* inner phase: (t0 + t1 b_f + t2 c_f)                 */
    double s0, s1;       /* middle phase: (s0 + s1 c_f)                         */
    double r0;           /* outer phase                                         */
    int gamma;           /* 1 when Nd = N-1 (mode f = 0 removed)                */
    int pad;
};

struct GridDev {
    int na, nb, nc;
    int N;
    int general;
    double ha, hb, hc;
};

/* Fourier symbols Lambda_x, Lambda_y, Lambda_z of mode (a_f, b_f, c_f)
 * (FAME_Matrix_Lambdas.cu:89-132 Simple, :309-375 General), as exact
 * functions of the mode index and the per-k phases.                          */
__device__ __forceinline__ void lambdas(const GridDev& g, const ProbDev& p, int af, int bf, int cf,
                                        double2& lx, double2& ly, double2& lz) {
    double pa = (double)af + (p.t0 + p.t1 * (double)bf + p.t2 * (double)cf);
    double pb = (double)bf + (p.s0 + p.s1 * (double)cf);
    double pc = (double)cf + p.r0;
    double2 ea = expipi(2.0 * pa / (double)g.na);
    double2 eb = expipi(2.0 * pb / (double)g.nb);
    double2 ec = expipi(2.0 * pc / (double)g.nc);
    double2 La = make_double2((ea.x - 1.0) / g.ha, ea.y / g.ha);
    double2 Lb = make_double2((eb.x - 1.0) / g.hb, eb.y / g.hb);
    double2 Lc = make_double2((ec.x - 1.0) / g.hc, ec.y / g.hc);
    if (g.general) { lx = Lc; ly = Lb; lz = La; }
    else { lx = La; ly = Lb; lz = Lc; }
}

__device__ __forceinline__ double2 cdiv_real(double2 a, double s) { return make_double2(a.x / s, a.y / s); }

/* Orthonormal basis Pi = [Pi1 Pi2] of Lambda-perp for one mode and |Lambda|
 * (Simple: FAME_Matrix_Lambdas.cu:184-215, General: :441-461).  Both radicands
 * are real and non-negative (Cauchy-Schwarz), so real square roots are used. */
__device__ __forceinline__ void pi_basis(const GridDev& g, double2 lx, double2 ly, double2 lz,
                                         double2 p1[3], double2 p2[3], double& lq_sqrt) {
    double Lq = lx.x * lx.x + lx.y * lx.y + ly.x * ly.x + ly.y * ly.y + lz.x * lz.x + lz.y * lz.y;
    lq_sqrt = sqrt(Lq);
    if (!g.general) {
        const double al = 3.0, be = 5.0;
        double2 t1 = csub(cscale(lz, be), ly);
        double2 t2 = csub(lx, cscale(lz, al));
        double2 t3 = csub(cscale(ly, al), cscale(lx, be));
        double n2 = t1.x * t1.x + t1.y * t1.y + t2.x * t2.x + t2.y * t2.y + t3.x * t3.x + t3.y * t3.y;
        double s1 = sqrt(n2 * Lq), s2 = sqrt(n2);
        p1[0] = cdiv_real(csub(cmulc(t3, ly), cmulc(t2, lz)), s1);
        p1[1] = cdiv_real(csub(cmulc(t1, lz), cmulc(t3, lx)), s1);
        p1[2] = cdiv_real(csub(cmulc(t2, lx), cmulc(t1, ly)), s1);
        p2[0] = cdiv_real(cconj(t1), s2);
        p2[1] = cdiv_real(cconj(t2), s2);
        p2[2] = cdiv_real(cconj(t3), s2);
    } else {
        double2 Ls = cadd(cadd(lx, ly), lz);
        double Lp = Ls.x * Ls.x + Ls.y * Ls.y;
        double n1 = sqrt(3.0 * Lq * Lq - Lq * Lp);
        double n2 = sqrt(3.0 * Lq - Lp);
        double2 lxs = cmulc(lx, Ls), lys = cmulc(ly, Ls), lzs = cmulc(lz, Ls);
        p1[0] = cdiv_real(make_double2(Lq - lxs.x, -lxs.y), n1);
        p1[1] = cdiv_real(make_double2(Lq - lys.x, -lys.y), n1);
        p1[2] = cdiv_real(make_double2(Lq - lzs.x, -lzs.y), n1);
        p2[0] = cdiv_real(csub(cconj(lz), cconj(ly)), n2);
        p2[1] = cdiv_real(csub(cconj(lx), cconj(lz)), n2);
        p2[2] = cdiv_real(csub(cconj(ly), cconj(lx)), n2);
    }
}

}  // namespace fame
