/*
 * fame_device.hpp -- device-side building blocks for gfx950:
 * complex arithmetic, the batched-line FFT engine (LDS Stockham autosort with
 * radices 2,3,4,5,7,11,13 and Bluestein chirp-z for lengths with larger prime
 * factors) and the on-the-fly Fourier symbols / Pi bases of the null-space-free
 * Maxwell operator.
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

/* (exp(2 pi i x) - 1) * inv_h evaluated as 2i sin(pi x) exp(i pi x) * inv_h:
 * one sincospi, no cancellation for small symbols (the reference's Lambda,
 * FAME_Matrix_Lambdas.cu:91-92, 312-313, 340-341, 373-374)                   */
__device__ __forceinline__ double2 lam_of(double x, double inv_h) {
    double s, c;
    sincospi(x, &s, &c);
    return make_double2(-2.0 * s * s * inv_h, 2.0 * s * c * inv_h);
}

/* exact quotient n / d for n*d < 2^32 from m = ceil(2^32 / d) */
__device__ __forceinline__ int fdiv(int n, unsigned long long m) { return (int)(((unsigned long long)(unsigned)n * m) >> 32); }

/* ------------------------------------------------------------------------- */
/*  FFT plan as seen by a kernel                                              */
/* ------------------------------------------------------------------------- */
#define FAME_MAX_PASSES 10
struct PassDev {
    int R, Ns, nb, tstep;           /* radix, span so far, butterflies per line, twiddle step */
    unsigned long long mNs, mNb;    /* magic multipliers for / Ns and / nb              */
};
struct FftPlanDev {
    int n;          /* transform length                                        */
    int M;          /* Stockham length: n (direct) or the Bluestein length     */
    int blue;       /* 1 when Bluestein chirp-z is used                        */
    int nrad;
    PassDev pass[FAME_MAX_PASSES];
    unsigned long long mM, mN;      /* magic for / M and / n                   */
    const void* W;      /* M entries exp(-2 pi i m / M), precision of the kernel */
    const void* chirp;  /* n entries exp(-i pi m^2 / n)                          */
    const void* bhat_m; /* M entries: FFT_M of the kernel for the forward sign   */
    const void* bhat_p; /* M entries: same for the inverse sign                  */
};

/* LDS-resident tables of a plan: W (M), and for Bluestein chirp (n), bhat_m, bhat_p (M each) */
template <typename T2>
struct PlanLds {
    T2* W; T2* chirp; T2* bm; T2* bp;
};
__host__ __device__ inline int plan_table_len(const FftPlanDev& p) { return p.blue ? (3 * p.M + p.n) : p.M; }

template <typename T2>
__device__ __forceinline__ PlanLds<T2> load_plan_tables(const FftPlanDev& pl, T2* base) {
    PlanLds<T2> t;
    t.W = base;
    const T2* W = reinterpret_cast<const T2*>(pl.W);
    for (int i = threadIdx.x; i < pl.M; i += blockDim.x) t.W[i] = W[i];
    if (pl.blue) {
        t.chirp = base + pl.M;
        t.bm = t.chirp + pl.n;
        t.bp = t.bm + pl.M;
        const T2* c = reinterpret_cast<const T2*>(pl.chirp);
        const T2* bm = reinterpret_cast<const T2*>(pl.bhat_m);
        const T2* bp = reinterpret_cast<const T2*>(pl.bhat_p);
        for (int i = threadIdx.x; i < pl.n; i += blockDim.x) t.chirp[i] = c[i];
        for (int i = threadIdx.x; i < pl.M; i += blockDim.x) { t.bm[i] = bm[i]; t.bp[i] = bp[i]; }
    } else {
        t.chirp = t.bm = t.bp = nullptr;
    }
    return t;
}

/* small DFT of radix R in registers: o[q] = sum_t v[t] wr[(q t) mod R] */
template <typename T2, int R>
__device__ __forceinline__ void small_dft(const T2 (&v)[R], T2 (&o)[R], const T2 (&wr)[R], int sign) {
    if (R == 2) {
        o[0] = cadd(v[0], v[1]); o[1] = csub(v[0], v[1]);
    } else if (R == 4) {
        T2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]), a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
        T2 j3;
        if (sign < 0) { j3.x = a3.y; j3.y = -a3.x; } else { j3.x = -a3.y; j3.y = a3.x; }
        o[0] = cadd(a0, a2); o[2] = csub(a0, a2); o[1] = cadd(a1, j3); o[3] = csub(a1, j3);
    } else if (R == 3) {
        typedef decltype(v[0].x) T;
        T2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        T2 m; m.x = v[0].x - (T)0.5 * s.x; m.y = v[0].y - (T)0.5 * s.y;
        /* i*sign*(sqrt(3)/2)*d */
        const T h = sign < 0 ? (T)-0.86602540378443864676 : (T)0.86602540378443864676;
        T2 r; r.x = -h * d.y; r.y = h * d.x;
        o[0] = cadd(v[0], s); o[1] = cadd(m, r); o[2] = csub(m, r);
    } else {
#pragma unroll
        for (int q = 0; q < R; ++q) {
            T2 acc = v[0];
#pragma unroll
            for (int t = 1; t < R; ++t) acc = cadd(acc, cmul(v[t], wr[(q * t) % R]));
            o[q] = acc;
        }
    }
}

/* One Stockham radix-R pass over L lines held in LDS (src -> dst).
 * Butterfly j of a line: inputs src[j + t*nb], twiddle W_{Ns R}^{t k},
 * outputs dst[(j/Ns)*Ns*R + k + t*Ns], k = j mod Ns. sign < 0: forward. */
template <typename T2, int R>
__device__ __forceinline__ void stockham_pass(const T2* __restrict__ src, T2* __restrict__ dst, int L, int pitch,
                                              const PassDev& ps, const T2* __restrict__ W, int sign) {
    const int nb = ps.nb, Ns = ps.Ns, tstep = ps.tstep;
    T2 wr[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
        T2 w = W[m * nb];
        if (sign > 0) w = cconj(w);
        wr[m] = w;
    }
    const int total = L * nb;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int line = fdiv(idx, ps.mNb);
        const int j = idx - line * nb;
        const int qd = fdiv(j, ps.mNs);
        const int k = j - qd * Ns;
        const T2* a = src + line * pitch;
        T2* b = dst + line * pitch;
        T2 v[R];
        v[0] = a[j];
        if (k > 0) {
#pragma unroll
            for (int t = 1; t < R; ++t) {
                T2 w = W[t * k * tstep];
                if (sign > 0) w = cconj(w);
                v[t] = cmul(a[j + t * nb], w);
            }
        } else {
#pragma unroll
            for (int t = 1; t < R; ++t) v[t] = a[j + t * nb];
        }
        T2 o[R];
        small_dft<T2, R>(v, o, wr, sign);
        const int base = qd * Ns * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) b[base + t * Ns] = o[t];
    }
}

/* Full Stockham transform of L lines of length pl.M held in LDS buffer A (B is
 * scratch of the same size).  Returns the buffer holding the result.          */
template <typename T2>
__device__ T2* stockham_lines(T2* A, T2* B, int L, int pitch, const FftPlanDev& pl, const T2* Wl, int sign) {
    for (int p = 0; p < pl.nrad; ++p) {
        const PassDev& ps = pl.pass[p];
        switch (ps.R) {
            case 2: stockham_pass<T2, 2>(A, B, L, pitch, ps, Wl, sign); break;
            case 3: stockham_pass<T2, 3>(A, B, L, pitch, ps, Wl, sign); break;
            case 4: stockham_pass<T2, 4>(A, B, L, pitch, ps, Wl, sign); break;
            case 5: stockham_pass<T2, 5>(A, B, L, pitch, ps, Wl, sign); break;
            case 7: stockham_pass<T2, 7>(A, B, L, pitch, ps, Wl, sign); break;
            case 11: stockham_pass<T2, 11>(A, B, L, pitch, ps, Wl, sign); break;
            case 13: stockham_pass<T2, 13>(A, B, L, pitch, ps, Wl, sign); break;
            default: break;
        }
        __syncthreads();
        T2* t = A; A = B; B = t;
    }
    return A;
}

/* Transform L lines of length pl.n that sit (unpadded) at the front of each
 * LDS line of `pitch` >= pl.M elements.  Handles both direct and Bluestein
 * plans.  Returns the buffer (A or B) holding the n outputs of each line.
 * Caller must have synchronised after filling A.                              */
template <typename T2, typename T>
__device__ T2* fft_lines(T2* A, T2* B, int L, int pitch, const FftPlanDev& pl, const PlanLds<T2>& tb, int sign) {
    if (!pl.blue) return stockham_lines<T2>(A, B, L, pitch, pl, tb.W, sign);
    const T2* bhat = sign < 0 ? tb.bm : tb.bp;
    const int n = pl.n, M = pl.M;
    /* pre-chirp and zero padding: a[j] = x[j] w[j], w = exp(sign i pi j^2/n) */
    for (int idx = threadIdx.x; idx < L * M; idx += blockDim.x) {
        const int line = fdiv(idx, pl.mM), j = idx - line * M;
        T2* a = A + line * pitch;
        if (j < n) {
            T2 w = tb.chirp[j];
            if (sign > 0) w = cconj(w);
            a[j] = cmul(a[j], w);
        } else {
            T2 z; z.x = 0; z.y = 0;
            a[j] = z;
        }
    }
    __syncthreads();
    T2* R = stockham_lines<T2>(A, B, L, pitch, pl, tb.W, -1);
    T2* S = (R == A) ? B : A;
    for (int idx = threadIdx.x; idx < L * M; idx += blockDim.x) {
        const int line = fdiv(idx, pl.mM), j = idx - line * M;
        R[line * pitch + j] = cmul(R[line * pitch + j], bhat[j]);
    }
    __syncthreads();
    R = stockham_lines<T2>(R, S, L, pitch, pl, tb.W, +1);
    const T invM = (T)1 / (T)M;
    for (int idx = threadIdx.x; idx < L * n; idx += blockDim.x) {
        const int line = fdiv(idx, pl.mN), j = idx - line * n;
        T2 w = tb.chirp[j];
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
    double iha, ihb, ihc;     /* 1/h */
    unsigned long long m_na, m_nb;   /* magic for / na, / nb */
};

/* Line constants of the inner axis for line gl = c*nb + b: the symbols of the
 * two outer axes and the inner phase (FAME_Matrix_Lambdas.cu:302-375 General,
 * :82-132 Simple)                                                             */
struct LineSym { double2 Lb, Lc; double phi; };

__device__ __forceinline__ LineSym line_sym(const GridDev& g, const ProbDev& p, int bf, int cf) {
    LineSym s;
    s.phi = p.t0 + p.t1 * (double)bf + p.t2 * (double)cf;
    s.Lb = lam_of(((double)bf + (p.s0 + p.s1 * (double)cf)) / (double)g.nb, g.ihb);
    s.Lc = lam_of(((double)cf + p.r0) / (double)g.nc, g.ihc);
    return s;
}

/* Lambda_x, Lambda_y, Lambda_z of mode a on a line (physical assignment of the axes) */
__device__ __forceinline__ void lambdas_line(const GridDev& g, const LineSym& s, int af,
                                             double2& lx, double2& ly, double2& lz) {
    const double2 La = lam_of(((double)af + s.phi) / (double)g.na, g.iha);
    if (g.general) { lx = s.Lc; ly = s.Lb; lz = La; }
    else { lx = La; ly = s.Lb; lz = s.Lc; }
}

/* Fourier symbols of mode (a_f, b_f, c_f) */
__device__ __forceinline__ void lambdas(const GridDev& g, const ProbDev& p, int af, int bf, int cf,
                                        double2& lx, double2& ly, double2& lz) {
    LineSym s = line_sym(g, p, bf, cf);
    lambdas_line(g, s, af, lx, ly, lz);
}

/* Orthonormal basis Pi = [Pi1 Pi2] of Lambda-perp for one mode
 * (Simple: FAME_Matrix_Lambdas.cu:184-215, General: :441-461).  Both radicands
 * are real and non-negative (Cauchy-Schwarz); reciprocal square roots replace
 * the reference's complex square roots and divisions.                         */
__device__ __forceinline__ void pi_basis(const GridDev& g, double2 lx, double2 ly, double2 lz,
                                         double2 p1[3], double2 p2[3]) {
    const double Lq = lx.x * lx.x + lx.y * lx.y + ly.x * ly.x + ly.y * ly.y + lz.x * lz.x + lz.y * lz.y;
    if (!g.general) {
        const double al = 3.0, be = 5.0;
        double2 t1 = csub(cscale(lz, be), ly);
        double2 t2 = csub(lx, cscale(lz, al));
        double2 t3 = csub(cscale(ly, al), cscale(lx, be));
        const double n2 = t1.x * t1.x + t1.y * t1.y + t2.x * t2.x + t2.y * t2.y + t3.x * t3.x + t3.y * t3.y;
        const double i2 = rsqrt(n2);
        const double i1 = i2 * rsqrt(Lq);
        p1[0] = cscale(csub(cmulc(t3, ly), cmulc(t2, lz)), i1);
        p1[1] = cscale(csub(cmulc(t1, lz), cmulc(t3, lx)), i1);
        p1[2] = cscale(csub(cmulc(t2, lx), cmulc(t1, ly)), i1);
        p2[0] = cscale(cconj(t1), i2);
        p2[1] = cscale(cconj(t2), i2);
        p2[2] = cscale(cconj(t3), i2);
    } else {
        const double2 Ls = cadd(cadd(lx, ly), lz);
        const double Lp = Ls.x * Ls.x + Ls.y * Ls.y;
        const double i1 = rsqrt(3.0 * Lq * Lq - Lq * Lp);
        const double i2 = rsqrt(3.0 * Lq - Lp);
        const double2 lxs = cmulc(lx, Ls), lys = cmulc(ly, Ls), lzs = cmulc(lz, Ls);
        p1[0] = make_double2((Lq - lxs.x) * i1, -lxs.y * i1);
        p1[1] = make_double2((Lq - lys.x) * i1, -lys.y * i1);
        p1[2] = make_double2((Lq - lzs.x) * i1, -lzs.y * i1);
        p2[0] = cscale(csub(cconj(lz), cconj(ly)), i2);
        p2[1] = cscale(csub(cconj(lx), cconj(lz)), i2);
        p2[2] = cscale(csub(cconj(ly), cconj(lx)), i2);
    }
}

}  // namespace fame
