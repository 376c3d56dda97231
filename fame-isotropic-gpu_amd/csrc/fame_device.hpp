/*
 * fame_device.hpp -- device-side building blocks for gfx950:
 * complex arithmetic, the batched-line FFT engine (in-place LDS Stockham
 * autosort with radices 2,3,4,5,7,8,11,13 and Bluestein chirp-z for lengths
 * with larger prime factors) and the on-the-fly Fourier symbols / Pi bases of the null-space-free
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


/* Workgroups go to the 8 XCDs round robin by linear block id; with gridDim.x a
 * multiple of 8, block x runs on XCD x % 8.  xcd_tile() renumbers the blocks so
 * that each XCD gets a run of consecutive tiles: tiles that share boundary
 * cache lines then meet in the same L2 instead of fetching them twice.     */
__device__ __forceinline__ int xcd_tile() {
    return (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
}
__host__ __device__ inline int xcd_round(int nblocks) { return (nblocks + 7) & ~7; }

/* quotient n / d from m = ceil(2^32 / d) (m = 0 encodes d = 1); exact for n*d < 2^32 */
__device__ __forceinline__ int udiv(int n, unsigned m) { return m ? (int)__umulhi((unsigned)n, m) : n; }

/* ------------------------------------------------------------------------- */
/*  Line transforms                                                           */
/* ------------------------------------------------------------------------- */
/* A kernel transforms `nl` lines of length M held in one LDS buffer (line pitch
 * P, odd, so that column accesses of consecutive lines hit different banks).
 * The work is a list of stages run by the whole block:
 *   R > 0 : one in-place Stockham autosort pass of radix R.  Every thread loads
 *           its butterflies' inputs into registers (applying the optional input
 *           multiplier and the twiddles), computes them, and after a barrier
 *           writes the outputs back to the same buffer.
 *   R = 0 : an elementwise step supplied by the kernel (1/eps, curl-curl, ...).
 * Bluestein chirp-z (lengths with a prime factor above 13) appears as two
 * length-M transforms in the list: the pre-chirp is the input multiplier of the
 * first pass, the kernel spectrum bhat that of the first inverse pass, and the
 * post-chirp (with 1/M) is applied by the kernel's store ("post").          */
#define FAME_MAX_STAGES 24
#define FFT_THREADS 256

/* A block-strided loop over idx = threadIdx.x + k * blockDim.x < total in two
 * phases per group of U elements: ld(idx) issues the global loads of each
 * element and returns them, then use(idx, value) consumes them.  The U loads
 * of a thread are in flight together instead of one load-wait-store round
 * trip per element, which is what keeps HBM busy in the strided gathers.   */
template <int U, typename Ld, typename Use>
__device__ __forceinline__ void batched_loop(int total, Ld ld, Use use) {
    typedef decltype(ld(0)) R;
    const int nt = (int)blockDim.x;
    for (int i0 = (int)threadIdx.x; i0 < total; i0 += U * nt) {
        R r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int idx = i0 + u * nt;
            if (idx < total) r[u] = ld(idx);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int idx = i0 + u * nt;
            if (idx < total) use(idx, r[u]);
        }
    }
}


enum { PRE_NONE = 0, PRE_CHIRP = 1, PRE_CHIRP_CONJ = 2, PRE_BM = 3, PRE_BP = 4 };
enum { POST_NONE = 0, POST_CHIRP = 1, POST_CHIRP_CONJ = 2 };

#define DFT_MFMA (-1)    /* StageDev.R of a direct DFT on the matrix cores       */
#define DFT_KS_MAX 16    /* K steps of 4 -> direct DFT lengths up to 127          */

struct StageDev {
    int R;              /* radix; 0 = kernel-supplied elementwise step; DFT_MFMA */
    int Ns;             /* product of the radices of the earlier passes          */
    int nb;             /* butterflies per line, M / R                          */
    int tstep;          /* twiddle table step, M / (Ns R)                       */
    unsigned mNs, mNb;  /* udiv magics for / Ns and / nb                        */
    int sign;           /* -1 forward (exp(-i..)), +1 inverse                  */
    int pre;            /* PRE_*: multiplier of the pass inputs (first passes) */
    int dn, dm, dQ, dK; /* direct DFT: length n, pairs m = (n-1)/2, outputs Q = n/2+1, K = m rounded up to 4 */
    int dqt, pad;       /* direct DFT: 16-row output tiles ceil(Q/16) (one per wave)                          */
};

struct LinePlanDev {
    int n;              /* transform length                                    */
    int M;              /* Stockham length (n, or the Bluestein length)        */
    int blue;           /* 1 for Bluestein                                     */
    int nst;            /* stages in st[]                                      */
    int post;           /* POST_*: multiplier the store applies                */
    int ntab;           /* LDS table entries: M (W) or 4M (W, chirp, bm, bp)   */
    int pad;
    unsigned mM, mN;    /* udiv magics for / M and / n                         */
    const void* W;      /* M entries exp(-2 pi i m / M) in the kernel precision */
    const void* chirp;  /* M entries exp(-i pi m^2 / n) for m < n, 0 beyond     */
    const void* bm;     /* M entries: DFT_M of the chirp kernel, forward sign   */
    const void* bp;     /* M entries: same for the inverse sign                 */
    StageDev st[FAME_MAX_STAGES];
};

template <typename T2>
struct Tabs { const T2* W; const T2* chirp; const T2* bm; const T2* bp; };

/* LDS layout of a line-transform block: tables (ntab), then the kernel's own
 * entries and lines */
__host__ __device__ inline int lds_head(const LinePlanDev& pl) { return pl.ntab; }

/* copy a plan's tables into LDS at `base` (caller synchronises) */
template <typename T2>
__device__ __forceinline__ Tabs<T2> load_tabs(const LinePlanDev& pl, T2* base) {
    const int M = pl.M;
    const T2* src[4] = {reinterpret_cast<const T2*>(pl.W), reinterpret_cast<const T2*>(pl.chirp),
                        reinterpret_cast<const T2*>(pl.bm), reinterpret_cast<const T2*>(pl.bp)};
    for (int i = threadIdx.x; i < pl.ntab; i += blockDim.x) {
        const int t = i >= M ? (i >= 2 * M ? (i >= 3 * M ? 3 : 2) : 1) : 0;
        base[i] = src[t][i - t * M];
    }
    Tabs<T2> tb;
    tb.W = base;
    tb.chirp = base + M; tb.bm = base + 2 * M; tb.bp = base + 3 * M;
    return tb;
}

/* store-side multiplier of output j of a line */
template <typename T2, typename T>
__device__ __forceinline__ T2 post_mul(const LinePlanDev& pl, const Tabs<T2>& tb, int j, T2 v) {
    if (pl.post == POST_NONE) return v;
    T2 w = tb.chirp[j];
    if (pl.post == POST_CHIRP_CONJ) w.y = -w.y;
    const T invM = (T)1 / (T)pl.M;
    return cscale(cmul(v, w), invM);
}

/* cos / sin (2 pi k / R), k = 1..(R-1)/2, for the symmetric odd-radix DFTs */
template <int R> struct OddTrig;
template <> struct OddTrig<3> {
    static constexpr double c[1] = {-0.5};
    static constexpr double s[1] = {0.8660254037844386};
};
template <> struct OddTrig<5> {
    static constexpr double c[2] = {0.30901699437494745, -0.8090169943749475};
    static constexpr double s[2] = {0.9510565162951535, 0.5877852522924731};
};
template <> struct OddTrig<7> {
    static constexpr double c[3] = {0.6234898018587335, -0.2225209339563144, -0.9009688679024191};
    static constexpr double s[3] = {0.7818314824680298, 0.9749279121818236, 0.4338837391175581};
};
template <> struct OddTrig<11> {
    static constexpr double c[5] = {0.8412535328311812, 0.41541501300188644, -0.14231483827328514,
                                    -0.6548607339452851, -0.9594929736144974};
    static constexpr double s[5] = {0.5406408174555976, 0.9096319953545183, 0.9898214418809327,
                                    0.7557495743542583, 0.28173255684142967};
};
template <> struct OddTrig<13> {
    static constexpr double c[6] = {0.8854560256532099, 0.5680647467311558, 0.12053668025532305,
                                    -0.3546048870425356, -0.7485107481711011, -0.970941817426052};
    static constexpr double s[6] = {0.46472317204376856, 0.8229838658936564, 0.992708874098054,
                                    0.9350162426854148, 0.6631226582407952, 0.23931566428755777};
};

/* in-place DFT of odd length R = 2m+1 with the pair sums / differences
 * s_t = v_t + v_{R-t}, d_t = v_t - v_{R-t}:
 *   X_q = v_0 + sum_t cos(2 pi q t/R) s_t  -/+  i sum_t sin(2 pi q t/R) d_t
 * (the sign of the i-term is the transform sign), X_{R-q} with the opposite.  */
template <typename T2, int R>
__device__ __forceinline__ void dft_odd(T2 (&v)[R], bool inv) {
    typedef decltype(v[0].x) T;
    constexpr int m = (R - 1) / 2;
    T2 sp[m], sm[m];
    const T2 v0 = v[0];
    T2 o0 = v0;
#pragma unroll
    for (int t = 1; t <= m; ++t) {
        sp[t - 1] = cadd(v[t], v[R - t]);
        sm[t - 1] = csub(v[t], v[R - t]);
        o0 = cadd(o0, sp[t - 1]);
    }
#pragma unroll
    for (int q = 1; q <= m; ++q) {
        T2 A = v0, B;
        B.x = 0; B.y = 0;
#pragma unroll
        for (int t = 1; t <= m; ++t) {
            const int r = (q * t) % R;
            const int k = r <= m ? r : R - r;
            const T c = (T)OddTrig<R>::c[k - 1];
            const T sn = r <= m ? (T)OddTrig<R>::s[k - 1] : -(T)OddTrig<R>::s[k - 1];
            A.x += c * sp[t - 1].x; A.y += c * sp[t - 1].y;
            B.x += sn * sm[t - 1].x; B.y += sn * sm[t - 1].y;
        }
        /* forward: X_q = A - iB, X_{R-q} = A + iB */
        T2 lo, hi;
        if (!inv) { lo.x = A.x + B.y; lo.y = A.y - B.x; hi.x = A.x - B.y; hi.y = A.y + B.x; }
        else      { lo.x = A.x - B.y; lo.y = A.y + B.x; hi.x = A.x + B.y; hi.y = A.y - B.x; }
        v[q] = lo;
        v[R - q] = hi;
    }
    v[0] = o0;
}

template <typename T2>
__device__ __forceinline__ void dft4(T2& v0, T2& v1, T2& v2, T2& v3, bool inv) {
    const T2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = csub(v1, v3);
    T2 j3;   /* -i a3 forward, +i a3 inverse */
    if (!inv) { j3.x = a3.y; j3.y = -a3.x; } else { j3.x = -a3.y; j3.y = a3.x; }
    v0 = cadd(a0, a2); v2 = csub(a0, a2); v1 = cadd(a1, j3); v3 = csub(a1, j3);
}

/* cos / sin (2 pi e / R), e = 0..R-1, for the inner twiddles of the compound radices */
template <int R> struct CTab;
template <> struct CTab<6> {
    static constexpr double c[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
    static constexpr double s[6] = {0.0, 0.8660254037844386, 0.8660254037844386, 0.0, -0.8660254037844386, -0.8660254037844386};
};
template <> struct CTab<9> {
    static constexpr double c[9] = {1.0, 0.766044443118978, 0.17364817766693036, -0.5, -0.9396926207859084, -0.9396926207859084, -0.5, 0.17364817766693036, 0.766044443118978};
    static constexpr double s[9] = {0.0, 0.6427876096865394, 0.984807753012208, 0.8660254037844386, 0.3420201433256687, -0.3420201433256687, -0.8660254037844386, -0.984807753012208, -0.6427876096865394};
};
template <> struct CTab<10> {
    static constexpr double c[10] = {1.0, 0.8090169943749475, 0.30901699437494745, -0.30901699437494745, -0.8090169943749475, -1.0, -0.8090169943749475, -0.30901699437494745, 0.30901699437494745, 0.8090169943749475};
    static constexpr double s[10] = {0.0, 0.5877852522924731, 0.9510565162951535, 0.9510565162951535, 0.5877852522924731, 0.0, -0.5877852522924731, -0.9510565162951535, -0.9510565162951535, -0.5877852522924731};
};
template <> struct CTab<12> {
    static constexpr double c[12] = {1.0, 0.8660254037844386, 0.5, 0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5, 0.0, 0.5, 0.8660254037844386};
    static constexpr double s[12] = {0.0, 0.5, 0.8660254037844386, 1.0, 0.8660254037844386, 0.5, 0.0, -0.5, -0.8660254037844386, -1.0, -0.8660254037844386, -0.5};
};
template <> struct CTab<15> {
    static constexpr double c[15] = {1.0, 0.9135454576426009, 0.6691306063588582, 0.30901699437494745, -0.10452846326765347, -0.5, -0.8090169943749475, -0.9781476007338057, -0.9781476007338057, -0.8090169943749475, -0.5, -0.10452846326765347, 0.30901699437494745, 0.6691306063588582, 0.9135454576426009};
    static constexpr double s[15] = {0.0, 0.4067366430758002, 0.7431448254773942, 0.9510565162951535, 0.9945218953682733, 0.8660254037844386, 0.5877852522924731, 0.20791169081775934, -0.20791169081775934, -0.5877852522924731, -0.8660254037844386, -0.9945218953682733, -0.9510565162951535, -0.7431448254773942, -0.4067366430758002};
};
template <> struct CTab<16> {
    static constexpr double c[16] = {1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898, 0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867};
    static constexpr double s[16] = {0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867, 1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898, 0.0, -0.3826834323650898, -0.7071067811865476, -0.9238795325112867, -1.0, -0.9238795325112867, -0.7071067811865476, -0.3826834323650898};
};

/* v * W_R^e with W_R = exp(-2 pi i / R) (forward) or its conjugate (INV); e is a
 * constant once the callers' loops are unrolled; 0, R/4, R/2, 3R/4 are exact */
template <typename T2, int R, bool INV>
__device__ __forceinline__ T2 tw_const(T2 v, int e) {
    typedef decltype(v.x) T;
    T2 o;
    if (e == 0) return v;
    if (2 * e == R) { o.x = -v.x; o.y = -v.y; return o; }
    if (4 * e == R) {          /* forward -i, inverse +i */
        if (!INV) { o.x = v.y; o.y = -v.x; } else { o.x = -v.y; o.y = v.x; }
        return o;
    }
    if (4 * e == 3 * R) {      /* forward +i, inverse -i */
        if (!INV) { o.x = -v.y; o.y = v.x; } else { o.x = v.y; o.y = -v.x; }
        return o;
    }
    const T c = (T)CTab<R>::c[e], sn = INV ? (T)CTab<R>::s[e] : -(T)CTab<R>::s[e];
    o.x = v.x * c - v.y * sn;
    o.y = v.x * sn + v.y * c;
    return o;
}

template <typename T2, int R>
__device__ __forceinline__ void dft(T2 (&v)[R], bool inv);

/* DFT of length R = R1 R2 in registers: input n = R2 n1 + n2, output
 * k = k1 + R1 k2,  X_k = sum_n2 W_R2^(n2 k2) W_R^(n2 k1) sum_n1 x_n W_R1^(n1 k1).
 * One radix-R pass replaces two passes (and an LDS round trip) of R1 and R2. */
template <typename T2, int R1, int R2, bool INV>
__device__ __forceinline__ void dft_pq(T2 (&v)[R1 * R2]) {
    constexpr int R = R1 * R2;
    T2 y[R2][R1];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
        T2 a[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[R2 * n1 + n2];
        dft<T2, R1>(a, INV);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) y[n2][k1] = tw_const<T2, R, INV>(a[k1], (n2 * k1) % R);
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
        T2 b[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) b[n2] = y[n2][k1];
        dft<T2, R2>(b, INV);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
    }
}

template <typename T2, int R>
__device__ __forceinline__ void dft(T2 (&v)[R], bool inv) {
    typedef decltype(v[0].x) T;
    if constexpr (R == 2) {
        const T2 a = v[0];
        v[0] = cadd(a, v[1]); v[1] = csub(a, v[1]);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3], inv);
    } else if constexpr (R == 8) {
        /* split into even / odd outputs: a_k = v_k + v_{k+4} -> X_{2q},
         * b_k = (v_k - v_{k+4}) W8^k -> X_{2q+1} */
        const T h = (T)0.7071067811865476;
        T2 a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) { a[k] = cadd(v[k], v[k + 4]); b[k] = csub(v[k], v[k + 4]); }
        T2 t;
        if (!inv) {
            t = b[1]; b[1].x = h * (t.x + t.y); b[1].y = h * (t.y - t.x);
            t = b[2]; b[2].x = t.y; b[2].y = -t.x;
            t = b[3]; b[3].x = h * (t.y - t.x); b[3].y = -h * (t.x + t.y);
        } else {
            t = b[1]; b[1].x = h * (t.x - t.y); b[1].y = h * (t.x + t.y);
            t = b[2]; b[2].x = -t.y; b[2].y = t.x;
            t = b[3]; b[3].x = -h * (t.x + t.y); b[3].y = h * (t.x - t.y);
        }
        dft4(a[0], a[1], a[2], a[3], inv);
        dft4(b[0], b[1], b[2], b[3], inv);
#pragma unroll
        for (int q = 0; q < 4; ++q) { v[2 * q] = a[q]; v[2 * q + 1] = b[q]; }
    } else if constexpr (R == 6) {
        if (inv) dft_pq<T2, 3, 2, true>(v); else dft_pq<T2, 3, 2, false>(v);
    } else if constexpr (R == 9) {
        if (inv) dft_pq<T2, 3, 3, true>(v); else dft_pq<T2, 3, 3, false>(v);
    } else if constexpr (R == 10) {
        if (inv) dft_pq<T2, 5, 2, true>(v); else dft_pq<T2, 5, 2, false>(v);
    } else if constexpr (R == 12) {
        if (inv) dft_pq<T2, 4, 3, true>(v); else dft_pq<T2, 4, 3, false>(v);
    } else if constexpr (R == 15) {
        if (inv) dft_pq<T2, 5, 3, true>(v); else dft_pq<T2, 5, 3, false>(v);
    } else if constexpr (R == 16) {
        if (inv) dft_pq<T2, 4, 4, true>(v); else dft_pq<T2, 4, 4, false>(v);
    } else {
        dft_odd<T2, R>(v, inv);
    }
}

/* butterflies per thread and pass: at most MAXE registers-resident elements */
__host__ __device__ constexpr int nbt_of(int R, int MAXE) { return MAXE / R >= 1 ? MAXE / R : 1; }

/* one in-place radix-R Stockham pass over nl lines.
 * Butterfly j of a line (k = j mod Ns): inputs x[j + t nb] (times the input
 * multiplier, then W_{Ns R}^{t k}), outputs y[(j / Ns) Ns R + k + t Ns].       */
/* A `mid` step (R = 0 stage) may instead be folded into the loads of the pass
 * that follows it: a Mid type with `fold = true` provides mul(line, m, v), the
 * step's value for element m of a line, and run_stages then skips the step's
 * own LDS round trip and barrier.                                            */
template <typename M, typename = void> struct FoldOf { static constexpr bool v = false; };
template <typename M> struct FoldOf<M, decltype((void)M::fold)> { static constexpr bool v = M::fold; };

struct NoFoldMul { };
/* Lane order of a pass.  Along the line (LF = false: butterfly j fastest,
 * then the line) or across the lines (LF = true: the line fastest, mNl the
 * udiv magic of nl).  With an odd line pitch P, lanes on consecutive lines at
 * one position hit distinct LDS banks, so LF makes every read and write of
 * every stage conflict-free once a wave spans enough lines (32 lines for
 * 8-byte elements); along the line, a stage with few butterflies per line
 * (nb = 8 or 12 at M = 96) puts neighbouring lines on the same banks (P = 1
 * mod 32) and its strided writes (Ns = 1) on every R-th bank.  LF also makes
 * a folded step's per-element loads consecutive across lanes.
 * returns the sum that a folded step's mul() accumulates (0 otherwise)     */
template <typename T2, int R, int MAXE, typename Mul = NoFoldMul, bool LF = false>
__device__ __forceinline__ double fft_pass(T2* S, int nl, int P, const StageDev& st, const Tabs<T2>& tb,
                                           const Mul& mul = Mul(), bool fold = false, unsigned mNl = 0) {
    double facc = 0.0;
    constexpr int NBT = nbt_of(R, MAXE);
    const int nb = st.nb, Ns = st.Ns, tstep = st.tstep;
    const int total = nl * nb;
    const bool inv = st.sign > 0;
    const T2* pre = st.pre == PRE_NONE ? nullptr
                  : (st.pre <= PRE_CHIRP_CONJ ? tb.chirp : (st.pre == PRE_BM ? tb.bm : tb.bp));
    const bool preconj = st.pre == PRE_CHIRP_CONJ;
    T2 v[NBT][R];
    int ob[NBT];
#pragma unroll
    for (int i = 0; i < NBT; ++i) {
        const int idx = threadIdx.x + i * FFT_THREADS;
        ob[i] = -1;
        if (idx < total) {
            int line, j;
            if constexpr (LF) {
                /* (line, j) do not depend on the stage here: keep the compiler
                 * from hoisting them for all NBT butterflies out of the stage
                 * loop, where they would stay live across it (and spill) */
                int id = idx;
                asm volatile("" : "+v"(id));
                j = udiv(id, mNl);
                line = id - j * nl;
            } else {
                line = udiv(idx, st.mNb);
                j = idx - line * nb;
            }
            const int qd = udiv(j, st.mNs);
            const int k = j - qd * Ns;
            const T2* a = S + line * P;
#pragma unroll
            for (int t = 0; t < R; ++t) v[i][t] = a[j + t * nb];
            if constexpr (FoldOf<Mul>::v) {
                if (fold) {
#pragma unroll
                    for (int t = 0; t < R; ++t) v[i][t] = mul.mul(line, j + t * nb, v[i][t], facc);
                }
            }
            if (pre) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    T2 w = pre[j + t * nb];
                    if (preconj) w.y = -w.y;
                    v[i][t] = cmul(v[i][t], w);
                }
            }
            if (k) {
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    T2 w = tb.W[t * k * tstep];
                    if (inv) w.y = -w.y;
                    v[i][t] = cmul(v[i][t], w);
                }
            }
            dft<T2, R>(v[i], inv);
            ob[i] = line * P + qd * Ns * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NBT; ++i) {
        if (ob[i] >= 0) {
#pragma unroll
            for (int t = 0; t < R; ++t) S[ob[i] + t * Ns] = v[i][t];
        }
    }
    __syncthreads();
    return facc;
}

/* ------------------------------------------------------------------------- */
/*  Direct DFT on the matrix cores                                            */
/* ------------------------------------------------------------------------- */
/* For a line x of length n (< 128), with s_t = x_t + x_{n-t}, d_t = x_t - x_{n-t}
 * (t = 1..m, m = (n-1)/2) and h = x_{n/2} for even n,
 *   X_q     = x_0 + (-1)^q h + A_q + sign i B_q,
 *   X_{n-q} = x_0 + (-1)^q h + A_q - sign i B_q,    q = 0..n/2,
 *   A_q = sum_t cos(2 pi q t / n) s_t,   B_q = sum_t sin(2 pi q t / n) d_t.
 * A and B are two real matrix products (Q x m) x (m x 2 columns per line) run
 * as 16x16x4 MFMAs: wave w owns output rows q = 16w..16w+15 and reads its cos /
 * sin operands from the twiddle table.  For 16 lines at a time, s_t and d_t
 * replace x_t and x_{n-t} in place in LDS (no workspace), and each lane reads
 * its B operands from its line.  About a quarter of the flops of the complex
 * DFT-GEMM. */
template <typename T> struct Mfma;
/* Operands of both 16x16x4 variants: lane l holds A[l % 16][l / 16] and
 * B[l / 16][l % 16].  Result row of accumulator element v in lane l (column
 * l % 16): 4 (l / 16) + v for f32, (l / 16) + 4 v for f64 (checked on gfx950). */
template <> struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kq, int v) { return 4 * kq + v; }
};
template <> struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kq, int v) { return kq + 4 * v; }
};


template <typename T2, bool GP>
__device__ __forceinline__ void dft_mfma(T2* S, int nl, int P, const StageDev& st, const Tabs<T2>& tb) {
    typedef decltype(S[0].x) T;
    typedef typename Mfma<T>::acc_t acc_t;
    const int n = st.dn, m = st.dm, Q = st.dQ, K = st.dK, nqt = st.dqt;
    const int nks = K >> 2;
    const bool inv = st.sign > 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    /* GP: waves take (output tile, line group) pairs, w = tile and gs = group
     * slot, so a block of nqt G waves transforms G groups of 16 lines at once.
     * Otherwise wave w takes tile w of one group at a time (waves past nqt
     * idle).  GP pays in FP64 (12% on the 83-point axis); in FP32 the plain
     * mapping is 4% faster (both measured on gfx950).                        */
    const int gpb = GP ? max(1, (int)(blockDim.x >> 6) / nqt) : 1;
    const int w = GP ? wave % nqt : wave, gs = GP ? wave / nqt : 0;
    const bool wact = GP ? gs < gpb : wave < nqt;
    /* A operands: lane holds row q = 16w + row, column k = 4 ks + kq (t = k + 1)
     * of the cos / sin matrices; r = q t mod n advances by 4q mod n per k step
     * and indexes the twiddle table exp(-2 pi i r / n) */
    const int q = w * 16 + row;
    const bool qon = q < Q;
    const int dr = (4 * q) % n;
    const int r0 = (q * (kq + 1)) % n;
    const bool even = (n & 1) == 0;
    const int ngr = (nl + 15) >> 4;
    for (int g0 = 0; g0 < ngr; g0 += gpb) {
        /* s_t and d_t in place for the groups g0..g0+gpb-1: s_t -> x_t,
         * d_t -> x_{n-t} (t = 1..m); x_0 and, for even n, h = x_{n/2} stay */
        if constexpr (GP) {
            /* thread -> (line slot jl, first pair k0); blockDim is a multiple of the
             * 16 gpb line slots, so each thread keeps its line and steps k */
            const int span = gpb * 16, jl = threadIdx.x % span, kstep = blockDim.x / span;
            const int line = g0 * 16 + jl;
            if (line < nl) {
                T2* Sl = S + line * P;
                for (int k = threadIdx.x / span; k < m; k += kstep) {
                    const int t = k + 1;
                    const T2 a = Sl[t], b = Sl[n - t];
                    Sl[t] = cadd(a, b);
                    Sl[n - t] = csub(a, b);
                }
            }
        } else {
            for (int idx = threadIdx.x; idx < m * 16; idx += blockDim.x) {
                const int k = idx >> 4, line = g0 * 16 + (idx & 15), t = k + 1;
                if (line < nl) {
                    T2* Sl = S + line * P;
                    const T2 a = Sl[t], b = Sl[n - t];
                    Sl[t] = cadd(a, b);
                    Sl[n - t] = csub(a, b);
                }
            }
        }
        __syncthreads();
        /* lane's column (line) row of its group; lines past nl read the last
         * line, whose outputs in those columns are discarded */
        const int gi = g0 + gs;
        const bool gact = wact && gi < ngr;
        const int line = gi * 16 + row;
        const T2* Sb = S + min(line, nl - 1) * P;
        acc_t cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0}, sr = {0, 0, 0, 0}, si = {0, 0, 0, 0};
        T2 x0, h;
        x0.x = 0; x0.y = 0; h.x = 0; h.y = 0;
        if (gact) {
            x0 = Sb[0];
            if (even) h = Sb[n >> 1];
            int r = r0;
            for (int ks = 0; ks < nks; ++ks) {
                /* B: s / d of k = 4 ks + kq (t = k + 1; t > m rows meet a zero A) */
                const int t = min(ks * 4 + kq + 1, max(m, 1));
                const T2 bs = Sb[t], bd = Sb[n - t];
                T a_c = 0, a_s = 0;
                if (qon && ks * 4 + kq + 1 <= m) {
                    const T2 wv = tb.W[r];
                    a_c = wv.x; a_s = wv.y;
                }
                r += dr;
                if (r >= n) r -= n;
                cr = Mfma<T>::mac(a_c, bs.x, cr);
                ci = Mfma<T>::mac(a_c, bs.y, ci);
                sr = Mfma<T>::mac(a_s, bd.x, sr);
                si = Mfma<T>::mac(a_s, bd.y, si);
            }
        }
        __syncthreads();   /* every wave has read its operands before S is overwritten */
        /* result: lane holds rows q = 16w + Mfma::row(kq, v), column (line) row */
        if (gact && line < nl) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int qq = w * 16 + Mfma<T>::row(kq, v);
                if (qq < Q) {
                    const T Ax = x0.x + ((qq & 1) ? -h.x : h.x) + cr[v];
                    const T Ay = x0.y + ((qq & 1) ? -h.y : h.y) + ci[v];
                    const T Bx = -sr[v], By = -si[v];
                    T2 lo, hi;   /* forward: X_q = A - iB, X_{n-q} = A + iB */
                    if (!inv) { lo.x = Ax + By; lo.y = Ay - Bx; hi.x = Ax - By; hi.y = Ay + Bx; }
                    else      { lo.x = Ax - By; lo.y = Ay + Bx; hi.x = Ax + By; hi.y = Ay - Bx; }
                    S[line * P + qq] = lo;
                    if (qq != 0 && 2 * qq != n) S[line * P + n - qq] = hi;
                }
            }
        }
        __syncthreads();
    }
}

struct NoMid { template <typename T2> __device__ void operator()(T2*) const {} };

/* run the whole stage list; `mid` handles R = 0 stages and must end with a
 * barrier.  ENG selects the stage kinds compiled in (a plan's lists are all
 * direct-DFT or all Stockham), so that each kernel variant is register-
 * allocated for its own engine only.                                        */
enum { ENG_STOCKHAM = 0, ENG_MFMA = 1 };
template <typename T2, int MAXE, int ENG, typename Mid, bool LF = false>
__device__ __forceinline__ double run_stages(T2* S, int nl, int P, const LinePlanDev& pl, const Tabs<T2>& tb,
                                             const Mid& mid, unsigned mNl = 0) {
    double facc = 0.0;   /* what a folded mid step accumulates */
    constexpr bool FOLD = FoldOf<Mid>::v && ENG != ENG_MFMA;
    bool fold_next = false;
    for (int s = 0; s < pl.nst; ++s) {
        const StageDev& st = pl.st[s];
        if (st.R == 0) {
            if (FOLD && s + 1 < pl.nst) { fold_next = true; continue; }
            mid(S);
            continue;
        }
        if constexpr (ENG == ENG_MFMA) {
            dft_mfma<T2, sizeof(S[0].x) == 8>(S, nl, P, st, tb);
        } else {
            const bool fm = fold_next;
            fold_next = false;
            /* the compound radices exist in the FP32 kernels only (the FP64
             * plans never use them, and their registers would cost waves) */
            constexpr bool CMP = MAXE >= 16;
            switch (st.R) {
                case 2: facc += fft_pass<T2, 2, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                case 3: facc += fft_pass<T2, 3, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                case 4: facc += fft_pass<T2, 4, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                case 5: facc += fft_pass<T2, 5, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                case 7: facc += fft_pass<T2, 7, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                case 8: facc += fft_pass<T2, 8, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                case 11: facc += fft_pass<T2, 11, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                case 13: facc += fft_pass<T2, 13, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                default:
                    if constexpr (CMP) {
                        switch (st.R) {
                            case 6: facc += fft_pass<T2, 6, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                            case 9: facc += fft_pass<T2, 9, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                            case 10: facc += fft_pass<T2, 10, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                            case 12: facc += fft_pass<T2, 12, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                            case 15: facc += fft_pass<T2, 15, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                            case 16: facc += fft_pass<T2, 16, MAXE, Mid, LF>(S, nl, P, st, tb, mid, fm, mNl); break;
                            default: break;
                        }
                    }
                    break;
            }
        }
    }
    return facc;
}

/* engine of a stage list (host and device) */
__host__ __device__ inline int plan_engine(const LinePlanDev& pl) {
    for (int s = 0; s < pl.nst; ++s) if (pl.st[s].R == DFT_MFMA) return ENG_MFMA;
    return ENG_STOCKHAM;
}

/* register budget (complex elements per thread and pass) of each precision */
template <typename T> struct MaxE;
template <> struct MaxE<double> { static constexpr int v = 8; };
template <> struct MaxE<float> { static constexpr int v = 16; };

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
    int pi_guard;             /* General with three equal mesh lengths: Lambda can be parallel to (1,1,1) */
    double ha, hb, hc;
    double iha, ihb, ihc;     /* 1/h */
    unsigned m_na, m_nb;      /* udiv magics for / na, / nb */
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

/* lambdas_line with exp(i pi (a_f + phi) / na) formed as the product of the
 * tabulated ea = exp(i pi a_f / na) and the line's ephi = exp(i pi phi / na):
 * one complex product instead of a division and a sincospi per mode.  The
 * product's absolute error is a few ulp, so sin stays accurate to a few ulp
 * relative to the largest symbol.                                           */
__device__ __forceinline__ void lambdas_line_tab(const GridDev& g, const LineSym& s, double2 ea, double2 ephi,
                                                 double2& lx, double2& ly, double2& lz) {
    const double2 e = cmul(ea, ephi);
    const double2 La = make_double2(-2.0 * e.y * e.y * g.iha, 2.0 * e.y * e.x * g.iha);
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
/* 1/sqrt(x): FP64, or (FAST) the FP32 hardware estimate (about 1 ulp of FP32),
 * which is all a basis that is rounded to FP32 afterwards needs.  A relative
 * error d in the norms scales each reduced mode of the FP32 operator by about
 * (1 + 2d), a perturbation that the FP64 refinement removes at rate ~2d.   */
template <bool FAST>
__device__ __forceinline__ double rsq(double x) {
    if (FAST) return (double)rsqrtf((float)x);
    return rsqrt(x);
}

/* General gauge (FAME_Matrix_Lambdas.cu:441-461) with S = Lx + Ly + Lz:
 *   p1 = (Lq - L_i conj(S)) / sqrt(Lq (3 Lq - |S|^2)),  p2 = conj(Lz - Ly, Lx - Lz, Ly - Lx) / sqrt(3 Lq - |S|^2).
 * Its normalisers vanish (0/0) where L is parallel to (1, 1, 1), and lose
 * accuracy near it.  Every symbol has the form (e^{i phi} - 1)/h_axis, so L can
 * be parallel to (1, 1, 1) (at a nonzero mode) only when the three mesh lengths
 * are equal.  Such grids (pi_guard) take the GUARD form: with d = (Lz - Ly,
 * Lx - Lz, Ly - Lx), 3 Lq - |S|^2 = |d|^2 and the p1 numerator is conj(L) x d,
 * both formed from d without cancellation, and a mode with |d|^2 <= 1e-30 Lq
 * takes the Simple gauge (w = (3, 5, 1)), which is well defined there.  Every
 * kernel evaluates the same form and branch from the same FP64 symbols, so the
 * basis is consistent; other grids keep the reference's arithmetic, whose
 * register footprint matters in the FP32 inner kernels.                     */
template <bool FAST = false, bool GUARD = false>
__device__ __forceinline__ void pi_basis(const GridDev& g, double2 lx, double2 ly, double2 lz,
                                         double2 p1[3], double2 p2[3]) {
    const double Lq = lx.x * lx.x + lx.y * lx.y + ly.x * ly.x + ly.y * ly.y + lz.x * lz.x + lz.y * lz.y;
    if (GUARD) {
        const double2 d1 = csub(lz, ly), d2 = csub(lx, lz), d3 = csub(ly, lx);
        const double nd = d1.x * d1.x + d1.y * d1.y + d2.x * d2.x + d2.y * d2.y + d3.x * d3.x + d3.y * d3.y;
        if (nd > 1e-30 * Lq) {
            const double i2 = rsq<FAST>(nd);
            const double i1 = i2 * rsq<FAST>(Lq);
            p1[0] = cscale(csub(cmul(cconj(ly), d3), cmul(cconj(lz), d2)), i1);
            p1[1] = cscale(csub(cmul(cconj(lz), d1), cmul(cconj(lx), d3)), i1);
            p1[2] = cscale(csub(cmul(cconj(lx), d2), cmul(cconj(ly), d1)), i1);
            p2[0] = cscale(cconj(d1), i2);
            p2[1] = cscale(cconj(d2), i2);
            p2[2] = cscale(cconj(d3), i2);
            return;
        }
    } else if (g.general) {
        const double2 Ls = cadd(cadd(lx, ly), lz);
        const double Lp = Ls.x * Ls.x + Ls.y * Ls.y;
        const double i1 = rsq<FAST>(3.0 * Lq * Lq - Lq * Lp);
        const double i2 = rsq<FAST>(3.0 * Lq - Lp);
        const double2 lxs = cmulc(lx, Ls), lys = cmulc(ly, Ls), lzs = cmulc(lz, Ls);
        p1[0] = make_double2((Lq - lxs.x) * i1, -lxs.y * i1);
        p1[1] = make_double2((Lq - lys.x) * i1, -lys.y * i1);
        p1[2] = make_double2((Lq - lzs.x) * i1, -lzs.y * i1);
        p2[0] = cscale(csub(cconj(lz), cconj(ly)), i2);
        p2[1] = cscale(csub(cconj(lx), cconj(lz)), i2);
        p2[2] = cscale(csub(cconj(ly), cconj(lx)), i2);
        return;
    }
    const double al = 3.0, be = 5.0;
    double2 t1 = csub(cscale(lz, be), ly);
    double2 t2 = csub(lx, cscale(lz, al));
    double2 t3 = csub(cscale(ly, al), cscale(lx, be));
    const double n2 = t1.x * t1.x + t1.y * t1.y + t2.x * t2.x + t2.y * t2.y + t3.x * t3.x + t3.y * t3.y;
    const double i2 = rsq<FAST>(n2);
    const double i1 = i2 * rsq<FAST>(Lq);
    p1[0] = cscale(csub(cmulc(t3, ly), cmulc(t2, lz)), i1);
    p1[1] = cscale(csub(cmulc(t1, lz), cmulc(t3, lx)), i1);
    p1[2] = cscale(csub(cmulc(t2, lx), cmulc(t1, ly)), i1);
    p2[0] = cscale(cconj(t1), i2);
    p2[1] = cscale(cconj(t2), i2);
    p2[2] = cscale(cconj(t3), i2);
}

/* The General gauge applied without forming Pi (pi_basis's non-GUARD General
 * branch, same normalisers i1, i2, same rounding of them).  With S = Lx + Ly + Lz,
 * Pi1_c = (Lq - L_c conj(S)) i1 and Pi2_c = conj(d_c) i2, d = (Lz - Ly, Lx - Lz, Ly - Lx):
 *   expansion   u_c  = Lq i1 x1 - L_c (i1 conj(S) x1) + conj(d_c) (i2 x2),
 *   contraction Pi1^H w = i1 (Lq sum_c w_c - S sum_c conj(L_c) w_c),
 *               Pi2^H w = i2 (Lx (w_y - w_z) + Ly (w_z - w_x) + Lz (w_x - w_y)).
 * Mathematically the same products as with pi_basis's vectors, with fewer
 * operations and live values (the contraction is issue-bound).             */
template <bool FAST>
__device__ __forceinline__ void pi_norms_general(double2 lx, double2 ly, double2 lz, double& Lq, double2& Ls,
                                                 double& i1, double& i2) {
    Lq = lx.x * lx.x + lx.y * lx.y + ly.x * ly.x + ly.y * ly.y + lz.x * lz.x + lz.y * lz.y;
    Ls = cadd(cadd(lx, ly), lz);
    const double Lp = Ls.x * Ls.x + Ls.y * Ls.y;
    i1 = rsq<FAST>(3.0 * Lq * Lq - Lq * Lp);
    i2 = rsq<FAST>(3.0 * Lq - Lp);
}
template <bool FAST>
__device__ __forceinline__ void pi_expand_general(double2 lx, double2 ly, double2 lz, double2 x1, double2 x2,
                                                  double2 u[3]) {
    double Lq, i1, i2;
    double2 Ls;
    pi_norms_general<FAST>(lx, ly, lz, Lq, Ls, i1, i2);
    const double2 al = cscale(x1, Lq * i1), be = cscale(cmulc(x1, Ls), i1), ga = cscale(x2, i2);
    u[0] = cadd(csub(al, cmul(lx, be)), cmulc(ga, csub(lz, ly)));
    u[1] = cadd(csub(al, cmul(ly, be)), cmulc(ga, csub(lx, lz)));
    u[2] = cadd(csub(al, cmul(lz, be)), cmulc(ga, csub(ly, lx)));
}
template <bool FAST>
__device__ __forceinline__ void pi_contract_general(double2 lx, double2 ly, double2 lz, double2 w0, double2 w1,
                                                    double2 w2, double2& y1, double2& y2) {
    double Lq, i1, i2;
    double2 Ls;
    pi_norms_general<FAST>(lx, ly, lz, Lq, Ls, i1, i2);
    const double2 sw = cadd(cadd(w0, w1), w2);
    const double2 t = cadd(cadd(cmulc(w0, lx), cmulc(w1, ly)), cmulc(w2, lz));
    y1 = cscale(csub(cscale(sw, Lq), cmul(Ls, t)), i1);
    y2 = cscale(cadd(cadd(cmul(lx, csub(w1, w2)), cmul(ly, csub(w2, w0))), cmul(lz, csub(w0, w1))), i2);
}

}  // namespace fame
