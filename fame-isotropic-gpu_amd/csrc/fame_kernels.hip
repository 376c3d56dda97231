/*
 * fame_kernels.hip -- gfx950 kernels of the null-space-free Maxwell band solver.
 *
 * One QBQ = Q_r^H B^{-1} Q_r apply (FAME_Matrix_Vector_Production_Isotropic_QBQ.cu)
 * is five axis passes over a batch of problems:
 *   P1 inner  : Pi expansion of the reduced vector + inverse FFT along a + inner twiddle
 *   P2 middle : inverse FFT along b + middle twiddle
 *   P3 outer  : inverse FFT along c, 1/eps and 1/N, forward FFT along c (the outer
 *               twiddle D and its conjugate cancel around the real diagonal 1/eps)
 *   P4 middle : conjugate middle twiddle + forward FFT along b
 *   P5 inner  : conjugate inner twiddle + forward FFT along a + Pi^H contraction
 * Every pass reads and writes each field element once; twiddles and Pi bases are
 * recomputed from closed forms instead of being read from N-sized tables.
 * All kernels take the problem index from blockIdx.y and skip problems whose
 * `active` flag is 0 (per-problem convergence masks of the batched solvers).
 */
#include "fame_device.hpp"
#include "fame_kernels.h"

namespace fame {

template <typename T> struct Vec;
template <> struct Vec<double> { typedef double2 T2; };
template <> struct Vec<float>  { typedef float2 T2; };

/* global loads in flight per thread: gathers of the axis kernels, the 1/eps
 * step of the outer kernel, and the inner kernels (several vectors per
 * element).  Larger batches cost registers, and so waves per SIMD; these are
 * the measured best on gfx950 for the 96x83x78 grid at 30 problems.        */
template <typename T> struct Inflight { static constexpr int v = sizeof(T) == 4 ? 8 : 4; };
template <typename T> struct InflightEps { static constexpr int v = sizeof(T) == 4 ? 4 : 2; };
template <typename T> struct InflightInner { static constexpr int v = 2; };

/* vector-access preconditions (a null pointer is never dereferenced) */
__device__ __forceinline__ bool aligned16(const void* q) { return ((uintptr_t)q & 15) == 0; }
__device__ __forceinline__ bool aligned8(const void* q) { return ((uintptr_t)q & 7) == 0; }

template <typename T2> __device__ __forceinline__ T2 from_d(double2 v);
template <> __device__ __forceinline__ double2 from_d<double2>(double2 v) { return v; }
template <> __device__ __forceinline__ float2 from_d<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ double2 to_d(double2 v) { return v; }
__device__ __forceinline__ double2 to_d(float2 v) { return make_double2(v.x, v.y); }

/* exp(i pi x) in the kernel precision.  FP32: the argument is reduced to
 * [-1, 1] in FP64 first, so that the FP32 evaluation loses nothing for large x. */
template <typename T2> __device__ __forceinline__ T2 tw_expipi(double x);
template <> __device__ __forceinline__ double2 tw_expipi<double2>(double x) { return expipi(x); }
template <> __device__ __forceinline__ float2 tw_expipi<float2>(double x) {
    const float r = (float)(x - 2.0 * rint(0.5 * x));
    float s, c;
    sincospif(r, &s, &c);
    return make_float2(c, s);
}

/* ------------------------------------------------------------------------- */
/*  P1 / P5: inner axis                                                       */
/* ------------------------------------------------------------------------- */
/* EXPAND = 1 : F = twid .* IFFT_a( Pi x )          (Qr / Pr first stage)
 * EXPAND = 0 : y = Pi^H FFT_a( conj(twid) .* F )   (Qrs / Prs last stage)
 * pikind = 0 : Pi_Qr = [Pi1 Pi2]; 1 : Pi_Pr = [-conj(Pi2) conj(Pi1)]
 *              (FAME_Matrix_Lambdas.cu:217-243)
 * A block owns L lines (b, c) and their three components: 3L LDS lines.      */
/* Pi forms compiled into an inner kernel: PI_ANY forms the basis vectors with
 * pi_basis (Simple gauge, or the Pr / Prs variants), PI_GUARD the same with the
 * equal-mesh guard, PI_GEN the General gauge of Qr / Qrs applied without forming
 * the vectors (pi_expand_general / pi_contract_general).  Each is its own
 * instance, so that registers are allocated for the one form it runs.      */
enum { PI_ANY = 0, PI_GUARD = 1, PI_GEN = 2 };

/* Stockham lane order of the outer pass (fft_pass): across its L lines in
 * FP64, along them in FP32.  On the FCC 96 grid (L = 16, M = 96) it took the
 * FP64 outer pass from 564 to 489 us; the FP32 one (L = 32) kept its time at
 * 5 waves per SIMD and spilled at its 6.  The inner passes, with 3L = 18 FP64
 * lines per block, were 1-3 us slower with it (gfx950).                    */
template <typename T> struct OuterLF { static constexpr bool v = sizeof(T) == 8; };

template <typename T, int EXPAND, int ENG, int PIK>
__global__ void __launch_bounds__(FFT_THREADS) k_inner(KInner args) {
    constexpr bool GUARD = PIK == PI_GUARD;
    typedef typename Vec<T>::T2 T2;
    const int p = blockIdx.y;
    if (args.active && !args.active[p]) return;
    const int nlines = args.g.nb * args.g.nc;
    const int nblk = (nlines + args.L - 1) / args.L;
    const int tile = xcd_tile();
    if (tile >= nblk) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ LineSym lsym[MAX_LINES_INNER];
    __shared__ double2 ephi[MAX_LINES_INNER];
    const GridDev& g = args.g;
    const ProbDev pr = args.prob[p];
    const LinePlanDev& pl = args.plan;
    const int na = g.na, L = args.L, P = args.P;
    T2* base = reinterpret_cast<T2*>(smem_raw);
    const Tabs<T2> tb = load_tabs<T2>(pl, base);
    T2* S = base + lds_head(pl);
    /* exp(i pi a / na) for a < na behind the lines, 16-byte aligned */
    double2* ea = reinterpret_cast<double2*>(base + inner_tab_offset(lds_head(pl) + 3 * L * P, (int)sizeof(T)));
    const int l0 = tile * L;
    if ((int)threadIdx.x < L) {
        const int gl = l0 + threadIdx.x;
        if (gl < nlines) {
            const int cf = udiv(gl, g.m_nb), bf = gl - cf * g.nb;
            const LineSym ls = line_sym(g, pr, bf, cf);
            lsym[threadIdx.x] = ls;
            ephi[threadIdx.x] = expipi(ls.phi / (double)na);
        }
    }
    for (int a = threadIdx.x; a < na; a += FFT_THREADS) ea[a] = expipi((double)a / (double)na);
    __syncthreads();
    /* inner twiddle exp(2 i pi phi a / na) = th[line][a >> 4] * tl[line][a & 15]:
     * two table reads and one complex product per element instead of a sincos */
    const int nh = (na + 15) >> 4, tpl = 16 + nh;
    T2* twt = reinterpret_cast<T2*>(ea + na);
    for (int t = threadIdx.x; t < L * tpl; t += FFT_THREADS) {
        const int line = t / tpl, q = t - line * tpl;
        if (l0 + line < nlines) {
            const double ph = 2.0 * lsym[line].phi / (double)na;
            twt[t] = tw_expipi<T2>(ph * (double)(q < 16 ? q : 16 * (q - 16)));
        }
    }
    __syncthreads();
    /* per-problem base pointers (uniform) with 32-bit element offsets below
     * (3N < 2^31 is checked at context creation) */
    const unsigned N = (unsigned)g.N;
    const T2* X = reinterpret_cast<const T2*>(args.x) + (size_t)p * args.xstride;
    T2* F = reinterpret_cast<T2*>(args.f) + (size_t)p * args.fstride;
    const double inv_na = 1.0 / (double)na;
    /* width written per LDS line: na, or the zero-padded Bluestein length */
    const int Wd = pl.blue ? pl.M : na;
    const unsigned mW = pl.blue ? pl.mM : g.m_na;

    if (EXPAND) {
        /* CG fusion: per-problem scalars of the direction / solution update */
        const bool cgf = args.cg_r != nullptr;
        T beta = 0, ar = 0, ai = 0;
        bool pend = false;
        if (cgf) {
            const CgScal& cs = args.cgsc[p];
            beta = (T)cs.beta; ar = (T)cs.alpha_re; ai = (T)cs.alpha_im; pend = cs.pend != 0;
        }
        T2* const Pw = const_cast<T2*>(X);
        const T2* const R = cgf ? reinterpret_cast<const T2*>(args.cg_r) + (size_t)p * args.xstride : nullptr;
        T2* const Xs = cgf ? reinterpret_cast<T2*>(args.cg_x) + (size_t)p * args.xstride : nullptr;
        /* r -= alpha Ap of the previous iteration rides on this pass too */
        const bool rup = cgf && pend && args.cg_ap != nullptr;
        const T2* const AP = rup ? reinterpret_cast<const T2*>(args.cg_ap) + (size_t)p * args.xstride : nullptr;
        /* refinement: x += fx32 + alpha fp32 of the FP32 solve that just ended */
        const bool FX = args.fx32 != nullptr;
        const float2* const FX32 = FX ? reinterpret_cast<const float2*>(args.fx32) + (size_t)p * args.xstride : nullptr;
        const float2* const FP32 = FX ? reinterpret_cast<const float2*>(args.fp32) + (size_t)p * args.xstride : nullptr;
        float far = 0.0f, fai = 0.0f;
        bool fpend = false;
        if (FX) {
            const CgScal& cs = args.cgsc[p];
            far = (float)cs.alpha_re; fai = (float)cs.alpha_im; fpend = cs.pend != 0;
        }
        struct XIn { T2 x1, x2, r1, r2, s1, s2, a1, a2; };
        double rnorm = 0.0;   /* <r,r> of the updated residual, this thread's share */
        /* one mode (line, a) at offset f: the CG / refinement updates of the
         * reduced vectors, then the Pi expansion into the three LDS lines */
        auto expand_mode = [&](int line, int a, unsigned f, const XIn& d) {
            T2 xv1 = d.x1, xv2 = d.x2;
            if (FX) {
                float2 u1 = make_float2((float)d.s1.x, (float)d.s1.y), u2 = make_float2((float)d.s2.x, (float)d.s2.y);
                if (fpend) {
                    const float2 v1 = make_float2((float)d.a1.x, (float)d.a1.y);
                    const float2 v2 = make_float2((float)d.a2.x, (float)d.a2.y);
                    u1.x += far * v1.x - fai * v1.y; u1.y += far * v1.y + fai * v1.x;
                    u2.x += far * v2.x - fai * v2.y; u2.y += far * v2.y + fai * v2.x;
                }
                xv1.x += (T)u1.x; xv1.y += (T)u1.y; xv2.x += (T)u2.x; xv2.y += (T)u2.y;
                Pw[f] = xv1; Pw[N + f] = xv2;
            }
            if (cgf) {
                if (pend) {
                    T2 s1 = d.s1, s2 = d.s2;
                    s1.x += ar * xv1.x - ai * xv1.y; s1.y += ar * xv1.y + ai * xv1.x;
                    s2.x += ar * xv2.x - ai * xv2.y; s2.y += ar * xv2.y + ai * xv2.x;
                    Xs[f] = s1; Xs[N + f] = s2;
                }
                T2 r1 = d.r1, r2 = d.r2;
                if (rup) {
                    r1.x -= ar * d.a1.x - ai * d.a1.y; r1.y -= ar * d.a1.y + ai * d.a1.x;
                    r2.x -= ar * d.a2.x - ai * d.a2.y; r2.y -= ar * d.a2.y + ai * d.a2.x;
                    T2* Rw = const_cast<T2*>(R);
                    Rw[f] = r1; Rw[N + f] = r2;
                }
                rnorm += (double)r1.x * r1.x + (double)r1.y * r1.y + (double)r2.x * r2.x + (double)r2.y * r2.y;
                xv1.x = r1.x + beta * xv1.x; xv1.y = r1.y + beta * xv1.y;
                xv2.x = r2.x + beta * xv2.x; xv2.y = r2.y + beta * xv2.y;
                Pw[f] = xv1; Pw[N + f] = xv2;
            }
            double2 u[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
            if (!(pr.gamma && f == 0)) {
                double2 lx, ly, lz;
                lambdas_line_tab(g, lsym[line], ea[a], ephi[line], lx, ly, lz);
                const double2 x1 = to_d(xv1), x2 = to_d(xv2);
                if constexpr (PIK == PI_GEN) {
                    pi_expand_general<sizeof(T) == 4>(lx, ly, lz, x1, x2, u);
                } else {
                    double2 p1[3], p2[3];
                    pi_basis<sizeof(T) == 4, GUARD>(g, lx, ly, lz, p1, p2);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        if (args.pikind == 0) u[c] = cadd(cmul(p1[c], x1), cmul(p2[c], x2));
                        else u[c] = cadd(cmul(make_double2(-p2[c].x, p2[c].y), x1), cmul(cconj(p1[c]), x2));
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) S[(c * L + line) * P + a] = from_d<T2>(u[c]);
        };
        /* FP32 without Bluestein and with na even: two consecutive modes of a
         * line per thread, each vector moved as one 16-byte access (a line
         * starts at an even offset, and every vector at a 16-byte boundary) */
        bool pairs = false;
        if constexpr (sizeof(T) == 4)
            pairs = !pl.blue && !(na & 1) && !FX && aligned16(X) && aligned16(R) && aligned16(Xs) && aligned16(AP) &&
                    aligned16(F);
        if constexpr (sizeof(T) == 4) {
            if (pairs) {
                struct XIn2 { float4 x1, x2, r1, r2, s1, s2, a1, a2; };
                auto lo = [](const float4& v) { return make_float2(v.x, v.y); };
                auto hi = [](const float4& v) { return make_float2(v.z, v.w); };
                batched_loop<1>(
                    (L * na) >> 1,
                    [&](int idx) -> XIn2 {
                        const int line = udiv(2 * idx, g.m_na), a = 2 * idx - line * na;
                        XIn2 d;
                        if (l0 + line < nlines) {
                            const unsigned f = (unsigned)((l0 + line) * na + a);
                            d.x1 = *reinterpret_cast<const float4*>(X + f);
                            d.x2 = *reinterpret_cast<const float4*>(X + N + f);
                            if (cgf) {
                                d.r1 = *reinterpret_cast<const float4*>(R + f);
                                d.r2 = *reinterpret_cast<const float4*>(R + N + f);
                                if (pend) {
                                    d.s1 = *reinterpret_cast<const float4*>(Xs + f);
                                    d.s2 = *reinterpret_cast<const float4*>(Xs + N + f);
                                }
                                if (rup) {
                                    d.a1 = *reinterpret_cast<const float4*>(AP + f);
                                    d.a2 = *reinterpret_cast<const float4*>(AP + N + f);
                                }
                            }
                        }
                        return d;
                    },
                    [&](int idx, const XIn2& d) {
                        const int line = udiv(2 * idx, g.m_na), a = 2 * idx - line * na;
                        if (l0 + line < nlines) {
                            const unsigned f = (unsigned)((l0 + line) * na + a);
                            XIn e0, e1;
                            e0.x1 = lo(d.x1); e0.x2 = lo(d.x2); e1.x1 = hi(d.x1); e1.x2 = hi(d.x2);
                            e0.r1 = lo(d.r1); e0.r2 = lo(d.r2); e1.r1 = hi(d.r1); e1.r2 = hi(d.r2);
                            e0.s1 = lo(d.s1); e0.s2 = lo(d.s2); e1.s1 = hi(d.s1); e1.s2 = hi(d.s2);
                            e0.a1 = lo(d.a1); e0.a2 = lo(d.a2); e1.a1 = hi(d.a1); e1.a2 = hi(d.a2);
                            expand_mode(line, a, f, e0);
                            expand_mode(line, a + 1, f + 1, e1);
                        } else {
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                S[(c * L + line) * P + a] = make_float2(0.f, 0.f);
                                S[(c * L + line) * P + a + 1] = make_float2(0.f, 0.f);
                            }
                        }
                    });
            }
        }
        if (!pairs)
            batched_loop<InflightInner<T>::v>(
                L * Wd,
                [&](int idx) -> XIn {
                    const int line = udiv(idx, mW), a = idx - line * Wd;
                    const int gl = l0 + line;
                    XIn d;
                    if (gl < nlines && a < na) {
                        const unsigned f = (unsigned)(gl * na + a);
                        d.x1 = X[f]; d.x2 = X[N + f];
                        if (FX) {
                            const float2 u1 = FX32[f], u2 = FX32[N + f];
                            d.s1.x = u1.x; d.s1.y = u1.y; d.s2.x = u2.x; d.s2.y = u2.y;
                            if (fpend) {
                                const float2 v1 = FP32[f], v2 = FP32[N + f];
                                d.a1.x = v1.x; d.a1.y = v1.y; d.a2.x = v2.x; d.a2.y = v2.y;
                            }
                        }
                        if (cgf) {
                            d.r1 = R[f]; d.r2 = R[N + f];
                            if (pend) { d.s1 = Xs[f]; d.s2 = Xs[N + f]; }
                            if (rup) { d.a1 = AP[f]; d.a2 = AP[N + f]; }
                        }
                    }
                    return d;
                },
                [&](int idx, const XIn& d) {
                    const int line = udiv(idx, mW), a = idx - line * Wd;
                    const int gl = l0 + line;
                    if (gl < nlines && a < na) {
                        expand_mode(line, a, (unsigned)(gl * na + a), d);
                    } else {
#pragma unroll
                        for (int c = 0; c < 3; ++c) S[(c * L + line) * P + a] = from_d<T2>(make_double2(0, 0));
                    }
                });
        /* inside CG with the stored Ap: block sum of <r,r> into slot 5 of
         * partials[p][tile] (the contracting pass fills slots 0..2); with the
         * residual updated by the contracting pass, that pass sums <r,r> */
        __shared__ double rred[FFT_THREADS / 64];
        const bool rsum = cgf && args.cg_ap != nullptr;
        if (rsum) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) rnorm += __shfl_xor(rnorm, o);
            if ((threadIdx.x & 63) == 0) rred[threadIdx.x >> 6] = rnorm;
        }
        __syncthreads();
        if (rsum && threadIdx.x == 0) {
            double sum = 0;
            for (int w = 0; w < FFT_THREADS / 64; ++w) sum += rred[w];
            args.partials[((size_t)p * nblk + tile) * 6 + 5] = sum;
        }
        run_stages<T2, MaxE<T>::v, ENG>(S, 3 * L, P, pl, tb, NoMid());
        if constexpr (sizeof(T) == 4) {
            if (pairs)
                for (int idx = threadIdx.x; idx < ((L * na) >> 1); idx += FFT_THREADS) {
                    const int line = udiv(2 * idx, g.m_na), a = 2 * idx - line * na;
                    const int gl = l0 + line;
                    if (gl >= nlines) continue;
                    /* a even: a and a + 1 share the high twiddle factor */
                    const float2 th = twt[line * tpl + 16 + (a >> 4)];
                    const float2 w0 = cmul(th, twt[line * tpl + (a & 15)]), w1 = cmul(th, twt[line * tpl + (a & 15) + 1]);
                    const unsigned f = (unsigned)(gl * na + a);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const float2 v0 = cmul(post_mul<T2, T>(pl, tb, a, S[(c * L + line) * P + a]), w0);
                        const float2 v1 = cmul(post_mul<T2, T>(pl, tb, a + 1, S[(c * L + line) * P + a + 1]), w1);
                        *reinterpret_cast<float4*>(F + c * N + f) = make_float4(v0.x, v0.y, v1.x, v1.y);
                    }
                }
        }
        if (!pairs)
            for (int idx = threadIdx.x; idx < L * na; idx += FFT_THREADS) {
                const int line = udiv(idx, g.m_na), a = idx - line * na;
                const int gl = l0 + line;
                if (gl >= nlines) continue;
                const T2 w = cmul(twt[line * tpl + 16 + (a >> 4)], twt[line * tpl + (a & 15)]);
                const unsigned f = (unsigned)(gl * na + a);
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    F[c * N + f] = cmul(post_mul<T2, T>(pl, tb, a, S[(c * L + line) * P + a]), w);
            }
    } else {
        /* FP32 without Bluestein, na even and 16-byte aligned vectors: two
         * consecutive modes of a line per thread, as in the expanding pass */
        T2* Y = reinterpret_cast<T2*>(args.y) + (size_t)p * args.xstride;
        const T2* D = args.dotvec ? reinterpret_cast<const T2*>(args.dotvec) + (size_t)p * args.xstride : nullptr;
        const bool RU = args.cg_r != nullptr;
        T2* const RW = RU ? reinterpret_cast<T2*>(args.cg_r) + (size_t)p * args.xstride : nullptr;
        const bool RB = args.rb != nullptr;
        bool pairs = false;
        if constexpr (sizeof(T) == 4)
            pairs = !pl.blue && !(na & 1) && !RB && aligned16(F) && aligned16(Y) && aligned16(D) && aligned16(RW);
        auto lo = [](const float4& v) { return make_float2(v.x, v.y); };
        auto hi = [](const float4& v) { return make_float2(v.z, v.w); };
        struct FIn { T2 v[3]; };
        if constexpr (sizeof(T) == 4) {
            if (pairs) {
                struct FIn2 { float4 v[3]; };
                batched_loop<1>(
                    (L * na) >> 1,
                    [&](int idx) -> FIn2 {
                        const int line = udiv(2 * idx, g.m_na), a = 2 * idx - line * na;
                        FIn2 d;
#pragma unroll
                        for (int c = 0; c < 3; ++c) d.v[c] = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (l0 + line < nlines) {
                            const unsigned f = (unsigned)((l0 + line) * na + a);
#pragma unroll
                            for (int c = 0; c < 3; ++c) d.v[c] = *reinterpret_cast<const float4*>(F + c * N + f);
                        }
                        return d;
                    },
                    [&](int idx, const FIn2& d) {
                        const int line = udiv(2 * idx, g.m_na), a = 2 * idx - line * na;
                        float2 w0 = make_float2(1.f, 0.f), w1 = w0;
                        if (l0 + line < nlines) {
                            const float2 th = twt[line * tpl + 16 + (a >> 4)];
                            w0 = cmul(th, twt[line * tpl + (a & 15)]);
                            w1 = cmul(th, twt[line * tpl + (a & 15) + 1]);
                        }
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            S[(c * L + line) * P + a] = cmulc(lo(d.v[c]), w0);
                            S[(c * L + line) * P + a + 1] = cmulc(hi(d.v[c]), w1);
                        }
                    });
            }
        }
        if (!pairs)
            batched_loop<InflightInner<T>::v>(
                L * Wd,
                [&](int idx) -> FIn {
                    const int line = udiv(idx, mW), a = idx - line * Wd;
                    const int gl = l0 + line;
                    FIn d;
#pragma unroll
                    for (int c = 0; c < 3; ++c) { d.v[c].x = 0; d.v[c].y = 0; }
                    if (gl < nlines && a < na) {
                        const unsigned f = (unsigned)(gl * na + a);
#pragma unroll
                        for (int c = 0; c < 3; ++c) d.v[c] = F[c * N + f];
                    }
                    return d;
                },
                [&](int idx, FIn d) {
                    const int line = udiv(idx, mW), a = idx - line * Wd;
                    const int gl = l0 + line;
                    if (gl < nlines && a < na) {
                        const T2 w = cmul(twt[line * tpl + 16 + (a >> 4)], twt[line * tpl + (a & 15)]);
#pragma unroll
                        for (int c = 0; c < 3; ++c) d.v[c] = cmulc(d.v[c], w);
                    }
#pragma unroll
                    for (int c = 0; c < 3; ++c) S[(c * L + line) * P + a] = d.v[c];
                });
        __syncthreads();
        run_stages<T2, MaxE<T>::v, ENG>(S, 3 * L, P, pl, tb, NoMid());
        double dre = 0.0, dim = 0.0, yy = 0.0, rn = 0.0;
        const bool CGD = args.cgdots != 0;
        /* inside CG with cg_r set: r -= alpha y (alpha of the problem's CgScal)
         * is stored in place of y, with <r,r> of the updated residual */
        T ar = 0, ai = 0;
        if (RU) { ar = (T)args.cgsc[p].alpha_re; ai = (T)args.cgsc[p].alpha_im; }
        /* refinement residual: rb - y becomes the FP32 start r32 = p32, x32 = 0 */
        const T2* const RBv = RB ? reinterpret_cast<const T2*>(args.rb) + (size_t)p * args.xstride : nullptr;
        float2* const R32 = RB ? reinterpret_cast<float2*>(args.r32) + (size_t)p * args.xstride : nullptr;
        float2* const P32 = RB ? reinterpret_cast<float2*>(args.p32) + (size_t)p * args.xstride : nullptr;
        float2* const X32 = RB ? reinterpret_cast<float2*>(args.x32) + (size_t)p * args.xstride : nullptr;
        /* Pi^H contraction of mode (line, a) at offset f from the three LDS lines */
        auto contract = [&](int line, int a, unsigned f, T2& o1, T2& o2) {
            double2 y1 = make_double2(0, 0), y2 = make_double2(0, 0);
            if (!(pr.gamma && f == 0)) {
                double2 lx, ly, lz;
                lambdas_line_tab(g, lsym[line], ea[a], ephi[line], lx, ly, lz);
                if constexpr (PIK == PI_GEN) {
                    const double2 w0 = to_d(post_mul<T2, T>(pl, tb, a, S[(0 * L + line) * P + a]));
                    const double2 w1 = to_d(post_mul<T2, T>(pl, tb, a, S[(1 * L + line) * P + a]));
                    const double2 w2 = to_d(post_mul<T2, T>(pl, tb, a, S[(2 * L + line) * P + a]));
                    pi_contract_general<sizeof(T) == 4>(lx, ly, lz, w0, w1, w2, y1, y2);
                } else {
                    double2 p1[3], p2[3];
                    pi_basis<sizeof(T) == 4, GUARD>(g, lx, ly, lz, p1, p2);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double2 w = to_d(post_mul<T2, T>(pl, tb, a, S[(c * L + line) * P + a]));
                        if (args.pikind == 0) {
                            y1 = cadd(y1, cmul(cconj(p1[c]), w));
                            y2 = cadd(y2, cmul(cconj(p2[c]), w));
                        } else {
                            y1 = csub(y1, cmul(p2[c], w));
                            y2 = cadd(y2, cmul(p1[c], w));
                        }
                    }
                }
            }
            o1 = from_d<T2>(y1); o2 = from_d<T2>(y2);
        };
        /* dot partials of one mode: conj(d) . y and y . y */
        auto dots = [&](const T2& o1, const T2& o2, const T2& e1, const T2& e2) {
            if (D || CGD) {
                const double2 q1 = to_d(o1), q2 = to_d(o2);
                if (D) {
                    const double2 d1 = to_d(e1), d2 = to_d(e2);
                    dre += d1.x * q1.x + d1.y * q1.y + d2.x * q2.x + d2.y * q2.y;
                    dim += d1.x * q1.y - d1.y * q1.x + d2.x * q2.y - d2.y * q2.x;
                }
                if (CGD) yy += q1.x * q1.x + q1.y * q1.y + q2.x * q2.x + q2.y * q2.y;
            }
        };
        /* r -= alpha y of one mode, with its <r,r> */
        auto rstep = [&](T2& r1, T2& r2, const T2& o1, const T2& o2) {
            r1.x -= ar * o1.x - ai * o1.y; r1.y -= ar * o1.y + ai * o1.x;
            r2.x -= ar * o2.x - ai * o2.y; r2.y -= ar * o2.y + ai * o2.x;
            rn += (double)r1.x * r1.x + (double)r1.y * r1.y + (double)r2.x * r2.x + (double)r2.y * r2.y;
        };
        if constexpr (sizeof(T) == 4) {
            if (pairs) {
                struct DIn2 { float4 d1, d2, r1, r2; };
                batched_loop<1>(
                    (L * na) >> 1,
                    [&](int idx) -> DIn2 {
                        const int line = udiv(2 * idx, g.m_na), a = 2 * idx - line * na;
                        DIn2 d;
                        if (l0 + line < nlines) {
                            const unsigned f = (unsigned)((l0 + line) * na + a);
                            if (D) {
                                d.d1 = *reinterpret_cast<const float4*>(D + f);
                                d.d2 = *reinterpret_cast<const float4*>(D + N + f);
                            }
                            if (RU) {
                                d.r1 = *reinterpret_cast<const float4*>(RW + f);
                                d.r2 = *reinterpret_cast<const float4*>(RW + N + f);
                            }
                        }
                        return d;
                    },
                    [&](int idx, const DIn2& dd) {
                        const int line = udiv(2 * idx, g.m_na), a = 2 * idx - line * na;
                        if (l0 + line >= nlines) return;
                        const unsigned f = (unsigned)((l0 + line) * na + a);
                        float2 o1a, o2a, o1b, o2b;
                        contract(line, a, f, o1a, o2a);
                        contract(line, a + 1, f + 1, o1b, o2b);
                        if (RU) {
                            float2 r1a = lo(dd.r1), r2a = lo(dd.r2), r1b = hi(dd.r1), r2b = hi(dd.r2);
                            rstep(r1a, r2a, o1a, o2a);
                            rstep(r1b, r2b, o1b, o2b);
                            *reinterpret_cast<float4*>(RW + f) = make_float4(r1a.x, r1a.y, r1b.x, r1b.y);
                            *reinterpret_cast<float4*>(RW + N + f) = make_float4(r2a.x, r2a.y, r2b.x, r2b.y);
                        } else {
                            *reinterpret_cast<float4*>(Y + f) = make_float4(o1a.x, o1a.y, o1b.x, o1b.y);
                            *reinterpret_cast<float4*>(Y + N + f) = make_float4(o2a.x, o2a.y, o2b.x, o2b.y);
                        }
                        dots(o1a, o2a, lo(dd.d1), lo(dd.d2));
                        dots(o1b, o2b, hi(dd.d1), hi(dd.d2));
                    });
            }
        }
        struct DIn { T2 d1, d2, r1, r2; };
        if (!pairs)
            batched_loop<InflightInner<T>::v>(
                L * na,
                [&](int idx) -> DIn {
                    const int line = udiv(idx, g.m_na), a = idx - line * na;
                    const int gl = l0 + line;
                    DIn d;
                    if (gl < nlines) {
                        const unsigned f = (unsigned)(gl * na + a);
                        if (D) { d.d1 = D[f]; d.d2 = D[N + f]; }
                        if (RU) { d.r1 = RW[f]; d.r2 = RW[N + f]; }
                        if (RB) { d.r1 = RBv[f]; d.r2 = RBv[N + f]; }
                    }
                    return d;
                },
                [&](int idx, DIn dd) {
                    const int line = udiv(idx, g.m_na), a = idx - line * na;
                    const int gl = l0 + line;
                    if (gl >= nlines) return;
                    const unsigned f = (unsigned)(gl * na + a);
                    T2 o1, o2;
                    contract(line, a, f, o1, o2);
                    if (RU) {
                        T2 r1 = dd.r1, r2 = dd.r2;
                        rstep(r1, r2, o1, o2);
                        RW[f] = r1; RW[N + f] = r2;
                    } else if (RB) {
                        const double2 b1 = to_d(dd.r1), b2 = to_d(dd.r2), q1 = to_d(o1), q2 = to_d(o2);
                        const float2 e1 = make_float2((float)(b1.x - q1.x), (float)(b1.y - q1.y));
                        const float2 e2 = make_float2((float)(b2.x - q2.x), (float)(b2.y - q2.y));
                        R32[f] = e1; R32[N + f] = e2;
                        P32[f] = e1; P32[N + f] = e2;
                        X32[f] = make_float2(0.0f, 0.0f); X32[N + f] = make_float2(0.0f, 0.0f);
                        const float s1 = e1.x * e1.x + e1.y * e1.y, s2 = e2.x * e2.x + e2.y * e2.y;
                        rn += (double)s1 + (double)s2;
                    } else {
                        Y[f] = o1;
                        Y[N + f] = o2;
                    }
                    dots(o1, o2, dd.d1, dd.d2);
                });
        if (D || CGD || RU || RB) {
            /* block sums into partials[p][tile]: conj(d) . y (slots 0, 1), and
             * inside CG y . y (slot 2) or <r,r> of the updated residual (slot
             * 3); wave shuffles, then the four wave sums in a fixed order */
            __shared__ double red[4][FFT_THREADS / 64];
            double v[4] = {dre, dim, yy, rn};
            const int stride = (CGD || RU) ? 6 : 2;
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
            if ((threadIdx.x & 63) == 0)
#pragma unroll
                for (int c = 0; c < 4; ++c) red[c][threadIdx.x >> 6] = v[c];
            __syncthreads();
            if (threadIdx.x == 0) {
                double* part = args.partials + ((size_t)p * nblk + tile) * stride;
                for (int c = 0; c < 4; ++c) {
                    if (!((c < 2 && D) || (c == 2 && CGD) || (c == 3 && (RU || RB)))) continue;
                    double sum = 0;
                    for (int w = 0; w < FFT_THREADS / 64; ++w) sum += red[c][w];
                    if (c == 3 && RB) args.rpart[(size_t)p * nblk + tile] = sum;
                    else part[c] = sum;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/*  P2 / P4: middle axis, per component                                      */
/* ------------------------------------------------------------------------- */
/* DIR = +1: F = twid_b .* IFFT_b(F) ; DIR = -1: F = FFT_b(conj(twid_b) .* F).
 * A block owns L consecutive a of one (problem, component, c) plane; the
 * middle twiddle depends on (c, b) only and is tabulated once per block.     */
/* largest block of a middle-axis kernel.  The FP64 matrix-core kernels may run
 * several 16-line groups at once with up to 8 waves; compiled for that bound
 * they also hold 84 instead of 116 VGPRs, which made the FP64 middle passes 12%
 * faster on gfx950.  The others keep FFT_THREADS. */
template <typename T, int ENG> struct MidThreads { static constexpr int v = (ENG == ENG_MFMA && sizeof(T) == 8) ? 512 : FFT_THREADS; };

/* waves per SIMD the compiler must leave room for, without spills: the FP32
 * general middle kernel on the matrix cores 6 (62 VGPRs, no AGPRs, instead of
 * 68 + 16), with Stockham 5 (96), and the FP32 outer kernel 6 (80 VGPRs
 * instead of 87; the FP32 outer pass ran 5% faster on gfx950 at 96x83x78).
 * The FP64 kernels need more registers than any such bound leaves.        */
template <typename T, int ENG> struct MidWaves { static constexpr int v = sizeof(T) == 4 ? (ENG == ENG_MFMA ? 6 : 5) : 1; };
template <typename T> struct OuterWaves { static constexpr int v = sizeof(T) == 4 ? 6 : 1; };

template <typename T, int DIR, int ENG>
__global__ void __launch_bounds__((MidThreads<T, ENG>::v)) __attribute__((amdgpu_waves_per_eu(MidWaves<T, ENG>::v)))
k_mid(KAxis args) {
    typedef typename Vec<T>::T2 T2;
    const int p = blockIdx.y / 3, comp = blockIdx.y - 3 * (blockIdx.y / 3);
    if (args.active && !args.active[p]) return;
    const int ntile = (args.g.na + args.L - 1) / args.L;
    const int tile = xcd_tile();
    if (tile >= ntile * args.g.nc) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const GridDev& g = args.g;
    const ProbDev pr = args.prob[p];
    const LinePlanDev& pl = args.plan;
    const int na = g.na, nb = g.nb, L = args.L, P = args.P;
    T2* base = reinterpret_cast<T2*>(smem_raw);
    const Tabs<T2> tb = load_tabs<T2>(pl, base);
    T2* twb = base + lds_head(pl);
    T2* S = twb + nb;
    const int cf = tile / ntile, a0 = (tile - cf * ntile) * L;
    T2* F = reinterpret_cast<T2*>(args.f) + (size_t)p * args.fstride + (size_t)comp * g.N + (size_t)cf * nb * na;
    /* 32-bit element offsets from the uniform plane pointer */
    const double phc = 2.0 * (pr.s0 + pr.s1 * (double)cf) / (double)nb;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) twb[b] = tw_expipi<T2>(phc * (double)b);
    __syncthreads();
    const int Wd = pl.blue ? pl.M : nb;
    /* FP32 with L and na even: each thread moves two consecutive a of a row
     * as one 16-byte access (every pair starts at an even element offset of a
     * 16-byte aligned plane), halving the index arithmetic per element */
    if constexpr (sizeof(T) == 4) {
        const int Lh = L >> 1;
        if (!(L & 1) && !(na & 1) && aligned16(F) && (int)blockDim.x % Lh == 0) {
            /* each thread keeps one pair column ta and steps the rows b by
             * blockDim / Lh: no division per element */
            const int ta = 2 * ((int)threadIdx.x % Lh), rs = (int)blockDim.x / Lh, a = a0 + ta;
            const bool aon = a < na;
            const unsigned rowq = (unsigned)(na >> 1);   /* row pitch in 16-byte units */
            const float4* Fa = reinterpret_cast<const float4*>(F + a);
            constexpr int U = Inflight<T>::v / 2;
            for (int b0 = (int)threadIdx.x / Lh; b0 < Wd; b0 += U * rs) {
                float4 v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int b = b0 + u * rs;
                    v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (aon && b < nb) v[u] = Fa[(unsigned)b * rowq];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int b = b0 + u * rs;
                    if (b < Wd) {
                        float2 v0 = make_float2(v[u].x, v[u].y), v1 = make_float2(v[u].z, v[u].w);
                        if (DIR < 0 && b < nb) { const float2 w = twb[b]; v0 = cmulc(v0, w); v1 = cmulc(v1, w); }
                        S[ta * P + b] = v0;
                        S[(ta + 1) * P + b] = v1;
                    }
                }
            }
            __syncthreads();
            run_stages<T2, MaxE<T>::v, ENG>(S, L, P, pl, tb, NoMid());
            if (aon) {
                float4* Fo = reinterpret_cast<float4*>(F + a);
                for (int b = (int)threadIdx.x / Lh; b < nb; b += rs) {
                    float2 v0 = post_mul<T2, T>(pl, tb, b, S[ta * P + b]);
                    float2 v1 = post_mul<T2, T>(pl, tb, b, S[(ta + 1) * P + b]);
                    if (DIR > 0) { const float2 w = twb[b]; v0 = cmul(v0, w); v1 = cmul(v1, w); }
                    Fo[(unsigned)b * rowq] = make_float4(v0.x, v0.y, v1.x, v1.y);
                }
            }
            return;
        }
    }
    batched_loop<Inflight<T>::v>(
        L * Wd,
        [&](int idx) -> T2 {
            const int b = udiv(idx, args.mL), a = a0 + (idx - b * L);
            T2 v;
            v.x = 0; v.y = 0;
            if (a < na && b < nb) v = F[(unsigned)(b * na + a)];
            return v;
        },
        [&](int idx, T2 v) {
            const int b = udiv(idx, args.mL), ta = idx - b * L;
            if (DIR < 0 && b < nb) v = cmulc(v, twb[b]);
            S[ta * P + b] = v;
        });
    __syncthreads();
    run_stages<T2, MaxE<T>::v, ENG>(S, L, P, pl, tb, NoMid());
    for (int idx = threadIdx.x; idx < L * nb; idx += blockDim.x) {
        const int b = udiv(idx, args.mL), ta = idx - b * L, a = a0 + ta;
        if (a >= na) continue;
        T2 v = post_mul<T2, T>(pl, tb, b, S[ta * P + b]);
        if (DIR > 0) v = cmul(v, twb[b]);
        F[(unsigned)(b * na + a)] = v;
    }
}

/* Middle pass whose plan is one direct DFT on the matrix cores (n = nb <
 * 128, no Bluestein), with the DFT's in-place prologue and its result step
 * fused into the memory phases of the pass:
 *   - the gather loads rows t and n - t of a column (FP32: a pair of columns)
 *     together and writes s_t = x_t + x_{n-t} and d_t = x_t - x_{n-t}
 *     (dft_mfma's first step) into the LDS lines, so the DFT starts on them
 *     after one barrier;
 *   - each wave stores its outputs X_q and X_{n-q} from the accumulators
 *     straight to HBM (16 lanes write 16 consecutive lines of one row), with
 *     the middle twiddle, instead of through LDS and a store loop.
 * The A operands, K steps and result layout are those of dft_mfma.  The host
 * launches it with L a multiple of 16, one wave per 16-row output tile
 * (blockDim = 64 dqt), and for FP32 na even and 16-byte aligned planes.    */
template <typename T, int DIR>
__global__ void __launch_bounds__(FFT_THREADS) __attribute__((amdgpu_waves_per_eu(6))) k_mid_dft(KAxis args) {
    typedef typename Vec<T>::T2 T2;
    typedef typename Mfma<T>::acc_t acc_t;
    const int p = blockIdx.y / 3, comp = blockIdx.y - 3 * (blockIdx.y / 3);
    if (args.active && !args.active[p]) return;
    const int ntile = (args.g.na + args.L - 1) / args.L;
    const int tile = xcd_tile();
    if (tile >= ntile * args.g.nc) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const GridDev& g = args.g;
    const ProbDev pr = args.prob[p];
    const LinePlanDev& pl = args.plan;
    const StageDev& st = pl.st[0];
    const int na = g.na, n = g.nb, L = args.L, P = args.P;
    T2* base = reinterpret_cast<T2*>(smem_raw);
    const Tabs<T2> tb = load_tabs<T2>(pl, base);
    T2* twb = base + lds_head(pl);
    T2* S = twb + n;
    const int cf = tile / ntile, a0 = (tile - cf * ntile) * L;
    T2* F = reinterpret_cast<T2*>(args.f) + (size_t)p * args.fstride + (size_t)comp * g.N + (size_t)cf * n * na;
    const double phc = 2.0 * (pr.s0 + pr.s1 * (double)cf) / (double)n;
    for (int b = threadIdx.x; b < n; b += blockDim.x) twb[b] = tw_expipi<T2>(phc * (double)b);
    __syncthreads();
    const int m = st.dm;
    const bool even = (n & 1) == 0;
    /* gather: unit u = 0 is row 0, u = 1..m the rows u and n - u, and for even
     * n u = m + 1 the row n / 2; each thread keeps one column (FP32: a pair of
     * columns, one 16-byte access) */
    const int nu = m + 1 + (even ? 1 : 0);
    if constexpr (sizeof(T) == 4) {
        const int Lh = L >> 1;
        const int ta = 2 * ((int)threadIdx.x % Lh), us = (int)blockDim.x / Lh;
        const bool aon = a0 + ta < na;
        const unsigned rowq = (unsigned)(na >> 1);
        const float4* Fa = reinterpret_cast<const float4*>(F + a0 + ta);
        for (int u0 = (int)threadIdx.x / Lh; u0 < nu; u0 += 2 * us) {
            float4 x[2], y[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int u = u0 + k * us;
                const int r1 = u <= m ? u : (n >> 1), r2 = (u >= 1 && u <= m) ? n - u : -1;
                x[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                y[k] = x[k];
                if (u < nu && aon) {
                    x[k] = Fa[(unsigned)r1 * rowq];
                    if (r2 >= 0) y[k] = Fa[(unsigned)r2 * rowq];
                }
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int u = u0 + k * us;
                if (u >= nu) continue;
                const int r1 = u <= m ? u : (n >> 1), r2 = (u >= 1 && u <= m) ? n - u : -1;
                float2 x0 = make_float2(x[k].x, x[k].y), x1 = make_float2(x[k].z, x[k].w);
                float2 y0 = make_float2(y[k].x, y[k].y), y1 = make_float2(y[k].z, y[k].w);
                if (DIR < 0) {
                    const float2 w1 = twb[r1];
                    x0 = cmulc(x0, w1); x1 = cmulc(x1, w1);
                }
                if (r2 >= 0) {
                    if (DIR < 0) {
                        const float2 w2 = twb[r2];
                        y0 = cmulc(y0, w2); y1 = cmulc(y1, w2);
                    }
                    S[ta * P + r1] = cadd(x0, y0);
                    S[ta * P + r2] = csub(x0, y0);
                    S[(ta + 1) * P + r1] = cadd(x1, y1);
                    S[(ta + 1) * P + r2] = csub(x1, y1);
                } else {
                    S[ta * P + r1] = x0;
                    S[(ta + 1) * P + r1] = x1;
                }
            }
        }
    } else {
        const int ta = (int)threadIdx.x % L, us = (int)blockDim.x / L;
        const bool aon = a0 + ta < na;
        const T2* Fa = F + a0 + ta;
        for (int u0 = (int)threadIdx.x / L; u0 < nu; u0 += 2 * us) {
            T2 x[2], y[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int u = u0 + k * us;
                const int r1 = u <= m ? u : (n >> 1), r2 = (u >= 1 && u <= m) ? n - u : -1;
                x[k].x = 0; x[k].y = 0;
                y[k] = x[k];
                if (u < nu && aon) {
                    x[k] = Fa[(unsigned)(r1 * na)];
                    if (r2 >= 0) y[k] = Fa[(unsigned)(r2 * na)];
                }
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int u = u0 + k * us;
                if (u >= nu) continue;
                const int r1 = u <= m ? u : (n >> 1), r2 = (u >= 1 && u <= m) ? n - u : -1;
                T2 x0 = x[k], y0 = y[k];
                if (DIR < 0) x0 = cmulc(x0, twb[r1]);
                if (r2 >= 0) {
                    if (DIR < 0) y0 = cmulc(y0, twb[r2]);
                    S[ta * P + r1] = cadd(x0, y0);
                    S[ta * P + r2] = csub(x0, y0);
                } else {
                    S[ta * P + r1] = x0;
                }
            }
        }
    }
    __syncthreads();
    /* the DFT: wave w owns output rows 16 w .. 16 w + 15 */
    const int Q = st.dQ, nks = st.dK >> 2;
    const bool inv = st.sign > 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int q = w * 16 + row;
    const bool qon = q < Q;
    const int dr = (4 * q) % n;
    const int r0 = (q * (kq + 1)) % n;
    T2* Fw = F + a0;
    for (int g0 = 0; g0 < (L >> 4); ++g0) {
        const int line = g0 * 16 + row;
        const T2* Sb = S + line * P;
        acc_t cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0}, sr = {0, 0, 0, 0}, si = {0, 0, 0, 0};
        const T2 x0 = Sb[0];
        T2 h;
        h.x = 0; h.y = 0;
        if (even) h = Sb[n >> 1];
        int r = r0;
        for (int ks = 0; ks < nks; ++ks) {
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
        if (a0 + line < na) {
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
                    if (DIR > 0) lo = cmul(lo, twb[qq]);
                    Fw[(unsigned)(qq * na + line)] = lo;
                    if (qq != 0 && 2 * qq != n) {
                        if (DIR > 0) hi = cmul(hi, twb[n - qq]);
                        Fw[(unsigned)((n - qq) * na + line)] = hi;
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/*  P3: outer axis, per component                                            */
/* ------------------------------------------------------------------------- */
/* mode OUTER_EPS   : F = FFT_c( invB/N .* IFFT_c(F) )
 * mode OUTER_QR    : F = D_c/N .* IFFT_c(F)            (end of Qr / Pr)
 * mode OUTER_QRS   : F = FFT_c( conj(D_c) .* F )       (start of Qrs / Prs)
 * A block owns L consecutive (a, b) positions q of one (problem, component).
 * In EPS mode with Bluestein the inverse post-chirp and the forward pre-chirp
 * are conjugate unit factors whose product is 1; both are left out and only
 * the 1/M remains in the elementwise step.                                  */
/* The outer pass's 1/eps step as a foldable stage of the Stockham engine:
 * mul() is its value for element c of line `line` (the scaled 1/eps product,
 * adding this thread's share of <v, eps^-1 v / N>), applied by the first
 * forward pass as it loads; operator() runs the same step as a pass of its
 * own (used only if no pass follows the step).                             */
template <typename T>
struct EpsFold {
    typedef typename Vec<T>::T2 T2;
    static constexpr bool fold = true;
    const T* invB;
    int q0, plane, nc, L, P, Wd;
    unsigned mL;
    T scale;
    bool pap;
    __device__ T2 mul(int line, int c, T2 v, double& acc) const {
        const int q = q0 + line;
        const T e = (c < nc && q < plane) ? invB[(unsigned)(c * plane + q)] : (T)0;
        const T2 w = cscale(v, e * scale);
        if (pap) acc += (double)v.x * (double)w.x + (double)v.y * (double)w.y;
        return w;
    }
    /* the step as a pass of its own; its sum is lost, so a plan that needs
     * this (a step with no pass after it) must not ask for <p, Ap> */
    __device__ void operator()(T2* Sx) const {
        double acc = 0.0;
        for (int idx = threadIdx.x; idx < L * Wd; idx += blockDim.x) {
            const int c = udiv(idx, mL), tq = idx - c * L;
            Sx[tq * P + c] = mul(tq, c, Sx[tq * P + c], acc);
        }
        __syncthreads();
    }
};

template <typename T, int MODE, int ENG>
__global__ void __launch_bounds__(FFT_THREADS) __attribute__((amdgpu_waves_per_eu(OuterWaves<T>::v))) k_outer(KAxis args) {
    typedef typename Vec<T>::T2 T2;
    const int p = blockIdx.y / 3, comp = blockIdx.y - 3 * (blockIdx.y / 3);
    if (args.active && !args.active[p]) return;
    const int tile = xcd_tile();
    double* pap = (MODE == OUTER_EPS && args.pap) ? args.pap + ((size_t)p * 3 + comp) * gridDim.x + tile : nullptr;
    if (tile * args.L >= args.g.na * args.g.nb) {
        if (pap && threadIdx.x == 0) *pap = 0.0;
        return;
    }
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const GridDev& g = args.g;
    const ProbDev pr = args.prob[p];
    const LinePlanDev& pl = args.plan;
    const int nc = g.nc, L = args.L, P = args.P;
    const int plane = g.na * g.nb;
    T2* base = reinterpret_cast<T2*>(smem_raw);
    const Tabs<T2> tb = load_tabs<T2>(pl, base);
    T2* S = base + lds_head(pl);
    const int q0 = tile * L;
    T2* F = reinterpret_cast<T2*>(args.f) + (size_t)p * args.fstride + (size_t)comp * g.N;
    const int Wd = pl.blue ? pl.M : nc;
    /* FP32 with L and the plane even: two consecutive q per 16-byte access (and
     * per 8-byte load of 1/eps), as in the middle-axis kernel */
    bool pairs = false;
    const int Lh = L >> 1;
    if constexpr (sizeof(T) == 4)
        pairs = !(L & 1) && !(plane & 1) && aligned16(F) && (MODE != OUTER_EPS || aligned8(args.invB));
    if constexpr (sizeof(T) == 4) {
        if (pairs)
            batched_loop<Inflight<T>::v / 2>(
                Lh * Wd,
                [&](int idx) -> float4 {
                    const int c = udiv(2 * idx, args.mL), q = q0 + (2 * idx - c * L);
                    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (q < plane && c < nc) v = *reinterpret_cast<const float4*>(F + (unsigned)(c * plane + q));
                    return v;
                },
                [&](int idx, float4 v) {
                    const int c = udiv(2 * idx, args.mL), tq = 2 * idx - c * L;
                    float2 v0 = make_float2(v.x, v.y), v1 = make_float2(v.z, v.w);
                    if (MODE == OUTER_QRS && c < nc) {
                        const float2 w = tw_expipi<T2>(2.0 * pr.r0 * (double)c / (double)nc);
                        v0 = cmulc(v0, w); v1 = cmulc(v1, w);
                    }
                    S[tq * P + c] = v0;
                    S[(tq + 1) * P + c] = v1;
                });
    }
    if (!pairs)
        batched_loop<Inflight<T>::v>(
            L * Wd,
            [&](int idx) -> T2 {
                const int c = udiv(idx, args.mL), q = q0 + (idx - c * L);
                T2 v;
                v.x = 0; v.y = 0;
                if (q < plane && c < nc) v = F[(unsigned)(c * plane + q)];
                return v;
            },
            [&](int idx, T2 v) {
                const int c = udiv(idx, args.mL), tq = idx - c * L;
                if (MODE == OUTER_QRS && c < nc) v = cmulc(v, tw_expipi<T2>(2.0 * pr.r0 * (double)c / (double)nc));
                S[tq * P + c] = v;
            });
    __syncthreads();
    if (MODE == OUTER_EPS) {
        const T* invB = reinterpret_cast<const T*>(args.invB) + (size_t)comp * g.N;
        const T scale = (T)(1.0 / (double)g.N / (pl.blue ? (double)pl.M : 1.0));
        const unsigned mL = args.mL;
        double vw = 0.0;
        auto eps_step = [&](T2* Sx) {
            if constexpr (sizeof(T) == 4) {
                if (pairs)
                    batched_loop<InflightEps<T>::v / 2>(
                        Lh * Wd,
                        [&](int idx) -> float2 {
                            const int c = udiv(2 * idx, mL), q = q0 + (2 * idx - c * L);
                            return (q < plane && c < nc) ? *reinterpret_cast<const float2*>(invB + (unsigned)(c * plane + q))
                                                         : make_float2(0.f, 0.f);
                        },
                        [&](int idx, float2 e) {
                            const int c = udiv(2 * idx, mL), tq = 2 * idx - c * L;
                            const float2 v0 = Sx[tq * P + c], w0 = cscale(v0, e.x * scale);
                            const float2 v1 = Sx[(tq + 1) * P + c], w1 = cscale(v1, e.y * scale);
                            Sx[tq * P + c] = w0;
                            Sx[(tq + 1) * P + c] = w1;
                            if (pap) vw += (double)v0.x * (double)w0.x + (double)v0.y * (double)w0.y +
                                           (double)v1.x * (double)w1.x + (double)v1.y * (double)w1.y;
                        });
            }
            if (!pairs)
                batched_loop<InflightEps<T>::v>(
                    L * Wd,
                    [&](int idx) -> T {
                        const int c = udiv(idx, mL), q = q0 + (idx - c * L);
                        return (q < plane && c < nc) ? invB[(unsigned)(c * plane + q)] : (T)0;
                    },
                    [&](int idx, T e) {
                        const int c = udiv(idx, mL), tq = idx - c * L;
                        const T2 v = Sx[tq * P + c], w = cscale(v, e * scale);
                        Sx[tq * P + c] = w;
                        if (pap) vw += (double)v.x * (double)w.x + (double)v.y * (double)w.y;
                    });
            __syncthreads();
        };
        /* FP32 Stockham plans fold the step into the loads of the first forward
         * pass (no LDS round trip of its own); 1/eps is shared by all problems,
         * so its scattered reads there are mostly served by the caches */
        bool folded = false;
        if constexpr (ENG != ENG_MFMA) {
            const EpsFold<T> em{invB, q0, plane, nc, L, P, Wd, mL, scale, pap != nullptr};
            vw += run_stages<T2, MaxE<T>::v, ENG, EpsFold<T>, OuterLF<T>::v>(S, L, P, pl, tb, em, args.mL);
            folded = true;
        }
        if (!folded) run_stages<T2, MaxE<T>::v, ENG, decltype(eps_step), OuterLF<T>::v>(S, L, P, pl, tb, eps_step, args.mL);
        if (pap) {
            /* block sum in a fixed order: wave shuffles, then the wave sums */
            __shared__ double red[FFT_THREADS / 64];
            for (int o = 32; o > 0; o >>= 1) vw += __shfl_xor(vw, o);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = vw;
            __syncthreads();
            if (threadIdx.x == 0) {
                double sum = 0.0;
                for (int w = 0; w < FFT_THREADS / 64; ++w) sum += red[w];
                *pap = sum;
            }
        }
    } else {
        run_stages<T2, MaxE<T>::v, ENG, NoMid, OuterLF<T>::v>(S, L, P, pl, tb, NoMid(), args.mL);
    }
    const double invN = 1.0 / (double)g.N;
    if constexpr (sizeof(T) == 4) {
        if (pairs)
            for (int idx = threadIdx.x; idx < Lh * nc; idx += FFT_THREADS) {
                const int c = udiv(2 * idx, args.mL), tq = 2 * idx - c * L, q = q0 + tq;
                if (q >= plane) continue;
                float2 v0 = post_mul<T2, T>(pl, tb, c, S[tq * P + c]);
                float2 v1 = post_mul<T2, T>(pl, tb, c, S[(tq + 1) * P + c]);
                if (MODE == OUTER_QR) {
                    const float2 w = from_d<T2>(cscale(expipi(2.0 * pr.r0 * (double)c / (double)nc), invN));
                    v0 = cmul(v0, w); v1 = cmul(v1, w);
                }
                *reinterpret_cast<float4*>(F + (unsigned)(c * plane + q)) = make_float4(v0.x, v0.y, v1.x, v1.y);
            }
    }
    if (!pairs)
        for (int idx = threadIdx.x; idx < L * nc; idx += FFT_THREADS) {
            const int c = udiv(idx, args.mL), tq = idx - c * L, q = q0 + tq;
            if (q >= plane) continue;
            T2 v = post_mul<T2, T>(pl, tb, c, S[tq * P + c]);
            if (MODE == OUTER_QR) v = cmul(v, from_d<T2>(cscale(expipi(2.0 * pr.r0 * (double)c / (double)nc), invN)));
            F[(unsigned)(c * plane + q)] = v;
        }
}

/* ------------------------------------------------------------------------- */
/*  Layout helpers                                                            */
/* ------------------------------------------------------------------------- */
/* internal [c][b][a] <-> reference real-space x-fastest for General lattices,
 * i.e. [x][y][z] (z fastest) <-> [z][y][x] (x fastest).  dir=+1: int -> ref.  */
template <typename T2>
__global__ void k_xz_transpose(const T2* __restrict__ in, T2* __restrict__ out, int nx, int ny, int nz, int dir,
                               size_t stride, int ncomp) {
    __shared__ T2 tile[32][33];
    const int comp = blockIdx.z % ncomp, p = blockIdx.z / ncomp;
    const int y = blockIdx.y;
    const int tilesz = (nz + 31) / 32;
    const int tx0 = (blockIdx.x / tilesz) * 32, tz0 = (blockIdx.x % tilesz) * 32;
    const size_t N = (size_t)nx * ny * nz;
    const T2* I = in + p * stride + comp * N;
    T2* O = out + p * stride + comp * N;
    const int lx = threadIdx.x, ly = threadIdx.y;
    if (dir > 0) {
        for (int r = ly; r < 32; r += 8) {
            int x = tx0 + r, z = tz0 + lx;
            if (x < nx && z < nz) tile[r][lx] = I[((size_t)x * ny + y) * nz + z];
        }
        __syncthreads();
        for (int r = ly; r < 32; r += 8) {
            int z = tz0 + r, x = tx0 + lx;
            if (x < nx && z < nz) O[x + (size_t)nx * (y + (size_t)ny * z)] = tile[lx][r];
        }
    } else {
        for (int r = ly; r < 32; r += 8) {
            int z = tz0 + r, x = tx0 + lx;
            if (x < nx && z < nz) tile[r][lx] = I[x + (size_t)nx * (y + (size_t)ny * z)];
        }
        __syncthreads();
        for (int r = ly; r < 32; r += 8) {
            int x = tx0 + r, z = tz0 + lx;
            if (x < nx && z < nz) O[((size_t)x * ny + y) * nz + z] = tile[lx][r];
        }
    }
}

/* reduced vector: reference layout [x1(Nd); x2(Nd)] <-> padded [x1(N); x2(N)] */
template <typename T2>
__global__ void k_pack(const T2* __restrict__ in, T2* __restrict__ out, int N, int Nd, int dir) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int off = N - Nd;
    if (i >= 2 * N) return;
    const int half = i / N, f = i - half * N;
    if (dir > 0) {           /* ref -> padded */
        T2 v; v.x = 0; v.y = 0;
        if (f >= off) v = in[half * Nd + (f - off)];
        out[i] = v;
    } else {                 /* padded -> ref */
        if (f >= off) out[half * Nd + (f - off)] = in[i];
    }
}

/* Sigma_r = |Lambda| per Fourier mode and its inverse (0 for the removed Gamma
 * mode), FAME_Matrix_Lambdas.cu:193-205 / :442-454, duplicated over [x1;x2] by
 * the callers through a modulo-N index.                                       */
__global__ void k_sigma(GridDev g, const ProbDev* prob, double* sig, double* sig_inv) {
    const int p = blockIdx.y;
    const ProbDev pr = prob[p];
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < g.N; f += gridDim.x * blockDim.x) {
        const int a = f % g.na, bc = f / g.na, b = bc % g.nb, c = bc / g.nb;
        double s = 0.0, si = 0.0;
        if (!(pr.gamma && f == 0)) {
            double2 lx, ly, lz;
            lambdas(g, pr, a, b, c, lx, ly, lz);
            double Lq = lx.x * lx.x + lx.y * lx.y + ly.x * ly.x + ly.y * ly.y + lz.x * lz.x + lz.y * lz.y;
            s = sqrt(Lq);
            si = 1.0 / s;
        }
        sig[(size_t)p * g.N + f] = s;
        sig_inv[(size_t)p * g.N + f] = si;
    }
}

}  // namespace fame

/* ========================================================================= */
/*  Host launchers                                                            */
/* ========================================================================= */
using namespace fame;

size_t fame_line_lds(const LinePlanDev& pl, int prec, int nl, int P, int extra) {
    return (size_t)(prec == 8 ? 16 : 8) * ((size_t)lds_head(pl) + (size_t)nl * P + (size_t)extra);
}

template <typename T, int ENG>
static hipError_t launch_inner(const KInner& a, int expand, int nprob, hipStream_t s) {
    const int nlines = a.g.nb * a.g.nc;
    dim3 grid(xcd_round((nlines + a.L - 1) / a.L), nprob);
    const size_t lds = fame_line_lds(a.plan, (int)sizeof(T), 3 * a.L, a.P, inner_lds_extra(a.g.na, (int)sizeof(T), a.L));
    if (a.g.pi_guard) {
        if (expand) hipLaunchKernelGGL((k_inner<T, 1, ENG, PI_GUARD>), grid, dim3(FFT_THREADS), lds, s, a);
        else hipLaunchKernelGGL((k_inner<T, 0, ENG, PI_GUARD>), grid, dim3(FFT_THREADS), lds, s, a);
    } else if (a.g.general && a.pikind == 0 && (sizeof(T) == 4 || !expand)) {
        /* the FP64 expansion keeps the vector form: 5% faster there on gfx950 */
        if constexpr (sizeof(T) == 4)
            if (expand) { hipLaunchKernelGGL((k_inner<T, 1, ENG, PI_GEN>), grid, dim3(FFT_THREADS), lds, s, a); return hipGetLastError(); }
        hipLaunchKernelGGL((k_inner<T, 0, ENG, PI_GEN>), grid, dim3(FFT_THREADS), lds, s, a);
    } else {
        if (expand) hipLaunchKernelGGL((k_inner<T, 1, ENG, PI_ANY>), grid, dim3(FFT_THREADS), lds, s, a);
        else hipLaunchKernelGGL((k_inner<T, 0, ENG, PI_ANY>), grid, dim3(FFT_THREADS), lds, s, a);
    }
    return hipGetLastError();
}

template <typename T, int ENG>
static hipError_t launch_mid(const KAxis& a, int dir, int nprob, hipStream_t s) {
    const int ntile = (a.g.na + a.L - 1) / a.L;
    dim3 grid(xcd_round(ntile * a.g.nc), 3 * nprob);
    const size_t lds = fame_line_lds(a.plan, (int)sizeof(T), a.L, a.P, a.g.nb);
    /* a single direct DFT: the fused kernel (its preconditions, see there; FP32
     * also needs na even and 16-byte aligned planes for its paired accesses) */
    if (ENG == ENG_MFMA && a.plan.nst == 1 && a.plan.st[0].R == DFT_MFMA && !a.plan.blue &&
        a.plan.post == POST_NONE && a.plan.st[0].dn == a.g.nb && a.L % 16 == 0 &&
        a.plan.st[0].dqt <= FFT_THREADS / 64 &&
        (sizeof(T) == 8 || (a.g.na % 2 == 0 && ((uintptr_t)a.f & 15) == 0 && a.fstride % 2 == 0))) {
        const dim3 blk(64 * a.plan.st[0].dqt);
        if (dir > 0) hipLaunchKernelGGL((k_mid_dft<T, 1>), grid, blk, lds, s, a);
        else hipLaunchKernelGGL((k_mid_dft<T, -1>), grid, blk, lds, s, a);
        return hipGetLastError();
    }
    /* the matrix-core DFT keeps one wave per (16-row output tile, 16-line group)
     * busy: launch those waves, at most MidThreads / 64 */
    int nthr = FFT_THREADS;
    if (ENG == ENG_MFMA) {
        int q = 1;
        for (int i = 0; i < a.plan.nst; ++i)
            if (a.plan.st[i].R == DFT_MFMA) q = a.plan.st[i].dqt > q ? a.plan.st[i].dqt : q;
        const int groups = (a.L + 15) / 16;
        int G = (MidThreads<T, ENG>::v / 64) / q;
        if (G > groups) G = groups;
        if (G < 1) G = 1;
        nthr = 64 * q * G;
    }
    if (dir > 0) hipLaunchKernelGGL((k_mid<T, 1, ENG>), grid, dim3(nthr), lds, s, a);
    else hipLaunchKernelGGL((k_mid<T, -1, ENG>), grid, dim3(nthr), lds, s, a);
    return hipGetLastError();
}

template <typename T, int ENG>
static hipError_t launch_outer(const KAxis& a, int mode, int nprob, hipStream_t s) {
    const int plane = a.g.na * a.g.nb;
    dim3 grid(xcd_round((plane + a.L - 1) / a.L), 3 * nprob);
    const size_t lds = fame_line_lds(a.plan, (int)sizeof(T), a.L, a.P, 0);
    switch (mode) {
        case OUTER_EPS: hipLaunchKernelGGL((k_outer<T, OUTER_EPS, ENG>), grid, dim3(FFT_THREADS), lds, s, a); break;
        case OUTER_QR: hipLaunchKernelGGL((k_outer<T, OUTER_QR, ENG>), grid, dim3(FFT_THREADS), lds, s, a); break;
        default: hipLaunchKernelGGL((k_outer<T, OUTER_QRS, ENG>), grid, dim3(FFT_THREADS), lds, s, a); break;
    }
    return hipGetLastError();
}

hipError_t fame_launch_inner(const KInner& a, int expand, int prec, int nprob, hipStream_t s) {
    if (plan_engine(a.plan) == ENG_MFMA)
        return prec == 8 ? launch_inner<double, ENG_MFMA>(a, expand, nprob, s) : launch_inner<float, ENG_MFMA>(a, expand, nprob, s);
    return prec == 8 ? launch_inner<double, ENG_STOCKHAM>(a, expand, nprob, s) : launch_inner<float, ENG_STOCKHAM>(a, expand, nprob, s);
}
hipError_t fame_launch_mid(const KAxis& a, int dir, int prec, int nprob, hipStream_t s) {
    if (plan_engine(a.plan) == ENG_MFMA)
        return prec == 8 ? launch_mid<double, ENG_MFMA>(a, dir, nprob, s) : launch_mid<float, ENG_MFMA>(a, dir, nprob, s);
    return prec == 8 ? launch_mid<double, ENG_STOCKHAM>(a, dir, nprob, s) : launch_mid<float, ENG_STOCKHAM>(a, dir, nprob, s);
}
hipError_t fame_launch_outer(const KAxis& a, int mode, int prec, int nprob, hipStream_t s) {
    if (plan_engine(a.plan) == ENG_MFMA)
        return prec == 8 ? launch_outer<double, ENG_MFMA>(a, mode, nprob, s) : launch_outer<float, ENG_MFMA>(a, mode, nprob, s);
    return prec == 8 ? launch_outer<double, ENG_STOCKHAM>(a, mode, nprob, s) : launch_outer<float, ENG_STOCKHAM>(a, mode, nprob, s);
}

hipError_t fame_launch_xz_transpose(const void* in, void* out, int nx, int ny, int nz, int dir, size_t stride,
                                    int ncomp, int nprob, int prec, hipStream_t s) {
    const int tiles = ((nx + 31) / 32) * ((nz + 31) / 32);
    dim3 grid(tiles, ny, ncomp * nprob), block(32, 8);
    if (prec == 8)
        hipLaunchKernelGGL(k_xz_transpose<double2>, grid, block, 0, s, (const double2*)in, (double2*)out, nx, ny, nz, dir, stride, ncomp);
    else
        hipLaunchKernelGGL(k_xz_transpose<float2>, grid, block, 0, s, (const float2*)in, (float2*)out, nx, ny, nz, dir, stride, ncomp);
    return hipGetLastError();
}

hipError_t fame_launch_pack(const void* in, void* out, int N, int Nd, int dir, int prec, hipStream_t s) {
    dim3 grid((2 * N + 255) / 256);
    if (prec == 8) hipLaunchKernelGGL(k_pack<double2>, grid, dim3(256), 0, s, (const double2*)in, (double2*)out, N, Nd, dir);
    else hipLaunchKernelGGL(k_pack<float2>, grid, dim3(256), 0, s, (const float2*)in, (float2*)out, N, Nd, dir);
    return hipGetLastError();
}

hipError_t fame_launch_sigma(const GridDev& g, const ProbDev* prob, double* sig, double* sig_inv, int nprob, hipStream_t s) {
    hipLaunchKernelGGL(k_sigma, dim3(512, nprob), dim3(256), 0, s, g, prob, sig, sig_inv);
    return hipGetLastError();
}
