/*
 * fame_blas.hip -- batched vector, reduction and small-GEMM kernels for the
 * CG (CG.cu:5-87) and Lanczos (Lanczos_decomp_gpu.cu, Lanczos_LockPurge_gpu.cu,
 * Lanczos_Isotropic.cu:93) layers.  Every kernel works on a batch of problems
 * (blockIdx.y) and honours a per-problem mask, so converged problems cost
 * nothing.  Scalars (alpha, beta, norms) stay on the device: no host round trip
 * per BLAS-1 call as with CUBLAS_POINTER_MODE_HOST in the reference.
 *
 * Reductions are two-stage: blocks write partial sums to `partials`
 * [problem][block][ncomp], one block per problem then sums them in a fixed
 * order (deterministic results run to run).
 */
#include "fame_kernels.h"

using namespace fame;

#define RB 256   /* threads per block for vector kernels */

template <typename T> struct V2;
template <> struct V2<double> { typedef double2 t; };
template <> struct V2<float> { typedef float2 t; };

__device__ __forceinline__ double2 d2(double2 v) { return v; }
__device__ __forceinline__ double2 d2(float2 v) { return make_double2(v.x, v.y); }

/* sum of `NC` doubles over the block (blockDim.x a multiple of 64, at most RB):
 * wave-level butterfly shuffles, then the per-wave sums in a fixed order, so
 * every thread receives the same, run-to-run deterministic total            */
template <int NC>
__device__ __forceinline__ void block_sum_all(double (&v)[NC]) {
    __shared__ double wred[NC][RB / 64];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) wred[c][w] = v[c];
    }
    __syncthreads();
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += wred[c][i];
        v[c] = s;
    }
}

/* block sum of `ncomp` doubles per thread -> partials[p][blockIdx.x][c] */
template <int NC>
__device__ __forceinline__ void block_partials(double (&v)[NC], double* partials, int p) {
    block_sum_all<NC>(v);
    if (threadIdx.x == 0) {
        double* o = partials + ((size_t)p * gridDim.x + blockIdx.x) * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) o[c] = v[c];
    }
}

/* sum over the nblk partials of problem p, component c of ncomp, by the whole block */
__device__ __forceinline__ double sum_partials(const double* partials, int p, int nblk, int ncomp, int c) {
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) v[0] += partials[((size_t)p * nblk + b) * ncomp + c];
    block_sum_all<1>(v);
    return v[0];
}

/* ---------------------------------------------------------------- CG ---- */
template <typename T>
__global__ void __launch_bounds__(RB) k_cg_begin(typename V2<T>::t* x, typename V2<T>::t* r, typename V2<T>::t* pv,
                                                 const typename V2<T>::t* b, double* partials, size_t n, size_t stride,
                                                 const int* mask) {
    typedef typename V2<T>::t T2;
    const int p = blockIdx.y;
    double acc[1] = {0.0};
    if (!mask || mask[p]) {
        const size_t o = (size_t)p * stride;
        for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
            T2 v = b[o + i];
            T2 z; z.x = 0; z.y = 0;
            x[o + i] = z; r[o + i] = v; pv[o + i] = v;
            T s = v.x * v.x + v.y * v.y;
            acc[0] += (double)s;
        }
    }
    block_partials<1>(acc, partials, p);
}

__global__ void k_cg_init_scal(CgScal* sc, const double* partials, int nblk, double tol, int maxit, int* active,
                               const int* mask) {
    const int p = blockIdx.x;
    const double s = sum_partials(partials, p, nblk, 1, 0);
    if (threadIdx.x != 0) return;
    CgScal c;
    c.rr = s; c.rr0 = s; c.pap_re = c.pap_im = 0; c.alpha_re = c.alpha_im = 0; c.beta = 0;
    c.tol2 = tol * tol; c.iters = 1; c.maxit = maxit;
    c.active = (!mask || mask[p]) && (s > c.tol2) && (1 <= maxit);
    c.pend = 0;
    sc[p] = c;
    active[p] = c.active;
}

/* partials per inner block (6 doubles): <p,Ap> (slots 0, 1) and <Ap,Ap>
 * (slot 2) from the contracting pass, <r,r> of the current residual (slot 5)
 * from the expanding pass; slots 3 and 4 are not used */
__global__ void k_cg_alpha_beta(CgScal* sc, const double* partials, int nblk, int* active) {
    const int p = blockIdx.x;
    if (!active[p]) return;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const double* q = partials + ((size_t)p * nblk + b) * 6;
        v[0] += q[0]; v[1] += q[1]; v[2] += q[2]; v[3] += q[5];
    }
    block_sum_all<4>(v);
    if (threadIdx.x != 0) return;
    CgScal& c = sc[p];
    const double re = v[0], im = v[1];          /* <p, Ap> */
    const double aa = v[2];                     /* <Ap, Ap> */
    const double rr = v[3];                     /* <r, r>, summed from the vector */
    c.pap_re = re; c.pap_im = im;
    /* a = r1 / dot  (CG.cu:60-63) */
    const double temp = re * re + im * im;
    const double ar = rr * re / temp, ai = -rr * im / temp;
    c.alpha_re = ar; c.alpha_im = ai;
    c.pend = 1;
    /* r1 of r - a Ap from one step of the recurrence off the summed <r,r>:
     * |r - a Ap|^2 = rr - 2 Re(a <r,Ap>) + |a|^2 <Ap,Ap>, where <r,Ap> = <p,Ap>
     * because p = r + beta p_old with p_old A-conjugate to p, and a <p,Ap> = rr,
     * so r1 = |a|^2 <Ap,Ap> - rr, to rounding and the loss of conjugacy between
     * consecutive directions.  b = r1 / r0 and the stopping rule (CG.cu:66-78)
     * use it, the next alpha the summed norm again */
    const double s = fmax(0.0, (ar * ar + ai * ai) * aa - rr);
    c.beta = s / rr;
    c.rr = s;
    c.iters += 1;
    c.active = (s > c.tol2) && (c.iters <= c.maxit);
    active[p] = c.active;
}

/* alpha = <r,r> / <p,Ap> with <p,Ap> summed from the outer pass's block sums
 * (3 npap per problem); the solution update x += alpha p becomes pending */
__global__ void k_cg_alpha(CgScal* sc, const double* pap, int npap, const int* active) {
    const int p = blockIdx.x;
    if (!active[p]) return;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < 3 * npap; b += blockDim.x) v[0] += pap[(size_t)p * 3 * npap + b];
    block_sum_all<1>(v);
    if (threadIdx.x != 0) return;
    CgScal& c = sc[p];
    /* a = r1 / dot  (CG.cu:60-63); the quadratic form is real */
    c.pap_re = v[0]; c.pap_im = 0.0;
    c.alpha_re = c.rr / v[0]; c.alpha_im = 0.0;
    c.pend = 1;
}

/* <r,r> of the residual the contracting pass updated (slot 3 of its 6-double
 * block partials): b = r1 / r0 and the stopping rule (CG.cu:66-78) */
__global__ void k_cg_beta(CgScal* sc, const double* partials, int nblk, int* active) {
    const int p = blockIdx.x;
    if (!active[p]) return;
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) v[0] += partials[((size_t)p * nblk + b) * 6 + 3];
    block_sum_all<1>(v);
    if (threadIdx.x != 0) return;
    CgScal& c = sc[p];
    const double s = v[0];
    c.beta = s / c.rr;
    c.rr = s;
    c.iters += 1;
    c.active = (s > c.tol2) && (c.iters <= c.maxit);
    active[p] = c.active;
}

/* the solution update x += alpha p still pending at the end of a solve */
template <typename T>
__global__ void __launch_bounds__(RB) k_cg_finish_x(typename V2<T>::t* x, const typename V2<T>::t* pv,
                                                    const CgScal* sc, size_t n, size_t stride) {
    typedef typename V2<T>::t T2;
    const int p = blockIdx.y;
    if (!sc[p].pend) return;
    const T ar = (T)sc[p].alpha_re, ai = (T)sc[p].alpha_im;
    const size_t o = (size_t)p * stride;
    for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
        T2 pp = pv[o + i], xx = x[o + i];
        xx.x += ar * pp.x - ai * pp.y; xx.y += ar * pp.y + ai * pp.x;
        x[o + i] = xx;
    }
}

/* start of an FP32 solve from an FP64 right-hand side (the refinement's
 * residual): x = 0, r = p = b rounded to FP32, and the partials of <r,r> */
__global__ void __launch_bounds__(RB) k_cg_begin_from64(float2* x, float2* r, float2* pv, const double2* b,
                                                        double* partials, size_t n, size_t stride, const int* mask) {
    const int p = blockIdx.y;
    double acc[1] = {0.0};
    if (!mask || mask[p]) {
        const size_t o = (size_t)p * stride;
        for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
            const double2 w = b[o + i];
            const float2 v = make_float2((float)w.x, (float)w.y);
            x[o + i] = make_float2(0.0f, 0.0f); r[o + i] = v; pv[o + i] = v;
            const float s = v.x * v.x + v.y * v.y;
            acc[0] += (double)s;
        }
    }
    block_partials<1>(acc, partials, p);
}

/* ------------------------------------------------------------ helpers ---- */
template <typename T>
__global__ void k_scale_real(typename V2<T>::t* y, size_t ystride, const typename V2<T>::t* x, size_t xstride,
                             const double* s, size_t n, size_t sstride, size_t smod, const int* mask) {
    typedef typename V2<T>::t T2;
    const int p = blockIdx.y;
    if (mask && !mask[p]) return;
    for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
        T2 v = x[(size_t)p * xstride + i];
        T f = (T)s[(size_t)p * sstride + (i % smod)];
        v.x *= f; v.y *= f;
        y[(size_t)p * ystride + i] = v;
    }
}

template <typename TD, typename TS>
__global__ void k_convert(typename V2<TD>::t* d, const typename V2<TS>::t* s, size_t n, size_t dstride, size_t sstride,
                          const int* mask) {
    const int p = blockIdx.y;
    if (mask && !mask[p]) return;
    for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
        typename V2<TS>::t v = s[(size_t)p * sstride + i];
        typename V2<TD>::t w; w.x = (TD)v.x; w.y = (TD)v.y;
        d[(size_t)p * dstride + i] = w;
    }
}

template <typename T>
__global__ void __launch_bounds__(RB) k_norm2(const typename V2<T>::t* x, double* partials, size_t n, size_t stride) {
    const int p = blockIdx.y;
    double acc[1] = {0.0};
    for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
        double2 v = d2(x[(size_t)p * stride + i]);
        acc[0] += v.x * v.x + v.y * v.y;
    }
    block_partials<1>(acc, partials, p);
}

/* grid (nprob, ncomp): one block per problem and component */
__global__ void __launch_bounds__(RB) k_reduce_partials(const double* partials, int nblk, int ncomp, double* out) {
    const int p = blockIdx.x, c = blockIdx.y;
    const double s = sum_partials(partials, p, nblk, ncomp, c);
    if (threadIdx.x == 0) out[(size_t)p * ncomp + c] = s;
}

/* ------------------------------------------------------------ Lanczos ---- */
/* a Lanczos basis element, loaded non-temporally: the products stream the
 * basis once per call, and keeping it out of the caches measured 13% (h) and
 * 5% (n) more bandwidth at FCC 96 sizes */
__device__ __forceinline__ double2 ld_basis(const double2* a) {
    double2 v;
    v.x = __builtin_nontemporal_load(&a->x);
    v.y = __builtin_nontemporal_load(&a->y);
    return v;
}

/* h[j] = sum_i conj(V[j][i]) w[i], j < m  (partials of 2m doubles per block).
 * Columns go CB at a time, so that w is read once per CB columns and CB + 1
 * loads are in flight per element; CB = 4 for m <= 4 (the three-term step,
 * where a wider group only adds reduction work), else 8.                    */
template <int CB>
__global__ void __launch_bounds__(RB) k_gemv_h(const double2* V, const double2* w, double* partials, int m, size_t n,
                                               size_t cstride, size_t pstride, size_t wstride, const int* mask) {
    const int p = blockIdx.y;
    double* out = partials + ((size_t)p * gridDim.x + blockIdx.x) * 2 * m;
    if (mask && !mask[p]) {
        for (int t = threadIdx.x; t < 2 * m; t += blockDim.x) out[t] = 0.0;
        return;
    }
    const double2* Vp = V + (size_t)p * pstride;
    const double2* wp = w + (size_t)p * wstride;
    for (int j0 = 0; j0 < m; j0 += CB) {
        const int nj = min(CB, m - j0);
        double acc[2 * CB];
#pragma unroll
        for (int q = 0; q < 2 * CB; ++q) acc[q] = 0.0;
        for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
            const double2 x = wp[i];
            double2 v[CB];
#pragma unroll
            for (int q = 0; q < CB; ++q) v[q] = q < nj ? ld_basis(Vp + (size_t)(j0 + q) * cstride + i) : make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < CB; ++q) {
                acc[2 * q] += v[q].x * x.x + v[q].y * x.y;
                acc[2 * q + 1] += v[q].x * x.y - v[q].y * x.x;
            }
        }
        block_sum_all<2 * CB>(acc);
        if ((int)threadIdx.x < 2 * nj) {
            double r = acc[0];
#pragma unroll
            for (int t = 1; t < 2 * CB; ++t) if ((int)threadIdx.x == t) r = acc[t];
            out[2 * j0 + threadIdx.x] = r;
        }
        __syncthreads();
    }
}

/* w[i] -= sum_j V[j][i] h[j]  (h complex, 2m doubles per problem) */
__global__ void k_gemv_n(const double2* V, double2* w, const double* h, int m, size_t n, size_t cstride,
                         size_t pstride, size_t wstride, size_t hstride, const int* mask) {
    const int p = blockIdx.y;
    if (mask && !mask[p]) return;
    const double2* Vp = V + (size_t)p * pstride;
    double2* wp = w + (size_t)p * wstride;
    const double* hp = h + (size_t)p * hstride;
    for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
        double2 acc = wp[i];
        int j = 0;
        /* four column loads in flight, accumulated in column order */
        for (; j + 4 <= m; j += 4) {
            double2 v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = ld_basis(Vp + (size_t)(j + q) * cstride + i);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double hr = hp[2 * (j + q)], hi = hp[2 * (j + q) + 1];
                acc.x -= v[q].x * hr - v[q].y * hi;
                acc.y -= v[q].x * hi + v[q].y * hr;
            }
        }
        for (; j < m; ++j) {
            double2 v = ld_basis(Vp + (size_t)j * cstride + i);
            double hr = hp[2 * j], hi = hp[2 * j + 1];
            acc.x -= v.x * hr - v.y * hi;
            acc.y -= v.x * hi + v.y * hr;
        }
        wp[i] = acc;
    }
}

/* Gram matrix G[j][l] = sum_i conj(V[j][i]) V[l][i] over the upper triangle,
 * on the matrix cores (v_mfma_f64_16x16x4f64).  A block stages GKT rows of all
 * m columns (m <= 16*MT, padded to 16-column tiles) in LDS with an odd pitch,
 * and each of the 4 waves accumulates a fixed set of upper-triangle 16x16 tile
 * pairs (pair w + 4q, q < QM; MT(MT+1)/2 <= 4 QM pairs):
 *   Re G += Vr^T Vr + Vi^T Vi,   Im G += Vr^T Vi - Vi^T Vr.
 * Operands: lane l holds A[l % 16][l / 16] = conj V_j(k) and B[l / 16][l % 16]
 * = V_l(k); result element v of lane l is row l / 16 + 4 v, column l % 16.
 * Instances: (GKT 32, MT 4, QM 3) for m <= 64 and (16, 8, 9) for m <= 128,
 * both 34-35 KB of LDS.  Partial per block: 2 m^2 doubles (lower triangle zero). */
typedef double gacc_t __attribute__((ext_vector_type(4)));
template <int GKT, int MT, int QM>
__global__ void __launch_bounds__(RB) k_gram(const double2* V, double* partials, int m, size_t n, size_t cstride,
                                             size_t pstride, const int* mask) {
    const int p = blockIdx.y;
    if (mask && !mask[p]) return;
    __shared__ double2 tile[16 * MT * (GKT + 1)];
    const int mt = (m + 15) >> 4, mp = mt * 16;
    const int npair = mt * (mt + 1) / 2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = lane & 15, kq = lane >> 4;
    int pj[QM], pl[QM];
#pragma unroll
    for (int q = 0; q < QM; ++q) {
        int t = w + 4 * q, j = 0;
        pj[q] = -1; pl[q] = -1;
        if (t < npair) {
            while (t >= mt - j) { t -= mt - j; ++j; }
            pj[q] = j; pl[q] = j + t;
        }
    }
    gacc_t re[QM], im[QM];
#pragma unroll
    for (int q = 0; q < QM; ++q) { re[q] = (gacc_t){0, 0, 0, 0}; im[q] = (gacc_t){0, 0, 0, 0}; }
    const double2* Vp = V + (size_t)p * pstride;
    const size_t chunk = (size_t)gridDim.x * GKT;
    for (size_t base = (size_t)blockIdx.x * GKT; base < n; base += chunk) {
        __syncthreads();
        for (int t = threadIdx.x; t < mp * GKT; t += RB) {
            const int j = t / GKT, r = t - j * GKT;
            const size_t i = base + r;
            tile[j * (GKT + 1) + r] = (j < m && i < n) ? Vp[(size_t)j * cstride + i] : make_double2(0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < QM; ++q) {
            if (pj[q] < 0) continue;
            const double2* A = tile + (pj[q] * 16 + row) * (GKT + 1) + kq;
            const double2* B = tile + (pl[q] * 16 + row) * (GKT + 1) + kq;
#pragma unroll
            for (int k0 = 0; k0 < GKT; k0 += 4) {
                const double2 a = A[k0], bb = B[k0];
                re[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bb.x, re[q], 0, 0, 0);
                re[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bb.y, re[q], 0, 0, 0);
                im[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bb.y, im[q], 0, 0, 0);
                im[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, bb.x, im[q], 0, 0, 0);
            }
        }
    }
    double* o = partials + ((size_t)p * gridDim.x + blockIdx.x) * 2 * m * m;
    for (int t = threadIdx.x; t < 2 * m * m; t += RB) o[t] = 0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QM; ++q) {
        if (pj[q] < 0) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = pj[q] * 16 + kq + 4 * v, l = pl[q] * 16 + row;
            if (j < m && l < m && j <= l) {
                o[2 * (j * m + l)] = re[q][v];
                o[2 * (j * m + l) + 1] = im[q][v];
            }
        }
    }
}

/* W[l][i] = sum_j V[j][i] Q[j][l], Q real m x q (column-major, leading dim m).
 * Output columns are produced 16 at a time; the 16 columns of Q they need
 * (16 m doubles, at most 16 KB for m <= 128) are staged in LDS per group.    */
__global__ void __launch_bounds__(RB) k_gemm_vq(const double2* V, const double* Q, double2* W, int m, int q, size_t n,
                                                size_t cstride, size_t pstride, size_t wcstride, size_t wpstride,
                                                size_t qstride, const int* mask) {
    const int p = blockIdx.y;
    if (mask && !mask[p]) return;
    extern __shared__ double qs[];
    const double2* Vp = V + (size_t)p * pstride;
    double2* Wp = W + (size_t)p * wpstride;
    const double* Qp = Q + (size_t)p * qstride;
    for (int l0 = 0; l0 < q; l0 += 16) {
        const int nl = min(16, q - l0);
        __syncthreads();
        for (int t = threadIdx.x; t < m * nl; t += RB) qs[t] = Qp[(size_t)l0 * m + t];
        __syncthreads();
        for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += (size_t)gridDim.x * RB) {
            double2 acc[16];
#pragma unroll
            for (int l = 0; l < 16; ++l) acc[l] = make_double2(0, 0);
            for (int j = 0; j < m; ++j) {
                double2 v = Vp[(size_t)j * cstride + i];
#pragma unroll
                for (int l = 0; l < 16; ++l) {
                    if (l < nl) {
                        double qq = qs[l * m + j];
                        acc[l].x += v.x * qq; acc[l].y += v.y * qq;
                    }
                }
            }
#pragma unroll
            for (int l = 0; l < 16; ++l)
                if (l < nl) Wp[(size_t)(l0 + l) * wcstride + i] = acc[l];
        }
    }
}

/* ================================================================ host ==== */
static inline dim3 vgrid(size_t n, int nprob, int nblk) { return dim3(nblk, nprob); }

hipError_t fame_cg_begin(void* x, void* r, void* p, const void* b, double* partials, int nblk, size_t n, size_t stride,
                         int nprob, int prec, const int* mask, hipStream_t s) {
    if (prec == 8)
        hipLaunchKernelGGL(k_cg_begin<double>, vgrid(n, nprob, nblk), dim3(RB), 0, s, (double2*)x, (double2*)r, (double2*)p,
                           (const double2*)b, partials, n, stride, mask);
    else
        hipLaunchKernelGGL(k_cg_begin<float>, vgrid(n, nprob, nblk), dim3(RB), 0, s, (float2*)x, (float2*)r, (float2*)p,
                           (const float2*)b, partials, n, stride, mask);
    return hipGetLastError();
}
hipError_t fame_cg_init_scal(CgScal* sc, const double* partials, int nblk, double tol, int maxit, int nprob, int* active,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_cg_init_scal, dim3(nprob), dim3(RB), 0, s, sc, partials, nblk, tol, maxit, active, (const int*)nullptr);
    return hipGetLastError();
}
hipError_t fame_cg_alpha(CgScal* sc, const double* pap, int npap, int nprob, const int* active, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_alpha, dim3(nprob), dim3(RB), 0, s, sc, pap, npap, active);
    return hipGetLastError();
}
hipError_t fame_cg_beta(CgScal* sc, const double* partials, int nblk, int nprob, int* active, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_beta, dim3(nprob), dim3(RB), 0, s, sc, partials, nblk, active);
    return hipGetLastError();
}
hipError_t fame_cg_alpha_beta(CgScal* sc, const double* partials, int nblk, int nprob, int* active, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_alpha_beta, dim3(nprob), dim3(RB), 0, s, sc, partials, nblk, active);
    return hipGetLastError();
}
hipError_t fame_cg_finish_x(void* x, const void* p, const CgScal* sc, size_t n, size_t stride, int nprob, int prec,
                            hipStream_t s) {
    const int nblk = 1024;
    if (prec == 8)
        hipLaunchKernelGGL(k_cg_finish_x<double>, vgrid(n, nprob, nblk), dim3(RB), 0, s, (double2*)x, (const double2*)p,
                           sc, n, stride);
    else
        hipLaunchKernelGGL(k_cg_finish_x<float>, vgrid(n, nprob, nblk), dim3(RB), 0, s, (float2*)x, (const float2*)p, sc,
                           n, stride);
    return hipGetLastError();
}

hipError_t fame_cg_begin_mixed(void* x32, void* r32, void* p32, const void* b64, double* partials, int nblk, size_t n,
                               size_t stride, int nprob, const int* mask, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_begin_from64, vgrid(n, nprob, nblk), dim3(RB), 0, s, (float2*)x32, (float2*)r32, (float2*)p32,
                       (const double2*)b64, partials, n, stride, mask);
    return hipGetLastError();
}

hipError_t fame_scale_real_mod2(void* y, size_t ystride, const void* x, size_t xstride, const double* sreal, size_t n,
                                size_t sstride, size_t smod, int nprob, int prec, hipStream_t s, const int* mask) {
    dim3 g(1024, nprob);
    if (prec == 8)
        hipLaunchKernelGGL(k_scale_real<double>, g, dim3(RB), 0, s, (double2*)y, ystride, (const double2*)x, xstride, sreal,
                           n, sstride, smod, mask);
    else
        hipLaunchKernelGGL(k_scale_real<float>, g, dim3(RB), 0, s, (float2*)y, ystride, (const float2*)x, xstride, sreal,
                           n, sstride, smod, mask);
    return hipGetLastError();
}
hipError_t fame_scale_real_mod(void* y, const void* x, const double* sreal, size_t n, size_t xstride, size_t sstride,
                               size_t smod, int nprob, int prec, hipStream_t s, const int* mask) {
    return fame_scale_real_mod2(y, xstride, x, xstride, sreal, n, sstride, smod, nprob, prec, s, mask);
}
hipError_t fame_scale_real(void* y, const void* x, const double* sreal, size_t n, size_t xstride, size_t sstride,
                           int nprob, int prec, hipStream_t s) {
    return fame_scale_real_mod(y, x, sreal, n, xstride, sstride, n, nprob, prec, s);
}
hipError_t fame_convert(void* dst, int dprec, const void* src, int sprec, size_t n, size_t dstride, size_t sstride,
                        int nprob, hipStream_t s, const int* mask) {
    dim3 g(1024, nprob);
    if (dprec == 4 && sprec == 8)
        hipLaunchKernelGGL((k_convert<float, double>), g, dim3(RB), 0, s, (float2*)dst, (const double2*)src, n, dstride, sstride, mask);
    else if (dprec == 8 && sprec == 4)
        hipLaunchKernelGGL((k_convert<double, float>), g, dim3(RB), 0, s, (double2*)dst, (const float2*)src, n, dstride, sstride, mask);
    else if (dprec == 8)
        hipLaunchKernelGGL((k_convert<double, double>), g, dim3(RB), 0, s, (double2*)dst, (const double2*)src, n, dstride, sstride, mask);
    else
        hipLaunchKernelGGL((k_convert<float, float>), g, dim3(RB), 0, s, (float2*)dst, (const float2*)src, n, dstride, sstride, mask);
    return hipGetLastError();
}
hipError_t fame_norm2_partials(const void* x, int prec, double* partials, int nblk, size_t n, size_t stride, int nprob,
                               hipStream_t s) {
    if (prec == 8)
        hipLaunchKernelGGL(k_norm2<double>, dim3(nblk, nprob), dim3(RB), 0, s, (const double2*)x, partials, n, stride);
    else
        hipLaunchKernelGGL(k_norm2<float>, dim3(nblk, nprob), dim3(RB), 0, s, (const float2*)x, partials, n, stride);
    return hipGetLastError();
}
hipError_t fame_reduce_partials(const double* partials, int nblk, int ncomp, double* out, int nprob, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(nprob, ncomp), dim3(RB), 0, s, partials, nblk, ncomp, out);
    return hipGetLastError();
}
hipError_t fame_gemv_h(const void* V, const void* w, double* partials, int nblk, int m, size_t n, size_t cstride,
                       size_t pstride, size_t wstride, int nprob, const int* mask, hipStream_t s) {
    if (m <= 4)
        hipLaunchKernelGGL(k_gemv_h<4>, dim3(nblk, nprob), dim3(RB), 0, s, (const double2*)V, (const double2*)w, partials,
                           m, n, cstride, pstride, wstride, mask);
    else
        hipLaunchKernelGGL(k_gemv_h<8>, dim3(nblk, nprob), dim3(RB), 0, s, (const double2*)V, (const double2*)w, partials,
                           m, n, cstride, pstride, wstride, mask);
    return hipGetLastError();
}
hipError_t fame_gemv_n(const void* V, void* w, const double* h, int m, size_t n, size_t cstride, size_t pstride,
                       size_t wstride, size_t hstride, int nprob, const int* mask, hipStream_t s) {
    hipLaunchKernelGGL(k_gemv_n, dim3(1024, nprob), dim3(RB), 0, s, (const double2*)V, (double2*)w, h, m, n, cstride,
                       pstride, wstride, hstride, mask);
    return hipGetLastError();
}
hipError_t fame_gram(const void* V, double* partials, int nblk, int m, size_t n, size_t cstride, size_t pstride, int nprob,
                     const int* mask, hipStream_t s) {
    if (m < 1 || m > FAME_LZ_MAXCOL) return hipErrorInvalidValue;
    if (m <= 64)
        hipLaunchKernelGGL((k_gram<32, 4, 3>), dim3(nblk, nprob), dim3(RB), 0, s, (const double2*)V, partials, m, n, cstride,
                           pstride, mask);
    else
        hipLaunchKernelGGL((k_gram<16, 8, 9>), dim3(nblk, nprob), dim3(RB), 0, s, (const double2*)V, partials, m, n, cstride,
                           pstride, mask);
    return hipGetLastError();
}
hipError_t fame_gemm_vq(const void* V, const double* Q, void* W, int m, int q, size_t n, size_t cstride, size_t pstride,
                        size_t wcstride, size_t wpstride, size_t qstride, int nprob, const int* mask, hipStream_t s) {
    if (m < 1 || m > FAME_LZ_MAXCOL || q < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gemm_vq, dim3(1024, nprob), dim3(RB), sizeof(double) * m * 16, s, (const double2*)V, Q, (double2*)W, m, q, n, cstride,
                       pstride, wcstride, wpstride, qstride, mask);
    return hipGetLastError();
}

/* Lanczos three-term coefficients (Lanczos_decomp_gpu.cu:161-169): alpha_j is
 * the real part of <U_j, U_{j+1}>; h receives the column weights that
 * gemv_n subtracts: [beta_{j-1}, alpha_j] for columns (j-1, j), or [alpha_0]. */
__global__ void k_lz_coeffs(const double* hdot, double* T0d, const double* T1d, int ld, int j, double* hout,
                            const int* mask) {
    const int p = blockIdx.x;
    if (threadIdx.x != 0 || (mask && !mask[p])) return;
    const double a = hdot[2 * p];
    T0d[(size_t)p * ld + j] = a;
    double* h = hout + 4 * (size_t)p;
    if (j >= 1) { h[0] = T1d[(size_t)p * ld + j - 1]; h[1] = 0; h[2] = a; h[3] = 0; }
    else { h[0] = a; h[1] = 0; h[2] = 0; h[3] = 0; }
}

/* out[p] = in[p] && sum_i h[p][i]^2 > tau2 (h: n doubles per problem) */
__global__ void k_lz_second_pass_mask(const double* h, int n, double tau2, const int* in, int* out) {
    const int p = blockIdx.x;
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) { const double x = h[(size_t)p * n + i]; v[0] += x * x; }
    block_sum_all<1>(v);
    if (threadIdx.x == 0) out[p] = (!in || in[p]) && v[0] > tau2;
}

/* beta_j = ||U_{j+1}|| and its inverse (Lanczos_decomp_gpu.cu:171-174) */
__global__ void k_lz_beta(const double* nrm2, double* T1d, int ld, int j, double* invb, const int* mask) {
    const int p = blockIdx.x;
    if (threadIdx.x != 0 || (mask && !mask[p])) return;
    const double b = sqrt(nrm2[p]);
    T1d[(size_t)p * ld + j] = b;
    invb[p] = 1.0 / b;
}

/* CG start scalars with an absolute tolerance and an optional relative one:
 * stop when <r,r> <= max(tol^2, reltol^2 <b,b>)                              */
__global__ void k_cg_init_scal2(CgScal* sc, const double* partials, int nblk, double tol, double reltol, int maxit,
                                int* active, const int* mask) {
    const int p = blockIdx.x;
    const double s = sum_partials(partials, p, nblk, 1, 0);
    if (threadIdx.x != 0) return;
    CgScal c;
    c.rr = s; c.rr0 = s; c.pap_re = c.pap_im = 0; c.alpha_re = c.alpha_im = 0; c.beta = 0;
    c.tol2 = fmax(tol * tol, reltol * reltol * s); c.iters = 1; c.maxit = maxit;
    c.active = (!mask || mask[p]) && (s > c.tol2) && (1 <= maxit);
    c.pend = 0;
    sc[p] = c;
    active[p] = c.active;
}

hipError_t fame_lz_coeffs(const double* hdot, double* T0d, const double* T1d, int ld, int j, double* hout, int nprob,
                          const int* mask, hipStream_t s) {
    hipLaunchKernelGGL(k_lz_coeffs, dim3(nprob), dim3(64), 0, s, hdot, T0d, T1d, ld, j, hout, mask);
    return hipGetLastError();
}
hipError_t fame_lz_second_pass_mask(const double* h, int n, double tau2, const int* in, int* out, int nprob,
                                    hipStream_t s) {
    hipLaunchKernelGGL(k_lz_second_pass_mask, dim3(nprob), dim3(64), 0, s, h, n, tau2, in, out);
    return hipGetLastError();
}
hipError_t fame_lz_beta(const double* nrm2, double* T1d, int ld, int j, double* invb, int nprob, const int* mask,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_lz_beta, dim3(nprob), dim3(64), 0, s, nrm2, T1d, ld, j, invb, mask);
    return hipGetLastError();
}
hipError_t fame_cg_init_scal2(CgScal* sc, const double* partials, int nblk, double tol, double reltol, int maxit,
                              int nprob, int* active, const int* mask, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_init_scal2, dim3(nprob), dim3(RB), 0, s, sc, partials, nblk, tol, reltol, maxit, active, mask);
    return hipGetLastError();
}
