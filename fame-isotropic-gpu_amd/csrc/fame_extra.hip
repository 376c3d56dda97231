/*
 * fame_extra.hip -- residual check and Hellmann-Feynman adjoint.
 *
 * fame_residuals: ||C^H C E - omega^2 B E|| / ||E|| on the GPU through the Fourier
 *   factorisation C_i = T Lambda_i T* (the identity Check_Eigendecomp asserts,
 *   FAME_Main_Code.cu:250-362): C^H C E = T( -conj(Lambda) x (Lambda x T* E) ).
 *   It replaces the host sparse-matrix Check_Residual (FAME_Main_Code.cu:383-421)
 *   and reports the residual relative to ||E||.
 * fame_eig_grad: d(sum_kj w_kj lambda_kj)/d eps_e with
 *   d lambda / d eps_e = -lambda |E_e|^2 / (E^H B E) for A E = lambda B E,
 *   A = C^H C independent of eps, B = diag(eps).  Within a degenerate cluster
 *   the weights are replaced by their mean, i.e. the derivative of the cluster
 *   trace, which is the only basis-independent quantity there.
 */
#include "fame_hip.h"
#include "fame_host.h"
#include "fame_kernels.h"

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

using namespace fame;

std::string& fame_err_ref();
#define FAIL(code, msg) do { fame_err_ref() = (msg); return (code); } while (0)
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    fame_err_ref() = std::string("HIP error: ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" + std::to_string(__LINE__); \
    return FAME_ERR_HIP; } } while (0)
#define CHK(x) do { fame_status s_ = (x); if (s_ != FAME_OK) return s_; } while (0)

/* inner pass of the curl-curl operator: F (after T* stages c and b) ->
 * conj(twid).F -> FFT_a -> -conj(L) x (L x .) -> IFFT_a -> twid . -> F
 * (one stage list: forward transform, the curl-curl step, inverse transform;
 * with Bluestein the forward post-chirp and inverse pre-chirp cancel and only
 * 1/M remains in the step) */
enum { STEP_CURLCURL = 0, STEP_LAMBDA = 1 };
template <int ENG, int STEP>
__global__ void __launch_bounds__(FFT_THREADS) k_curlcurl(KInner args) {
    typedef double2 T2;
    const int p = blockIdx.y;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const GridDev& g = args.g;
    const ProbDev pr = args.prob[p];
    const LinePlanDev& pl = args.plan;
    const int na = g.na, L = args.L, P = args.P;
    T2* base = reinterpret_cast<T2*>(smem_raw);
    const Tabs<T2> tb = load_tabs<T2>(pl, base);
    T2* S = base + lds_head(pl);
    const int nlines = g.nb * g.nc;
    const int l0 = blockIdx.x * L;
    const size_t N = (size_t)g.N;
    T2* F = reinterpret_cast<T2*>(args.f) + (size_t)p * args.fstride;
    const int Wd = pl.blue ? pl.M : na;
    const unsigned mW = pl.blue ? pl.mM : g.m_na;
    for (int idx = threadIdx.x; idx < L * Wd; idx += FFT_THREADS) {
        const int line = udiv(idx, mW), a = idx - line * Wd, gl = l0 + line;
        T2 v[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
        if (gl < nlines && a < na) {
            const int bf = gl % g.nb, cf = gl / g.nb;
            const T2 w = expipi(2.0 * (pr.t0 + pr.t1 * (double)bf + pr.t2 * (double)cf) * (double)a / (double)na);
            for (int c = 0; c < 3; ++c) v[c] = cmulc(F[c * N + (size_t)gl * na + a], w);
        }
        for (int c = 0; c < 3; ++c) S[(c * L + line) * P + a] = v[c];
    }
    __syncthreads();
    const double sc = pl.blue ? 1.0 / (double)pl.M : 1.0;
    auto curlcurl = [&](T2* Sx) {
        for (int idx = threadIdx.x; idx < L * Wd; idx += FFT_THREADS) {
            const int line = udiv(idx, mW), a = idx - line * Wd, gl = l0 + line;
            double2 r[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
            if (gl < nlines && a < na) {
                const int bf = gl % g.nb, cf = gl / g.nb;
                double2 l[3];
                lambdas(g, pr, a, bf, cf, l[0], l[1], l[2]);
                double2 e[3];
                for (int c = 0; c < 3; ++c) e[c] = cscale(Sx[(c * L + line) * P + a], sc);
                if (STEP == STEP_LAMBDA) {   /* component c times Lambda_c */
                    for (int c = 0; c < 3; ++c) Sx[(c * L + line) * P + a] = cmul(l[c], e[c]);
                    continue;
                }
                double2 u[3];   /* u = L x e */
                u[0] = csub(cmul(l[1], e[2]), cmul(l[2], e[1]));
                u[1] = csub(cmul(l[2], e[0]), cmul(l[0], e[2]));
                u[2] = csub(cmul(l[0], e[1]), cmul(l[1], e[0]));
                double2 m[3] = {make_double2(-l[0].x, l[0].y), make_double2(-l[1].x, l[1].y), make_double2(-l[2].x, l[2].y)};
                /* r = (-conj L) x u */
                r[0] = csub(cmul(m[1], u[2]), cmul(m[2], u[1]));
                r[1] = csub(cmul(m[2], u[0]), cmul(m[0], u[2]));
                r[2] = csub(cmul(m[0], u[1]), cmul(m[1], u[0]));
            }
            for (int c = 0; c < 3; ++c) Sx[(c * L + line) * P + a] = r[c];
        }
        __syncthreads();
    };
    run_stages<T2, MaxE<double>::v, ENG>(S, 3 * L, P, pl, tb, curlcurl);
    for (int idx = threadIdx.x; idx < L * na; idx += FFT_THREADS) {
        const int line = udiv(idx, g.m_na), a = idx - line * na, gl = l0 + line;
        if (gl >= nlines) continue;
        const int bf = gl % g.nb, cf = gl / g.nb;
        const T2 w = expipi(2.0 * (pr.t0 + pr.t1 * (double)bf + pr.t2 * (double)cf) * (double)a / (double)na);
        for (int c = 0; c < 3; ++c)
            F[c * N + (size_t)gl * na + a] = cmul(post_mul<T2, double>(pl, tb, a, S[(c * L + line) * P + a]), w);
    }
}

/* partials of ||F - w2 eps E||^2 and ||E||^2 (kernel layout, one problem per blockIdx.y) */
__global__ void __launch_bounds__(256) k_resid(const double2* F, const double2* E, const double* eps, const double* w2,
                                               double* partials, size_t n3, size_t stride) {
    const int p = blockIdx.y;
    double a = 0, b = 0;
    const double om2 = w2[p];
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n3; i += (size_t)gridDim.x * 256) {
        double2 f = F[p * stride + i], e = E[p * stride + i];
        double s = om2 * eps[i];
        double rx = f.x - s * e.x, ry = f.y - s * e.y;
        a += rx * rx + ry * ry;
        b += e.x * e.x + e.y * e.y;
    }
    __shared__ double ra[256], rb[256];
    ra[threadIdx.x] = a; rb[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { ra[threadIdx.x] += ra[threadIdx.x + s]; rb[threadIdx.x] += rb[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[((size_t)p * gridDim.x + blockIdx.x) * 2] = ra[0];
        partials[((size_t)p * gridDim.x + blockIdx.x) * 2 + 1] = rb[0];
    }
}

/* partials of sum eps |E|^2 for one field (reference layout) */
__global__ void __launch_bounds__(256) k_wnorm(const double2* E, const double* eps, double* partials, size_t n3) {
    double a = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n3; i += (size_t)gridDim.x * 256) {
        double2 e = E[i];
        a += eps[i] * (e.x * e.x + e.y * e.y);
    }
    __shared__ double ra[256];
    ra[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) ra[threadIdx.x] += ra[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = ra[0];
}

/* g += coef * |E|^2 */
__global__ void k_grad_acc(double* g, const double2* E, double coef, size_t n3) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n3; i += (size_t)gridDim.x * 256) {
        double2 e = E[i];
        g[i] += coef * (e.x * e.x + e.y * e.y);
    }
}

/* Relative residuals of the bands of wave vectors p0..p0+np-1, through the
 * FFT curl-curl pipeline on the context's field buffer, with Ecopy (np x 3N
 * complex, kernel layout) holding the fields themselves for the residual sums. */
static fame_status residuals_chunk(fame_ctx* c, int nwant, const double* freq, const void* E_dev, double* res,
                                   double2* Ecopy, int p0, int np) {
    const size_t N3 = 3 * (size_t)c->N;
    const int NB = 256;
    std::vector<double> w2(np), hp((size_t)np * NB * 2);
    double* w2d = c->scal;
    for (int i = 0; i < nwant; ++i) {
        for (int q = 0; q < np; ++q) {
            const int p = p0 + q;
            const double2* src = (const double2*)E_dev + ((size_t)p * nwant + i) * N3;
            CHK(c->field_from_ref(src, c->F64 + (size_t)p * N3, 1));
            CHK(c->field_from_ref(src, Ecopy + (size_t)q * N3, 1));
            w2[q] = freq[(size_t)p * nwant + i] * freq[(size_t)p * nwant + i];
        }
        KInner ki = c->kin(8, p0, nullptr, LIST_FI);
        HIPCHK(hipMemcpyAsync(w2d, w2.data(), sizeof(double) * np, hipMemcpyHostToDevice, c->stream));
        HIPCHK(fame_launch_outer(c->kout(8, p0, nullptr, LIST_FWD), OUTER_QRS, 8, np, c->stream));
        HIPCHK(fame_launch_mid(c->kmid(8, p0, nullptr, LIST_FWD), -1, 8, np, c->stream));
        dim3 grid((c->g.nb * c->g.nc + ki.L - 1) / ki.L, np);
        const size_t lds = fame_line_lds(ki.plan, 8, 3 * ki.L, ki.P, 0);
        if (plan_engine(ki.plan) == ENG_MFMA)
            hipLaunchKernelGGL((k_curlcurl<ENG_MFMA, STEP_CURLCURL>), grid, dim3(FFT_THREADS), lds, c->stream, ki);
        else hipLaunchKernelGGL((k_curlcurl<ENG_STOCKHAM, STEP_CURLCURL>), grid, dim3(FFT_THREADS), lds, c->stream, ki);
        HIPCHK(hipGetLastError());
        HIPCHK(fame_launch_mid(c->kmid(8, p0, nullptr, LIST_INV), +1, 8, np, c->stream));
        HIPCHK(fame_launch_outer(c->kout(8, p0, nullptr, LIST_INV), OUTER_QR, 8, np, c->stream));
        hipLaunchKernelGGL(k_resid, dim3(NB, np), dim3(256), 0, c->stream, c->F64 + (size_t)p0 * N3, Ecopy, c->eps_int,
                           w2d, c->partials, N3, N3);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hp.data(), c->partials, sizeof(double) * np * NB * 2, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int q = 0; q < np; ++q) {
            double a = 0, b = 0;
            for (int t = 0; t < NB; ++t) { a += hp[((size_t)q * NB + t) * 2]; b += hp[((size_t)q * NB + t) * 2 + 1]; }
            res[(size_t)(p0 + q) * nwant + i] = b > 0 ? std::sqrt(a / b) : 0.0;
        }
    }
    return FAME_OK;
}

extern "C" fame_status fame_residuals(fame_ctx* c, int nwant, const double* freq, const void* E_dev, double* res) {
    if (!c || !freq || !E_dev || !res || nwant < 1) FAIL(FAME_ERR_ARG, "fame_residuals: bad arguments");
    if (!c->have_eps) FAIL(FAME_ERR_STATE, "fame_residuals: permittivity not set");
    HIPCHK(hipSetDevice(c->device));
    const int nk = c->nk;
    const size_t N3 = 3 * (size_t)c->N;
    double2* Ecopy = nullptr;
    HIPCHK(hipMalloc(&Ecopy, 16 * N3 * nk));
    const fame_status st = residuals_chunk(c, nwant, freq, E_dev, res, Ecopy, 0, nk);
    (void)hipFree(Ecopy);
    return st;
}

/* Scratch of fame_residuals_noalloc: the fields of as many wave vectors as fit,
 * in the kernel layout.  A file-scope device array, so that the entry point
 * allocates nothing; every device the library runs on holds it (128 MiB) until
 * the process ends.  Host code never passes the array's name to a kernel: it
 * uses the device address that hipGetSymbolAddress returns, once per device. */
#define FAME_NOALLOC_ELEMS (8u << 20)    /* complex doubles: 128 MiB */
static __device__ __align__(256) double2 scratch[FAME_NOALLOC_ELEMS];

static std::mutex g_scratch_mu;
static double2* g_scratch_dev[64] = {nullptr};

/* the current device's address of `scratch` (the current device must be dev) */
static fame_status scratch_address(int dev, double2** out) {
    if (dev < 0 || dev >= 64) FAIL(FAME_ERR_ARG, "fame_residuals_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (double2*)ptr;
    }
    *out = g_scratch_dev[dev];
    return FAME_OK;
}

extern "C" fame_status fame_residuals_noalloc(fame_ctx* c, int nwant, const double* freq, const void* E_dev,
                                              double* res) {
    if (!c || !freq || !E_dev || !res || nwant < 1) FAIL(FAME_ERR_ARG, "fame_residuals_noalloc: bad arguments");
    if (!c->have_eps) FAIL(FAME_ERR_STATE, "fame_residuals_noalloc: permittivity not set");
    const size_t N3 = 3 * (size_t)c->N;
    if (N3 > FAME_NOALLOC_ELEMS)
        FAIL(FAME_ERR_UNSUPPORTED, "fame_residuals_noalloc: 3N = " + std::to_string(N3) + " exceeds the static scratch of " +
                                       std::to_string(FAME_NOALLOC_ELEMS) + " complex values (N <= " +
                                       std::to_string(FAME_NOALLOC_ELEMS / 3) + ")");
    HIPCHK(hipSetDevice(c->device));
    double2* sc = nullptr;
    CHK(scratch_address(c->device, &sc));
    const int chunk = (int)std::min<size_t>((size_t)c->nk, FAME_NOALLOC_ELEMS / N3);
    for (int p0 = 0; p0 < c->nk; p0 += chunk)
        CHK(residuals_chunk(c, nwant, freq, E_dev, res, sc, p0, std::min(chunk, c->nk - p0)));
    return FAME_OK;
}

/* Average duration (microseconds, hipEvent timing on the context stream) of each
 * of the five QBQ passes and of a whole QBQ, for all nk problems of the context
 * in one launch, in precision prec (8 or 4).  Used by the benchmark to price the
 * dominant kernel against the HBM roofline.  us6 = {P1, P2, P3, P4, P5, QBQ}.  */
extern "C" fame_status fame_time_qbq_passes(fame_ctx* c, int prec, int reps, double* us6) {
    if (!c || !us6 || reps < 1 || (prec != 8 && prec != 4)) FAIL(FAME_ERR_ARG, "fame_time_qbq_passes: bad arguments");
    if (!c->have_eps || c->nk < 1) FAIL(FAME_ERR_STATE, "fame_time_qbq_passes: permittivity and wave vectors must be set");
    HIPCHK(hipSetDevice(c->device));
    const int nk = c->nk;
    void* x = prec == 8 ? c->cg64[1] : c->cg32[1];
    void* y = prec == 8 ? c->cg64[2] : c->cg32[2];
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    KInner ki = c->kin(prec, 0, nullptr, LIST_INV), kf = c->kin(prec, 0, nullptr, LIST_FWD);
    const KAxis kmi = c->kmid(prec, 0, nullptr, LIST_INV), kmf = c->kmid(prec, 0, nullptr, LIST_FWD);
    const KAxis ko = c->kout(prec, 0, nullptr, LIST_EPS);
    ki.x = x; ki.pikind = 0;
    kf.y = y; kf.pikind = 0;
    for (int which = 0; which < 6; ++which) {
        for (int rep = -1; rep < reps; ++rep) {
            if (rep == 0) HIPCHK(hipEventRecord(e0, c->stream));
            switch (which) {
                case 0: HIPCHK(fame_launch_inner(ki, 1, prec, nk, c->stream)); break;
                case 1: HIPCHK(fame_launch_mid(kmi, +1, prec, nk, c->stream)); break;
                case 2: HIPCHK(fame_launch_outer(ko, OUTER_EPS, prec, nk, c->stream)); break;
                case 3: HIPCHK(fame_launch_mid(kmf, -1, prec, nk, c->stream)); break;
                case 4: HIPCHK(fame_launch_inner(kf, 0, prec, nk, c->stream)); break;
                default: CHK(c->qbq(prec, x, y, 0, nk, nullptr, nullptr)); break;
            }
        }
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        us6[which] = 1000.0 * ms / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return FAME_OK;
}

/* Average durations (microseconds, hipEvent timing on the context stream) of
 * the kernels of one fused CG iteration as the solver launches them
 * (fame_ctx::cg_step_rc), with every problem of the context active:
 *   us8 = {inner expand + x, p updates, middle inverse, outer eps + <p,Ap> sums,
 *          alpha, middle forward, inner contract + r update and <r,r> sums,
 *          beta and flags, whole iteration}.
 * The CG state starts from x = 0, r = p = b for a constant nonzero b, with a
 * zero tolerance, so that no problem stops during the timing, and one untimed
 * iteration leaves the solution update pending as inside a solve.  Returns
 * FAME_ERR_UNSUPPORTED when an axis plan is Bluestein (the iteration then
 * stores Ap; time it with fame_time_qbq_passes).  Measurement helper for the
 * benchmark, no reference counterpart.                                       */
extern "C" fame_status fame_time_cg_passes(fame_ctx* c, int prec, int reps, double* us8) {
    if (!c || !us8 || reps < 1 || (prec != 8 && prec != 4)) FAIL(FAME_ERR_ARG, "fame_time_cg_passes: bad arguments");
    if (!c->have_eps || c->nk < 1) FAIL(FAME_ERR_STATE, "fame_time_cg_passes: permittivity and wave vectors must be set");
    const int npap = c->pap_blocks(prec);
    if (!npap) FAIL(FAME_ERR_UNSUPPORTED, "fame_time_cg_passes: an axis plan is Bluestein");
    HIPCHK(hipSetDevice(c->device));
    const int nk = c->nk;
    const size_t st = 2 * (size_t)c->N, es = prec == 8 ? 16 : 8;
    void** V = prec == 8 ? c->cg64 : c->cg32;
    void* r = V[0];
    void* pv = V[1];
    void* b = V[3];
    void* x = prec == 8 ? c->ref64[1] : c->cg32[4];
    /* b: every byte 0x3c (1.15e-2 in FP32, 1.6e-19 in FP64), nonzero and finite */
    HIPCHK(hipMemsetAsync(b, 0x3c, es * st * nk, c->stream));
    KInner ki = c->kin(prec, 0, c->active, LIST_INV);
    ki.x = pv; ki.pikind = 0;
    ki.cg_r = r; ki.cg_x = x; ki.cgsc = c->cgsc; ki.cg_ap = nullptr; ki.partials = c->partials2;
    const KAxis kmi = c->kmid(prec, 0, c->active, LIST_INV), kmf = c->kmid(prec, 0, c->active, LIST_FWD);
    KAxis ko = c->kout(prec, 0, c->active, LIST_EPS);
    ko.pap = c->pap_part;
    KInner kc = c->kin(prec, 0, c->active, LIST_FWD);
    kc.x = pv; kc.y = r; kc.pikind = 0;
    kc.cg_r = r; kc.cgsc = c->cgsc; kc.partials = c->partials2;
    const int nbi = c->nblk_in[prec == 8 ? 0 : 1];
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const long long c64 = c->cnt_qbq64, c32 = c->cnt_qbq32;
    for (int which = 0; which < 8; ++which) {
        /* a fresh state for every kernel (repeating one kernel alone drifts the
         * vectors, and a non-finite <r,r> would clear the active flags), and
         * one iteration: alpha leaves the x update pending, so that the first
         * pass reads and writes x as it does inside a solve */
        HIPCHK(fame_cg_begin(x, r, pv, b, c->partials, NBLK_VEC, st, st, nk, prec, nullptr, c->stream));
        HIPCHK(fame_cg_init_scal2(c->cgsc, c->partials, NBLK_VEC, 0.0, 0.0, 1 << 30, nk, c->active, nullptr, c->stream));
        CHK(c->cg_step_rc(prec, pv, r, x, 0, nk, npap));
        for (int rep = -1; rep < reps; ++rep) {
            if (rep == 0) HIPCHK(hipEventRecord(e0, c->stream));
            switch (which) {
                case 0: HIPCHK(fame_launch_inner(ki, 1, prec, nk, c->stream)); break;
                case 1: HIPCHK(fame_launch_mid(kmi, +1, prec, nk, c->stream)); break;
                case 2: HIPCHK(fame_launch_outer(ko, OUTER_EPS, prec, nk, c->stream)); break;
                case 3: HIPCHK(fame_cg_alpha(c->cgsc, c->pap_part, npap, nk, c->active, c->stream)); break;
                case 4: HIPCHK(fame_launch_mid(kmf, -1, prec, nk, c->stream)); break;
                case 5: HIPCHK(fame_launch_inner(kc, 0, prec, nk, c->stream)); break;
                case 6: HIPCHK(fame_cg_beta(c->cgsc, c->partials2, nbi, nk, c->active, c->stream)); break;
                default: CHK(c->cg_step_rc(prec, pv, r, x, 0, nk, npap)); break;
            }
        }
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        us8[which] = 1000.0 * ms / reps;
    }
    c->cnt_qbq64 = c64; c->cnt_qbq32 = c32;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return FAME_OK;
}

extern "C" fame_status fame_eig_grad(fame_ctx* c, int nwant, const double* freq, const void* E_dev, const double* weights,
                                     void* grad_dev) {
    if (!c || !freq || !E_dev || !weights || !grad_dev || nwant < 1) FAIL(FAME_ERR_ARG, "fame_eig_grad: bad arguments");
    if (!c->have_eps) FAIL(FAME_ERR_STATE, "fame_eig_grad: permittivity not set");
    HIPCHK(hipSetDevice(c->device));
    const int nk = c->nk;
    const size_t N3 = 3 * (size_t)c->N;
    const int NB = 256;
    std::vector<double> hp(NB);
    HIPCHK(hipMemsetAsync(grad_dev, 0, sizeof(double) * N3, c->stream));
    for (int p = 0; p < nk; ++p) {
        const double* f = freq + (size_t)p * nwant;
        /* degenerate clusters: mean weight over bands whose lambda agree to 1e-8 relative */
        std::vector<double> w(weights + (size_t)p * nwant, weights + (size_t)(p + 1) * nwant);
        for (int i = 0; i < nwant;) {
            int j = i + 1;
            const double li = f[i] * f[i];
            while (j < nwant && std::fabs(f[j] * f[j] - li) <= 1e-8 * std::max(li, 1e-300)) ++j;
            double m = 0;
            for (int q = i; q < j; ++q) m += w[q];
            m /= (j - i);
            for (int q = i; q < j; ++q) w[q] = m;
            i = j;
        }
        for (int i = 0; i < nwant; ++i) {
            const double lam = f[i] * f[i];
            if (w[i] == 0.0 || lam == 0.0) continue;
            const double2* E = (const double2*)E_dev + ((size_t)p * nwant + i) * N3;
            hipLaunchKernelGGL(k_wnorm, dim3(NB), dim3(256), 0, c->stream, E, c->eps_ref, c->partials, N3);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hp.data(), c->partials, sizeof(double) * NB, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            double s = 0;
            for (int q = 0; q < NB; ++q) s += hp[q];
            hipLaunchKernelGGL(k_grad_acc, dim3(1024), dim3(256), 0, c->stream, (double*)grad_dev, E, -w[i] * lam / s, N3);
            HIPCHK(hipGetLastError());
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return FAME_OK;
}

/* ------------------------------------------------------------------------- */
/*  Sparse curl and the check of the factorisation                            */
/* ------------------------------------------------------------------------- */
/* The forward differences C_1, C_2, C_3 of FAME_Matrix_Curl
 * (FAME_Main/FAME_Matrix_Curl.cpp:90-443) as a rule per grid point: row (x, y, z)
 * (x fastest) holds -1/h on the diagonal and w/h on one neighbour, w = 1
 * inside the cell and a Bloch factor exp(i theta) across its boundary.
 *   C_1: x + 1; at x = n1 - 1, x = 0 with theta = 2 pi a1.k.
 *   C_2: y + 1; at y = n2 - 1, y = 0, and for General lattices x shifted by m1:
 *        x < m1 -> x + n1 - m1 with theta = 2 pi (a2 + (rho - 1) a1).k, else
 *        x - m1 with 2 pi (a2 + rho a1).k, rho = 1 when the computational a2
 *        points to negative x (:259-280).
 *   C_3: z + 1; at z = n3 - 1, z = 0, and for General lattices (y, x) shifted
 *        by (m3, m4 or m2) with the boundary-point phases t1..t4 (:296-373):
 *        y < m3 -> y + n2 - m3, x < m4 ? x + n1 - m4 (t3) : x - m4 (t4);
 *        else   -> y - m3,      x < m2 ? x + n1 - m2 (t2) : x - m2 (t1).
 * The full curl is [0, -C3, C2; C3, 0, -C1; -C2, C1, 0] (:105-129).        */
struct CurlGeo {
    int n1, n2, n3, general, m1, m2, m3, m4, rho;
    double ih[3];
    double ta[3];   /* 2 pi a_i . k */
    double tt[4];   /* 2 pi t_j . k */
};

/* the boundary shifts keep every neighbour inside the grid */
static bool curl_geo_ok(const fame_lattice& L) {
    const int n1 = L.grid_nums[0], n2 = L.grid_nums[1], n3 = L.grid_nums[2];
    if (n1 < 1 || n2 < 1 || n3 < 1) return false;
    if (!L.general) return true;
    return L.m1 >= 0 && L.m1 <= n1 && L.m2 >= 0 && L.m2 <= n1 && L.m4 >= 0 && L.m4 <= n1 && L.m3 >= 0 && L.m3 <= n2;
}

static CurlGeo curl_geo(const fame_lattice& L, const double* k) {
    CurlGeo G;
    G.n1 = L.grid_nums[0]; G.n2 = L.grid_nums[1]; G.n3 = L.grid_nums[2];
    G.general = L.general ? 1 : 0;
    G.m1 = L.m1; G.m2 = L.m2; G.m3 = L.m3; G.m4 = L.m4;
    G.rho = L.lattice_vec_a[3] < 0 ? 1 : 0;
    const double tp = 2.0 * 3.14159265358979323846;
    for (int i = 0; i < 3; ++i) {
        G.ih[i] = 1.0 / L.mesh_len[i];
        G.ta[i] = tp * (L.lattice_vec_a[3 * i] * k[0] + L.lattice_vec_a[3 * i + 1] * k[1] + L.lattice_vec_a[3 * i + 2] * k[2]);
    }
    const double* t[4] = {L.t1, L.t2, L.t3, L.t4};
    for (int j = 0; j < 4; ++j) G.tt[j] = tp * (t[j][0] * k[0] + t[j][1] * k[1] + t[j][2] * k[2]);
    return G;
}

/* neighbour of (x, y, z) in C_{ax+1} and the angle of its factor (0 inside);
 * `wrap` tells whether the factor is exp(i angle) or 1 */
__host__ __device__ inline int curl_nbr(const CurlGeo& G, int ax, int x, int y, int z, double& ang, bool& wrap) {
    const int n1 = G.n1, n2 = G.n2, n3 = G.n3;
    ang = 0.0;
    wrap = false;
    if (ax == 0) {
        if (x < n1 - 1) return x + 1 + n1 * (y + n2 * z);
        wrap = true; ang = G.ta[0];
        return n1 * (y + n2 * z);
    }
    if (ax == 1) {
        if (y < n2 - 1) return x + n1 * (y + 1 + n2 * z);
        wrap = true;
        if (!G.general) { ang = G.ta[1]; return x + n1 * n2 * z; }
        if (x < G.m1) { ang = G.ta[1] + G.ta[0] * (double)(G.rho - 1); return x + n1 - G.m1 + n1 * n2 * z; }
        ang = G.ta[1] + G.ta[0] * (double)G.rho;
        return x - G.m1 + n1 * n2 * z;
    }
    if (z < n3 - 1) return x + n1 * (y + n2 * (z + 1));
    wrap = true;
    if (!G.general) { ang = G.ta[2]; return x + n1 * y; }
    int xs, ys;
    if (y < G.m3) {
        ys = y + n2 - G.m3;
        if (x < G.m4) { xs = x + n1 - G.m4; ang = G.ta[2] + G.tt[2]; }
        else { xs = x - G.m4; ang = G.ta[2] + G.tt[3]; }
    } else {
        ys = y - G.m3;
        if (x < G.m2) { xs = x + n1 - G.m2; ang = G.ta[2] + G.tt[1]; }
        else { xs = x - G.m2; ang = G.ta[2] + G.tt[0]; }
    }
    return xs + n1 * ys;
}

/* dif = C_i x_i (component i) and cur = C x, reference layout, one point per thread */
__global__ void __launch_bounds__(256) k_curl_stencil(const double2* __restrict__ X, double2* __restrict__ dif,
                                                      double2* __restrict__ cur, CurlGeo G) {
    const size_t N = (size_t)G.n1 * G.n2 * G.n3;
    for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < N; p += (size_t)gridDim.x * 256) {
        const int x = (int)(p % G.n1), y = (int)((p / G.n1) % G.n2), z = (int)(p / ((size_t)G.n1 * G.n2));
        double2 d[3][3];   /* d[ax][c] = (C_{ax+1} x_c)(p) */
        for (int ax = 0; ax < 3; ++ax) {
            double ang;
            bool wrap;
            const int q = curl_nbr(G, ax, x, y, z, ang, wrap);
            double2 w = make_double2(1.0, 0.0);
            if (wrap) sincos(ang, &w.y, &w.x);
            for (int c = 0; c < 3; ++c)
                d[ax][c] = cscale(csub(cmul(w, X[c * N + q]), X[c * N + p]), G.ih[ax]);
        }
        for (int c = 0; c < 3; ++c) dif[c * N + p] = d[c][c];
        cur[p] = csub(d[1][2], d[2][1]);
        cur[N + p] = csub(d[2][0], d[0][2]);
        cur[2 * N + p] = csub(d[0][1], d[1][0]);
    }
}

/* partials of ||a_c - b_c||^2 per component c (one block row per component) */
__global__ void __launch_bounds__(256) k_diff_norm(const double2* a, const double2* b, double* partials, size_t n) {
    const int c = blockIdx.y;
    double s = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double2 u = a[c * n + i], v = b[c * n + i];
        const double dx = u.x - v.x, dy = u.y - v.y;
        s += dx * dx + dy * dy;
    }
    __shared__ double r[256];
    r[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) r[threadIdx.x] += r[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[(size_t)c * gridDim.x + blockIdx.x] = r[0];
}

extern "C" fame_status fame_curl_coo(const fame_lattice* lat, const double* k, int* rows, int* cols, double* vals) {
    if (!lat || !k || !rows || !cols || !vals) FAIL(FAME_ERR_ARG, "fame_curl_coo: bad arguments");
    if (!curl_geo_ok(*lat)) FAIL(FAME_ERR_UNSUPPORTED, "fame_curl_coo: boundary shifts outside the grid");
    const CurlGeo G = curl_geo(*lat, k);
    const int N = G.n1 * G.n2 * G.n3;
    if (N < 1) FAIL(FAME_ERR_ARG, "fame_curl_coo: empty grid");
    /* block (row component, column component, sign) of each C_{ax+1} in C */
    const int blk[3][2][3] = {{{1, 2, -1}, {2, 1, +1}},    /* C_1: -C1 at (y, z), +C1 at (z, y) */
                              {{0, 2, +1}, {2, 0, -1}},    /* C_2: +C2 at (x, z), -C2 at (z, x) */
                              {{0, 1, -1}, {1, 0, +1}}};   /* C_3: -C3 at (x, y), +C3 at (y, x) */
    size_t e = 0;
    for (int ax = 0; ax < 3; ++ax)
        for (int z = 0; z < G.n3; ++z)
            for (int y = 0; y < G.n2; ++y)
                for (int x = 0; x < G.n1; ++x) {
                    const int p = x + G.n1 * (y + G.n2 * z);
                    double ang;
                    bool wrap;
                    const int q = curl_nbr(G, ax, x, y, z, ang, wrap);
                    const double wr = wrap ? std::cos(ang) : 1.0, wi = wrap ? std::sin(ang) : 0.0;
                    for (int b = 0; b < 2; ++b) {
                        const int rc = blk[ax][b][0], cc = blk[ax][b][1];
                        const double sg = blk[ax][b][2] * G.ih[ax];
                        rows[e] = rc * N + p; cols[e] = cc * N + p; vals[2 * e] = -sg; vals[2 * e + 1] = 0.0; ++e;
                        rows[e] = rc * N + p; cols[e] = cc * N + q; vals[2 * e] = sg * wr; vals[2 * e + 1] = sg * wi; ++e;
                    }
                }
    return FAME_OK;
}

extern "C" fame_status fame_check_eigendecomp(fame_ctx* c, int ik, unsigned long long seed, double* err) {
    if (!c || !err || ik < 0 || ik >= c->nk) FAIL(FAME_ERR_ARG, "fame_check_eigendecomp: bad arguments");
    if (!curl_geo_ok(c->lat)) FAIL(FAME_ERR_UNSUPPORTED, "fame_check_eigendecomp: boundary shifts outside the grid");
    HIPCHK(hipSetDevice(c->device));
    const size_t N = (size_t)c->N, N3 = 3 * N;
    /* the seeded field: uniform real and imaginary parts in [0, 1) (splitmix64) */
    std::vector<double2> xh(N3);
    unsigned long long st = seed;
    auto next = [&st]() {
        unsigned long long z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    for (size_t i = 0; i < N3; ++i) { xh[i].x = next(); xh[i].y = next(); }
    double2 *X = nullptr, *D = nullptr, *Cx = nullptr, *Y = nullptr;
    double* part = nullptr;
    const int NB = 256;
    fame_status rs = FAME_OK;
    auto run = [&]() -> fame_status {
        HIPCHK(hipMalloc(&X, 16 * N3));
        HIPCHK(hipMalloc(&D, 16 * N3));
        HIPCHK(hipMalloc(&Cx, 16 * N3));
        HIPCHK(hipMalloc(&Y, 16 * N3));
        HIPCHK(hipMalloc(&part, sizeof(double) * 3 * NB));
        HIPCHK(hipMemcpyAsync(X, xh.data(), 16 * N3, hipMemcpyHostToDevice, c->stream));
        /* sparse side */
        const CurlGeo G = curl_geo(c->lat, &c->kvec[3 * (size_t)ik]);
        hipLaunchKernelGGL(k_curl_stencil, dim3(1024), dim3(256), 0, c->stream, X, D, Cx, G);
        HIPCHK(hipGetLastError());
        /* T Lambda_i T* x_i for the three components at once (the curl-curl
         * pipeline with the Lambda step) */
        CHK(c->field_from_ref(X, c->F64 + (size_t)ik * N3, 1));
        KInner ki = c->kin(8, ik, nullptr, LIST_FI);
        HIPCHK(fame_launch_outer(c->kout(8, ik, nullptr, LIST_FWD), OUTER_QRS, 8, 1, c->stream));
        HIPCHK(fame_launch_mid(c->kmid(8, ik, nullptr, LIST_FWD), -1, 8, 1, c->stream));
        dim3 grid((c->g.nb * c->g.nc + ki.L - 1) / ki.L, 1);
        const size_t lds = fame_line_lds(ki.plan, 8, 3 * ki.L, ki.P, 0);
        if (plan_engine(ki.plan) == ENG_MFMA)
            hipLaunchKernelGGL((k_curlcurl<ENG_MFMA, STEP_LAMBDA>), grid, dim3(FFT_THREADS), lds, c->stream, ki);
        else hipLaunchKernelGGL((k_curlcurl<ENG_STOCKHAM, STEP_LAMBDA>), grid, dim3(FFT_THREADS), lds, c->stream, ki);
        HIPCHK(hipGetLastError());
        HIPCHK(fame_launch_mid(c->kmid(8, ik, nullptr, LIST_INV), +1, 8, 1, c->stream));
        HIPCHK(fame_launch_outer(c->kout(8, ik, nullptr, LIST_INV), OUTER_QR, 8, 1, c->stream));
        CHK(c->field_to_ref(c->F64 + (size_t)ik * N3, Y, 1, N3));
        hipLaunchKernelGGL(k_diff_norm, dim3(NB, 3), dim3(256), 0, c->stream, D, Y, part, N);
        HIPCHK(hipGetLastError());
        std::vector<double> hp(3 * NB);
        HIPCHK(hipMemcpyAsync(hp.data(), part, sizeof(double) * 3 * NB, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < 3; ++i) {
            double a = 0;
            for (int t = 0; t < NB; ++t) a += hp[(size_t)i * NB + t];
            err[i] = std::sqrt(a);
        }
        /* P Sigma Q^H x: Qrs, the singular values |Lambda|, Pr */
        double2* q = (double2*)c->ref64[1];
        CHK(c->field_from_ref(X, c->F64 + (size_t)ik * N3, 1));
        CHK(c->qrs_field(q, ik, 1, 0, nullptr));
        HIPCHK(fame_scale_real_mod(q, q, c->sig + (size_t)ik * N, 2 * N, 2 * N, N, N, 1, 8, c->stream));
        CHK(c->qr_field(q, ik, 1, 1, nullptr));
        CHK(c->field_to_ref(c->F64 + (size_t)ik * N3, Y, 1, N3));
        hipLaunchKernelGGL(k_diff_norm, dim3(NB, 3), dim3(256), 0, c->stream, Cx, Y, part, N);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hp.data(), part, sizeof(double) * 3 * NB, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double a = 0;
        for (double v : hp) a += v;
        err[3] = std::sqrt(a);
        return FAME_OK;
    };
    rs = run();
    (void)hipStreamSynchronize(c->stream);
    for (void* ptr : {(void*)X, (void*)D, (void*)Cx, (void*)Y, (void*)part}) if (ptr) (void)hipFree(ptr);
    return rs;
}
