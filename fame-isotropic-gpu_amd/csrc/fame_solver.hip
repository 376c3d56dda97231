/*
 * fame_solver.hip -- host side of the gfx950 band solver behind include/fame_hip.h.
 *
 * Owns the FFT plans and twiddle tables, the per-wave-vector phases, the
 * permittivity in the kernels' layout and all work buffers, and drives:
 *   QBQ         five axis passes (fame_kernels.hip)
 *   CG          batched, device-resident scalars, per-problem masks (CG.cu:5-87)
 *   invAr       Sigma^-1 CG(QBQ, Sigma^-1 x) in FP64, or FP32 CG inside FP64
 *               iterative refinement (..._Isotropic_invAr.cu:8-48)
 *   Lanczos     e_n start, three-term recurrence, CGS2 full reorthogonalisation,
 *               exact-shift lock/purge restart applied as one U*Q product,
 *               Ritz vectors by U*z (Lanczos_decomp_gpu.cu, Lanczos_LockPurge_gpu.cu,
 *               Lanczos_Isotropic.cu)
 *   restoration E = invB .* Qr(Sigma y) scaled by ||E|| and the Gamma rule
 *               (FAME_Fast_Algorithms_Isotropic.cu:57-125)
 */
#include "fame_hip.h"
#include "fame_kernels.h"
#include "fame_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace fame;

static thread_local std::string g_err;
extern "C" const char* fame_last_error(void) { return g_err.c_str(); }
std::string& fame_err_ref() { return g_err; }

#define FAIL(code, msg) do { g_err = (msg); return (code); } while (0)
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    g_err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" + std::to_string(__LINE__); \
    return FAME_ERR_HIP; } } while (0)
#define CHK(x) do { fame_status s_ = (x); if (s_ != FAME_OK) return s_; } while (0)

/* ======================================================================== */
/*  FFT plans                                                                */
/* ======================================================================== */
/* relative cost of one element through one pass of radix r (LDS round trip,
 * twiddle and butterfly arithmetic); 6, 9, 10, 12, 15 and 16 are two-level
 * in-register butterflies that save a pass */
static double pass_cost(int r) {
    switch (r) {
        case 2: return 18; case 3: return 20; case 4: return 20; case 5: return 23; case 6: return 24;
        case 7: return 26; case 8: return 23; case 9: return 26; case 10: return 26; case 11: return 30;
        case 12: return 26; case 13: return 32; case 15: return 29; case 16: return 27;
        default: return 1e9;
    }
}
/* radices the FP64 kernels are compiled for (the single-level butterflies) */
static bool is_prime_radix(int r) { return r == 2 || r == 3 || r == 4 || r == 5 || r == 7 || r == 8 || r == 11 || r == 13; }

/* cheapest factorisation of n into the supported radices, largest radix first
 * (the twiddle-free first pass gets the largest butterfly).  compound = false
 * leaves out 6, 9, 10, 12, 15 and 16: the FP64 kernels hold 8 complex elements
 * per thread, and the extra registers of the two-level butterflies cost them
 * more waves than the saved pass gives back (measured on gfx950).           */
static bool factor_gpu(int n, std::vector<int>* rad, bool compound = true) {
    static const int RADS[] = {16, 15, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2};
    std::vector<double> cost(n + 1, 1e30);
    std::vector<int> pick(n + 1, 0);
    cost[1] = 0;
    for (int m = 2; m <= n; ++m) {
        if (n % m) continue;                   /* only divisors of n are needed */
        for (int r : RADS)
            if ((compound || is_prime_radix(r)) && m % r == 0 && cost[m / r] < 1e29 &&
                cost[m / r] + pass_cost(r) < cost[m]) {
                cost[m] = cost[m / r] + pass_cost(r);
                pick[m] = r;
            }
    }
    if (n < 1 || cost[n] > 1e29) return false;
    std::vector<int> r;
    for (int m = n; m > 1; m /= pick[m]) r.push_back(pick[m]);
    std::sort(r.begin(), r.end(), [](int x, int y) { return x > y; });
    if (rad) *rad = r;
    return true;
}
static double plan_cost(const std::vector<int>& rad) { double c = 0; for (int r : rad) c += pass_cost(r); return c; }

unsigned fame_magic32(int d) {
    return d <= 1 ? 0u : (unsigned)(((1ULL << 32) + (unsigned long long)d - 1) / (unsigned long long)d);
}

typedef std::complex<long double> cld;

AxisPlan::~AxisPlan() {
    for (void* p : bufs) if (p) (void)hipFree(p);
}

/* append one transform of sign `sign` to a stage list */
static bool add_transform(LinePlanDev& d, const std::vector<int>& rad, int M, int sign, int blue, bool prechirp) {
    auto passes = [&](int sgn, int pre0) -> bool {
        int Ns = 1;
        for (size_t i = 0; i < rad.size(); ++i) {
            if (d.nst >= FAME_MAX_STAGES) return false;
            StageDev& s = d.st[d.nst++];
            s.R = rad[i]; s.Ns = Ns; s.nb = M / s.R; s.tstep = M / (Ns * s.R);
            s.mNs = fame_magic32(Ns); s.mNb = fame_magic32(s.nb);
            s.sign = sgn; s.pre = i == 0 ? pre0 : PRE_NONE;
            Ns *= s.R;
        }
        return true;
    };
    if (!blue) return passes(sign, PRE_NONE);
    /* chirp-z: X = post .* IFFT_M( bhat_sign .* FFT_M( pre .* x ) ) */
    return passes(-1, prechirp ? (sign < 0 ? PRE_CHIRP : PRE_CHIRP_CONJ) : PRE_NONE) &&
           passes(+1, sign < 0 ? PRE_BM : PRE_BP);
}
/* append one direct DFT (matrix cores) of length n */
static bool add_direct(LinePlanDev& d, int n, int sign) {
    if (d.nst >= FAME_MAX_STAGES) return false;
    StageDev& s = d.st[d.nst++];
    std::memset(&s, 0, sizeof(s));
    s.R = DFT_MFMA; s.sign = sign;
    s.dn = n; s.dm = (n - 1) / 2; s.dQ = n / 2 + 1;
    s.dK = std::max(4, (s.dm + 3) / 4 * 4);
    s.dqt = (s.dQ + 15) / 16;
    return true;
}
static bool add_mid(LinePlanDev& d) {
    if (d.nst >= FAME_MAX_STAGES) return false;
    StageDev& s = d.st[d.nst++];
    std::memset(&s, 0, sizeof(s));
    return true;
}

fame_status AxisPlan::build(int n_, int engine) {
    n = n_;
    /* lengths 16..127: direct DFT on the matrix cores (any n, no Bluestein);
     * otherwise Stockham passes, or Bluestein when n has a prime factor > 13
     * or the chirp-z transform is cheaper */
    mfma = (engine == FFT_ENGINE_MFMA && n >= 2 && n <= 127) || (engine == FFT_ENGINE_AUTO && n >= 16 && n <= 127);
    std::vector<int> rad;
    double direct = 1e30;
    if (factor_gpu(n, &rad)) direct = n * plan_cost(rad);
    int bestM = 0; std::vector<int> bestR; double bestC = 1e30;
    for (int m = 2 * n - 1; m <= 4 * n + 8 && !mfma; ++m) {
        std::vector<int> r2;
        if (!factor_gpu(m, &r2)) continue;
        double c = 2.0 * m * plan_cost(r2) + 12.0 * m;
        if (c < bestC) { bestC = c; bestM = m; bestR = r2; }
    }
    if (mfma) { blue = 0; M = n; radices.clear(); }
    else if (direct <= bestC && direct < 1e29) { blue = 0; M = n; radices = rad; }
    else if (bestM) { blue = 1; M = bestM; radices = bestR; }
    else FAIL(FAME_ERR_UNSUPPORTED, "no FFT plan for length " + std::to_string(n));
    /* tables in long double, rounded once to each precision */
    const long double PI_L = 3.14159265358979323846264338327950288L;
    std::vector<cld> W(M), ch, bm, bp;
    for (int m = 0; m < M; ++m) W[m] = cld(cosl(2 * PI_L * m / M), -sinl(2 * PI_L * m / M));
    if (blue) {
        ch.assign(M, 0);
        for (int m = 0; m < n; ++m) {
            long long q = ((long long)m * m) % (2LL * n);
            ch[m] = cld(cosl(PI_L * q / n), -sinl(PI_L * q / n));     /* exp(-i pi m^2/n) */
        }
        /* kernels b = conj(chirp) wrapped, for each sign; bhat = DFT_M(b) (exp(-)) */
        std::vector<cld> b1(M, 0), b2(M, 0);
        for (int m = 0; m < n; ++m) {
            b1[m] = std::conj(ch[m]);            /* sign -1: chirp = exp(-i..) -> b = exp(+i..) */
            b2[m] = ch[m];                       /* sign +1: chirp = exp(+i..) -> b = exp(-i..) */
            if (m) { b1[M - m] = b1[m]; b2[M - m] = b2[m]; }
        }
        bm.assign(M, 0); bp.assign(M, 0);
        for (int k = 0; k < M; ++k) {
            cld s1 = 0, s2 = 0;
            for (int j = 0; j < M; ++j) {
                cld w = W[(long long)j * k % M];
                s1 += b1[j] * w; s2 += b2[j] * w;
            }
            bm[k] = s1; bp[k] = s2;
        }
    }
    auto upload = [&](const std::vector<cld>& v, int prec, const void** dst) -> fame_status {
        if (v.empty()) { *dst = nullptr; return FAME_OK; }
        void* d = nullptr;
        size_t bytes = v.size() * (prec == 8 ? 16 : 8);
        HIPCHK(hipMalloc(&d, bytes));
        bufs.push_back(d);
        if (prec == 8) {
            std::vector<double> h(2 * v.size());
            for (size_t i = 0; i < v.size(); ++i) { h[2 * i] = (double)v[i].real(); h[2 * i + 1] = (double)v[i].imag(); }
            HIPCHK(hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice));
        } else {
            std::vector<float> h(2 * v.size());
            for (size_t i = 0; i < v.size(); ++i) { h[2 * i] = (float)v[i].real(); h[2 * i + 1] = (float)v[i].imag(); }
            HIPCHK(hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice));
        }
        *dst = d;
        return FAME_OK;
    };
    /* FP64 lists: the same length M factored without the compound radices */
    std::vector<int> rad64 = radices;
    if (!mfma && !factor_gpu(M, &rad64, false)) FAIL(FAME_ERR_UNSUPPORTED, "no FP64 FFT plan for length " + std::to_string(M));
    for (int r : rad64)
        if (!is_prime_radix(r)) FAIL(FAME_ERR_UNSUPPORTED, "FP64 FFT plan with a compound radix");
    for (int pi = 0; pi < 2; ++pi) {
        const int prec = pi == 0 ? 8 : 4;
        const std::vector<int>& rads = pi == 0 ? rad64 : radices;
        LinePlanDev base;
        std::memset(&base, 0, sizeof(base));
        base.n = n; base.M = M; base.blue = blue;
        base.ntab = blue ? 4 * M : M;
        base.mM = fame_magic32(M); base.mN = fame_magic32(n);
        CHK(upload(W, prec, &base.W));
        CHK(upload(ch, prec, &base.chirp));
        CHK(upload(bm, prec, &base.bm));
        CHK(upload(bp, prec, &base.bp));
        bool ok = true;
        for (int l = 0; l < NLIST && mfma; ++l) {
            LinePlanDev& d = lists[pi][l];
            d = base;
            switch (l) {
                case LIST_FWD: ok = ok && add_direct(d, n, -1); break;
                case LIST_INV: ok = ok && add_direct(d, n, +1); break;
                case LIST_EPS: ok = ok && add_direct(d, n, +1) && add_mid(d) && add_direct(d, n, -1); break;
                default: ok = ok && add_direct(d, n, -1) && add_mid(d) && add_direct(d, n, +1); break;
            }
        }
        for (int l = 0; l < NLIST && !mfma; ++l) {
            LinePlanDev& d = lists[pi][l];
            d = base;
            switch (l) {
                case LIST_FWD: ok = ok && add_transform(d, rads, M, -1, blue, true); d.post = blue ? POST_CHIRP : POST_NONE; break;
                case LIST_INV: ok = ok && add_transform(d, rads, M, +1, blue, true); d.post = blue ? POST_CHIRP_CONJ : POST_NONE; break;
                case LIST_EPS:
                    ok = ok && add_transform(d, rads, M, +1, blue, true) && add_mid(d) && add_transform(d, rads, M, -1, blue, false);
                    d.post = blue ? POST_CHIRP : POST_NONE;
                    break;
                default:
                    ok = ok && add_transform(d, rads, M, -1, blue, true) && add_mid(d) && add_transform(d, rads, M, +1, blue, false);
                    d.post = blue ? POST_CHIRP_CONJ : POST_NONE;
                    break;
            }
        }
        if (!ok) FAIL(FAME_ERR_UNSUPPORTED, "too many FFT passes for length " + std::to_string(n));
        /* register budget per thread: 8 (FP64) / 16 (FP32) complex elements */
        const int maxe = prec == 8 ? MaxE<double>::v : MaxE<float>::v;
        int c = 1 << 12;
        for (int r : rads) c = std::min(c, r * nbt_of(r, maxe));
        cap[pi] = FFT_THREADS * c;     /* direct DFT stages hold no lines in registers */
    }
    return FAME_OK;
}

/* LDS entries of one block beyond the lines: the middle-axis kernels keep one
 * twiddle per line position, the inner-axis kernels a phase per mode */
static int tile_extra(int axis, const GridDev& g, int prec, int L) {
    return axis == AX_MID ? g.nb : (axis == AX_INNER ? inner_lds_extra(g.na, prec, L) : 0);
}
static int tile_ncomp(int axis) { return axis == AX_INNER ? 3 : 1; }

/* is L lines per block feasible for this plan: register budget of the Stockham
 * passes (ncomp x L lines of M elements within cap), at most 32 lines for the
 * inner kernels, LDS within 60 KiB (+ the kernels' static LDS < 64 KiB) */
static bool tile_ok(const AxisPlan& p, int axis, int prec, int L, const GridDev& g) {
    const int pi = prec == 8 ? 0 : 1, nc = tile_ncomp(axis);
    if (L < 1 || (axis == AX_INNER && L > MAX_LINES_INNER)) return false;
    if (!p.mfma && (long long)nc * L * p.M > p.cap[pi]) return false;
    return fame_line_lds(p.lists[pi][0], prec, nc * L, p.M | 1, tile_extra(axis, g, prec, L)) <= 60 * 1024;
}

/* default lines per block before tuning: as many as feasible up to maxL, rounded
 * down to a full 128-byte row (8 FP64 / 16 FP32 lines) on the strided axes */
static int default_lines(const AxisPlan& p, int axis, int prec, const GridDev& g) {
    const int maxL = axis == AX_INNER ? MAX_LINES_INNER : (p.mfma ? 32 : 64);
    const int align = axis == AX_INNER ? 1 : (prec == 8 ? 8 : 16);
    int l = maxL;
    while (l > 1 && !tile_ok(p, axis, prec, l, g)) --l;
    if (!tile_ok(p, axis, prec, l, g)) return 0;
    if (align > 1 && l >= align) l = (l / align) * align;
    return l;
}

/* ======================================================================== */
/*  Context                                                                  */
/* ======================================================================== */
static fame_status dalloc(fame_ctx* c, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *p = nullptr;
        FAIL(FAME_ERR_NOMEM, std::string("device allocation of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    c->allocs.push_back(*p);
    return FAME_OK;
}

void fame_ctx::drop_graphs() {
    for (auto& kv : cg_graphs) (void)hipGraphExecDestroy(kv.second);
    cg_graphs.clear();
}

fame_ctx::~fame_ctx() {
    drop_graphs();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (void* p : allocs) if (p) (void)hipFree(p);
    for (void* p : big) if (p) (void)hipFree(p);
}

extern "C" fame_status fame_create(const fame_lattice* lat, int max_batch, int device, fame_ctx** out) {
    if (!lat || !out || max_batch < 1) FAIL(FAME_ERR_ARG, "fame_create: null argument or max_batch < 1");
    for (int i = 0; i < 3; ++i) if (lat->grid_nums[i] < 1 || lat->grid_nums[i] > 1024) FAIL(FAME_ERR_ARG, "grid_nums out of range");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) FAIL(FAME_ERR_ARG, "fame_create: no such device");
    HIPCHK(hipSetDevice(device));
    fame_ctx* c = new fame_ctx();
    c->device = device;
    c->lat = *lat;
    c->maxb = max_batch;
    const int Nx = lat->grid_nums[0], Ny = lat->grid_nums[1], Nz = lat->grid_nums[2];
    GridDev& g = c->g;
    g.general = lat->general ? 1 : 0;
    {
        const double* h = lat->mesh_len;
        auto close = [](double u, double v) { return std::fabs(u - v) <= 1e-12 * std::fabs(u); };
        g.pi_guard = g.general && close(h[0], h[1]) && close(h[1], h[2]) ? 1 : 0;
    }
    if (g.general) { g.na = Nz; g.nb = Ny; g.nc = Nx; g.ha = lat->mesh_len[2]; g.hb = lat->mesh_len[1]; g.hc = lat->mesh_len[0]; }
    else { g.na = Nx; g.nb = Ny; g.nc = Nz; g.ha = lat->mesh_len[0]; g.hb = lat->mesh_len[1]; g.hc = lat->mesh_len[2]; }
    if ((long long)Nx * Ny * Nz * 3 >= (1LL << 31)) { delete c; FAIL(FAME_ERR_ARG, "grid too large: 3N must stay below 2^31"); }
    g.N = Nx * Ny * Nz;
    g.iha = 1.0 / g.ha; g.ihb = 1.0 / g.hb; g.ihc = 1.0 / g.hc;
    g.m_na = fame_magic32(g.na);
    g.m_nb = fame_magic32(g.nb);
    c->N = g.N;
    fame_status st;
    /* FAME_FFT=stockham|mfma restricts the line-transform engine (default: both
     * are planned, the matrix cores are preferred for lengths 16..127 and the
     * autotuner may change the choice); FAME_AUTOTUNE=0 disables tuning */
    if (const char* e = std::getenv("FAME_FFT")) {
        if (!std::strcmp(e, "stockham")) c->eng_forced = 0;
        else if (!std::strcmp(e, "mfma")) c->eng_forced = 1;
    }
    if (const char* e = std::getenv("FAME_AUTOTUNE")) c->autotune = std::strcmp(e, "0") != 0;
    if (const char* e = std::getenv("FAME_GRAPH")) c->use_graphs = std::strcmp(e, "0") != 0;
    const int len[3] = {g.na, g.nb, g.nc};
    for (int a = 0; a < 3; ++a) {
        if ((st = c->ax[a][0].build(len[a], FFT_ENGINE_STOCKHAM)) != FAME_OK) { delete c; return st; }
        if (c->eng_forced != 0 && len[a] >= 2 && len[a] <= 127 &&
            (st = c->ax[a][1].build(len[a], FFT_ENGINE_MFMA)) != FAME_OK) { delete c; return st; }
    }
    for (int pi = 0; pi < 2; ++pi) {
        const int prec = pi == 0 ? 8 : 4;
        for (int a = 0; a < 3; ++a) {
            int eng = 0;
            if (c->ax[a][1].n > 0 && (c->eng_forced == 1 || (c->eng_forced < 0 && len[a] >= 16))) eng = 1;
            PassTile& t = c->tile[a][pi];
            t.eng = eng;
            t.L = default_lines(c->ax[a][eng], a, prec, g);
            t.P = c->ax[a][eng].M | 1;
            if (t.L < 1) {
                delete c;
                FAIL(FAME_ERR_UNSUPPORTED, "FFT length too large for the LDS-resident line transforms");
            }
        }
    }
    const size_t N = g.N, B = max_batch;
    c->nblk_in[0] = (g.nb * g.nc + c->tile[AX_INNER][0].L - 1) / c->tile[AX_INNER][0].L;
    c->nblk_in[1] = (g.nb * g.nc + c->tile[AX_INNER][1].L - 1) / c->tile[AX_INNER][1].L;
#define A(ptr, bytes) do { fame_status s_ = dalloc(c, (void**)&(ptr), (bytes)); if (s_ != FAME_OK) { delete c; return s_; } } while (0)
    A(c->prob_d, sizeof(ProbDev) * B);
    A(c->sig, sizeof(double) * N * B);
    A(c->sig_inv, sizeof(double) * N * B);
    A(c->invB64, sizeof(double) * 3 * N);
    A(c->invB32, sizeof(float) * 3 * N);
    A(c->eps_int, sizeof(double) * 3 * N);
    A(c->eps_ref, sizeof(double) * 3 * N);
    A(c->F64, 16 * 3 * N * B);
    A(c->F32, 8 * 3 * N * B);
    for (int i = 0; i < 4; ++i) A(c->cg64[i], 16 * 2 * N * B);
    for (int i = 0; i < 5; ++i) A(c->cg32[i], 8 * 2 * N * B);
    for (int i = 0; i < 3; ++i) A(c->ref64[i], 16 * 2 * N * B);
    /* inner-kernel partials: room for one line per block (any tuned tile) */
    int nbmax = std::max(g.nb * g.nc, NBLK_VEC);
    A(c->partials, sizeof(double) * 4 * nbmax * B);
    A(c->partials2, sizeof(double) * 6 * nbmax * B);
    A(c->pap_part, sizeof(double) * 3 * xcd_round(g.na * g.nb) * B);   /* the most outer blocks: one line each */
    A(c->cgsc, sizeof(CgScal) * B);
    A(c->active, sizeof(int) * B);
    A(c->mask, sizeof(int) * B);
    A(c->scal, sizeof(double) * 64 * B);
#undef A
    HIPCHK(hipHostMalloc((void**)&c->sc_h, sizeof(CgScal) * B, 0));
    c->big.clear();
    *out = c;
    return FAME_OK;
}

extern "C" fame_status fame_destroy(fame_ctx* c) {
    if (!c) return FAME_OK;
    for (fame_ctx* k : c->kids) (void)fame_destroy(k);
    c->kids.clear();
    if (c->owns_stream && c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        c->stream = nullptr;
    }
    if (c->sc_h) (void)hipHostFree(c->sc_h);
    delete c;
    return FAME_OK;
}

extern "C" fame_status fame_set_stream(fame_ctx* c, void* stream) {
    if (!c) FAIL(FAME_ERR_ARG, "null context");
    c->stream = (hipStream_t)stream;
    c->drop_graphs();
    return FAME_OK;
}

extern "C" fame_status fame_query(const fame_ctx* c, int* N, int* nk, int* len3) {
    if (!c) FAIL(FAME_ERR_ARG, "null context");
    if (N) *N = c->N;
    if (nk) *nk = c->nk;
    if (len3) for (int a = 0; a < 3; ++a) len3[a] = c->plan(a, 8).M;
    return FAME_OK;
}

/* reference layout (component-stacked, x fastest) -> kernel layout [c][b][a] */
static size_t int_index(const GridDev& g, int x, int y, int z) {
    if (g.general) return ((size_t)x * g.nb + y) * g.na + z;    /* [x][y][z] */
    return x + (size_t)g.na * (y + (size_t)g.nb * z);            /* [z][y][x] */
}

extern "C" fame_status fame_set_eps(fame_ctx* c, const double* eps) {
    if (!c || !eps) FAIL(FAME_ERR_ARG, "fame_set_eps: null argument");
    const int Nx = c->lat.grid_nums[0], Ny = c->lat.grid_nums[1], Nz = c->lat.grid_nums[2];
    const size_t N = c->N;
    std::vector<double> e(3 * N), ib(3 * N);
    std::vector<float> ib32(3 * N);
    for (int comp = 0; comp < 3; ++comp)
        for (int z = 0; z < Nz; ++z)
            for (int y = 0; y < Ny; ++y)
                for (int x = 0; x < Nx; ++x) {
                    size_t r = x + (size_t)Nx * (y + (size_t)Ny * z);
                    double v = eps[comp * N + r];
                    if (!(v > 0)) FAIL(FAME_ERR_ARG, "fame_set_eps: permittivity must be positive");
                    size_t q = comp * N + int_index(c->g, x, y, z);
                    e[q] = v; ib[q] = 1.0 / v; ib32[q] = (float)(1.0 / v);
                }
    /* the context's stream may still be running kernels that read the old values */
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    /* ordered on the context's stream (contexts of fame_eigs_many run on their
     * own streams); the host staging vectors live until the stream is drained */
    HIPCHK(hipMemcpyAsync(c->eps_int, e.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->eps_ref, eps, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->invB64, ib.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->invB32, ib32.data(), sizeof(float) * 3 * N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_eps = true;
    ++c->eps_ver;
    return FAME_OK;
}

/* per-k phases, FAME_Matrix_Lambdas.cu:40-76 (theta's) and :155-157 / :402-406 (Gamma) */
void fame_phases(const fame_lattice& L, const double* k, ProbDev* p) {
    const double* a = L.lattice_vec_a;
    auto dot3 = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    const int Nx = L.grid_nums[0], Ny = L.grid_nums[1], Nz = L.grid_nums[2];
    const double kn = std::sqrt(dot3(k, k));
    std::memset(p, 0, sizeof(*p));
    if (!L.general) {
        p->t0 = dot3(k, a); p->t1 = 0; p->t2 = 0;
        p->s0 = dot3(k, a + 3); p->s1 = 0;
        p->r0 = dot3(k, a + 6);
        p->gamma = kn < 1e-4;
    } else {
        double c1 = a[3] < 0 ? 1. - (double)L.m1 / (double)Nx : -(double)L.m1 / (double)Nx;
        double c2 = (double)L.m3 / (double)Ny, c3 = (double)L.m2 / (double)Nx;
        double a2h[3], a3h[3];
        for (int i = 0; i < 3; ++i) a2h[i] = c1 * a[i] + a[3 + i];
        for (int i = 0; i < 3; ++i) a3h[i] = a[6 + i] - c2 * a2h[i] - c3 * a[i] + L.t1[i];
        p->t0 = dot3(k, a3h);
        p->t1 = -((double)L.m3 / (double)Ny);
        p->t2 = ((double)L.m1 * (double)L.m3) / ((double)Ny * Nx) - ((double)L.m2 / (double)Nx);
        p->s0 = dot3(k, a2h);
        p->s1 = -(double)L.m1 / (double)Nx;
        p->r0 = dot3(k, a);
        p->gamma = kn < 1e-3;
        (void)Nz;
    }
}

extern "C" fame_status fame_set_wave_vectors(fame_ctx* c, int nk, const double* k, int* nd_out) {
    if (!c || !k || nk < 1) FAIL(FAME_ERR_ARG, "fame_set_wave_vectors: bad arguments");
    if (nk > c->maxb) FAIL(FAME_ERR_ARG, "fame_set_wave_vectors: more wave vectors than max_batch");
    HIPCHK(hipSetDevice(c->device));
    c->nk = nk;
    c->kvec.assign(k, k + 3 * (size_t)nk);
    c->prob_h.resize(nk);
    c->nd.resize(nk);
    for (int i = 0; i < nk; ++i) {
        fame_phases(c->lat, k + 3 * i, &c->prob_h[i]);
        c->nd[i] = c->prob_h[i].gamma ? c->N - 1 : c->N;
        if (nd_out) nd_out[i] = c->nd[i];
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpyAsync(c->prob_d, c->prob_h.data(), sizeof(ProbDev) * nk, hipMemcpyHostToDevice, c->stream));
    HIPCHK(fame_launch_sigma(c->g, c->prob_d, c->sig, c->sig_inv, nk, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    ++c->k_ver;
    return FAME_OK;
}

/* ======================================================================== */
/*  Operators                                                                */
/* ======================================================================== */
static inline int pix(int prec) { return prec == 8 ? 0 : 1; }

KInner fame_ctx::kin(int prec, int p0, const int* act, int list) const {
    KInner a;
    std::memset(&a, 0, sizeof(a));
    const PassTile& t = tile[AX_INNER][pix(prec)];
    a.g = g; a.prob = prob_d + p0; a.plan = ax[AX_INNER][t.eng].lists[pix(prec)][list]; a.active = act;
    a.f = prec == 8 ? (void*)(F64 + (size_t)p0 * 3 * N) : (void*)(F32 + (size_t)p0 * 3 * N);
    a.xstride = 2 * (size_t)N; a.fstride = 3 * (size_t)N;
    a.L = t.L; a.P = t.P;
    return a;
}
KAxis fame_ctx::kmid(int prec, int p0, const int* act, int list) const {
    KAxis a;
    std::memset(&a, 0, sizeof(a));
    const PassTile& t = tile[AX_MID][pix(prec)];
    a.g = g; a.prob = prob_d + p0; a.plan = ax[AX_MID][t.eng].lists[pix(prec)][list]; a.active = act;
    a.f = prec == 8 ? (void*)(F64 + (size_t)p0 * 3 * N) : (void*)(F32 + (size_t)p0 * 3 * N);
    a.fstride = 3 * (size_t)N; a.L = t.L; a.P = t.P;
    a.mL = fame_magic32(a.L);
    return a;
}
KAxis fame_ctx::kout(int prec, int p0, const int* act, int list) const {
    KAxis a = kmid(prec, p0, act, list);
    const PassTile& t = tile[AX_OUTER][pix(prec)];
    a.plan = ax[AX_OUTER][t.eng].lists[pix(prec)][list];
    a.invB = prec == 8 ? (const void*)invB64 : (const void*)invB32;
    a.L = t.L; a.P = t.P;
    a.mL = fame_magic32(a.L);
    return a;
}

int fame_ctx::pap_blocks(int prec) const {
    for (int axis = 0; axis < 3; ++axis)
        if (plan(axis, prec).lists[pix(prec)][LIST_FWD].blue) return 0;
    const int L = tile[AX_OUTER][pix(prec)].L;
    return xcd_round((g.na * g.nb + L - 1) / L);
}

/* y = QBQ x on nprob problems starting at p0 (padded reduced layout, 2N each);
 * with dotvec, partial sums of conj(dotvec).y go to partials2 per inner block. */
fame_status fame_ctx::qbq(int prec, const void* x, void* y, int p0, int nprob, const int* act, const void* dotvec,
                          void* cg_r, void* cg_x, const void* cg_ap, bool cgdots) {
    KInner ki = kin(prec, p0, act, LIST_INV);
    ki.x = x; ki.pikind = 0;
    if (cg_r) { ki.cg_r = cg_r; ki.cg_x = cg_x; ki.cgsc = cgsc; ki.cg_ap = cg_ap; ki.partials = partials2; }
    HIPCHK(fame_launch_inner(ki, 1, prec, nprob, stream));
    ki.cg_r = nullptr; ki.cg_x = nullptr; ki.cgsc = nullptr; ki.cg_ap = nullptr;
    ki.cgdots = cgdots;
    HIPCHK(fame_launch_mid(kmid(prec, p0, act, LIST_INV), +1, prec, nprob, stream));
    HIPCHK(fame_launch_outer(kout(prec, p0, act, LIST_EPS), OUTER_EPS, prec, nprob, stream));
    HIPCHK(fame_launch_mid(kmid(prec, p0, act, LIST_FWD), -1, prec, nprob, stream));
    ki.plan = plan(AX_INNER, prec).lists[pix(prec)][LIST_FWD];
    ki.y = y; ki.dotvec = dotvec; ki.partials = partials2;
    HIPCHK(fame_launch_inner(ki, 0, prec, nprob, stream));
    if (prec == 8) cnt_qbq64 += nprob; else cnt_qbq32 += nprob;
    return FAME_OK;
}

fame_status fame_ctx::qbq_residual32(const void* xr, const void* b, void* x32, int p0, int nprob, const int* act,
                                     int nact) {
    const size_t es = 8 * 2 * (size_t)N;   /* FP32 complex vector bytes per problem */
    KInner ki = kin(8, p0, act, LIST_INV);
    ki.x = xr; ki.pikind = 0;
    ki.fx32 = x32; ki.fp32 = (char*)cg32[1] + es * p0; ki.cgsc = cgsc;
    HIPCHK(fame_launch_inner(ki, 1, 8, nprob, stream));
    HIPCHK(fame_launch_mid(kmid(8, p0, act, LIST_INV), +1, 8, nprob, stream));
    HIPCHK(fame_launch_outer(kout(8, p0, act, LIST_EPS), OUTER_EPS, 8, nprob, stream));
    HIPCHK(fame_launch_mid(kmid(8, p0, act, LIST_FWD), -1, 8, nprob, stream));
    KInner kc = kin(8, p0, act, LIST_FWD);
    kc.x = xr; kc.y = const_cast<void*>(b); kc.pikind = 0;
    kc.rb = b; kc.x32 = x32;
    kc.r32 = (char*)cg32[0] + es * p0; kc.p32 = (char*)cg32[1] + es * p0;
    kc.rpart = partials;
    HIPCHK(fame_launch_inner(kc, 0, 8, nprob, stream));
    cnt_qbq64 += nact;
    return FAME_OK;
}

/* field (kernel layout, in F64) = Qr/Pr x  for problems p0.. */
fame_status fame_ctx::qr_field(const void* x, int p0, int nprob, int pikind, const int* act) {
    KInner ki = kin(8, p0, act, LIST_INV);
    ki.x = x; ki.pikind = pikind;
    HIPCHK(fame_launch_inner(ki, 1, 8, nprob, stream));
    HIPCHK(fame_launch_mid(kmid(8, p0, act, LIST_INV), +1, 8, nprob, stream));
    HIPCHK(fame_launch_outer(kout(8, p0, act, LIST_INV), OUTER_QR, 8, nprob, stream));
    return FAME_OK;
}
/* y (padded reduced) = Qrs/Prs of the field in F64 (kernel layout) */
fame_status fame_ctx::qrs_field(void* y, int p0, int nprob, int pikind, const int* act) {
    HIPCHK(fame_launch_outer(kout(8, p0, act, LIST_FWD), OUTER_QRS, 8, nprob, stream));
    HIPCHK(fame_launch_mid(kmid(8, p0, act, LIST_FWD), -1, 8, nprob, stream));
    KInner ki = kin(8, p0, act, LIST_FWD);
    ki.y = y; ki.pikind = pikind;
    HIPCHK(fame_launch_inner(ki, 0, 8, nprob, stream));
    return FAME_OK;
}

/* copy a field between the kernel layout (F64 problem slots) and the reference layout */
fame_status fame_ctx::field_to_ref(const void* fint, void* fref, int nprob, size_t dst_stride) {
    if (g.general) {
        HIPCHK(fame_launch_xz_transpose(fint, fref, g.nc, g.nb, g.na, +1, 3 * (size_t)N, 3, nprob, 8, stream));
        (void)dst_stride;
    } else {
        HIPCHK(hipMemcpyAsync(fref, fint, 16 * 3 * (size_t)N * nprob, hipMemcpyDeviceToDevice, stream));
    }
    return FAME_OK;
}
fame_status fame_ctx::field_from_ref(const void* fref, void* fint, int nprob) {
    if (g.general) {
        HIPCHK(fame_launch_xz_transpose(fref, fint, g.nc, g.nb, g.na, -1, 3 * (size_t)N, 3, nprob, 8, stream));
    } else {
        HIPCHK(hipMemcpyAsync(fint, fref, 16 * 3 * (size_t)N * nprob, hipMemcpyDeviceToDevice, stream));
    }
    return FAME_OK;
}

extern "C" fame_status fame_apply(fame_ctx* c, int op, int ik, const void* x, void* y) {
    if (!c || !x || !y) FAIL(FAME_ERR_ARG, "fame_apply: null argument");
    if (ik < 0 || ik >= c->nk) FAIL(FAME_ERR_ARG, "fame_apply: wave vector index out of range");
    if ((op == FAME_OP_QBQ || op == FAME_OP_INVAR || op == FAME_OP_INVAR_MIXED || op == FAME_OP_QBQ32) && !c->have_eps)
        FAIL(FAME_ERR_STATE, "fame_apply: permittivity not set");
    HIPCHK(hipSetDevice(c->device));
    const int N = c->N, Nd = c->nd[ik];
    double2* xin = (double2*)c->ref64[0];
    double2* yout = (double2*)c->ref64[1];
    switch (op) {
        case FAME_OP_QR:
        case FAME_OP_PR:
            HIPCHK(fame_launch_pack(x, xin, N, Nd, +1, 8, c->stream));
            CHK(c->qr_field(xin, ik, 1, op == FAME_OP_PR ? 1 : 0, nullptr));
            CHK(c->field_to_ref(c->F64 + (size_t)ik * 3 * N, y, 1, 3 * (size_t)N));
            break;
        case FAME_OP_QRS:
        case FAME_OP_PRS:
            CHK(c->field_from_ref(x, c->F64 + (size_t)ik * 3 * N, 1));
            CHK(c->qrs_field(yout, ik, 1, op == FAME_OP_PRS ? 1 : 0, nullptr));
            HIPCHK(fame_launch_pack(yout, y, N, Nd, -1, 8, c->stream));
            break;
        case FAME_OP_QBQ:
            HIPCHK(fame_launch_pack(x, xin, N, Nd, +1, 8, c->stream));
            /* the field slot used by qbq is that of problem ik */
            CHK(c->qbq(8, xin, yout, ik, 1, nullptr, nullptr));
            HIPCHK(fame_launch_pack(yout, y, N, Nd, -1, 8, c->stream));
            break;
        case FAME_OP_QBQ32: {
            /* FP32 QBQ (the inner operator of the mixed-precision solver), FP64 in/out */
            float2* x32 = (float2*)c->cg32[3] + (size_t)ik * 2 * N;
            float2* y32 = (float2*)c->cg32[4] + (size_t)ik * 2 * N;
            HIPCHK(fame_launch_pack(x, xin, N, Nd, +1, 8, c->stream));
            HIPCHK(fame_convert(x32, 4, xin, 8, 2 * (size_t)N, 0, 0, 1, c->stream));
            CHK(c->qbq(4, x32, y32, ik, 1, nullptr, nullptr));
            HIPCHK(fame_convert(yout, 8, y32, 4, 2 * (size_t)N, 0, 0, 1, c->stream));
            HIPCHK(fame_launch_pack(yout, y, N, Nd, -1, 8, c->stream));
            break;
        }
        case FAME_OP_INVAR:
        case FAME_OP_INVAR_MIXED: {
            HIPCHK(fame_launch_pack(x, xin, N, Nd, +1, 8, c->stream));
            int it = 0;
            fame_ls ls = c->ls_last.maxit > 0 ? c->ls_last : fame_ls{1000, 1e-12};
            CHK(c->invar(op == FAME_OP_INVAR ? FAME_PREC_F64 : FAME_PREC_F32_REFINE, xin, yout, ik, 1, nullptr, ls, &it));
            HIPCHK(fame_launch_pack(yout, y, N, Nd, -1, 8, c->stream));
            c->last_iters = it;
            break;
        }
        default: FAIL(FAME_ERR_ARG, "fame_apply: unknown operator");
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return FAME_OK;
}

/* ======================================================================== */
/*  Autotuner                                                                */
/* ======================================================================== */
/* For each pass kernel (inner, middle, outer) of precision prec, time every
 * feasible (engine, lines per block) on the current batch (all nk problems, the
 * shape of the production launches) and keep the fastest.  The field and
 * vector buffers are zeroed first; the results are discarded.               */
fame_status fame_ctx::tune(int prec) {
    const int pi = pix(prec);
    if (nk < 1) FAIL(FAME_ERR_STATE, "autotune: wave vectors must be set");
    const size_t es = prec == 8 ? 16 : 8;
    void* F = prec == 8 ? (void*)F64 : (void*)F32;
    void* x = prec == 8 ? cg64[1] : cg32[1];
    void* y = prec == 8 ? cg64[2] : cg32[2];
    HIPCHK(hipMemsetAsync(F, 0, es * 3 * (size_t)N * nk, stream));
    HIPCHK(hipMemsetAsync(x, 0, es * 2 * (size_t)N * nk, stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    static const int candL_inner[] = {2, 4, 6, 8, 10, 12, 16, 20, 24, 32};
    static const int candL_axis[] = {4, 8, 16, 24, 32, 48, 64};
    fame_status st = FAME_OK;
    for (int a = 0; a < 3 && st == FAME_OK; ++a) {
        PassTile best = tile[a][pi];
        float best_ms = 1e30f;
        for (int eng = 0; eng < 2; ++eng) {
            if (ax[a][eng].n == 0 || (eng_forced >= 0 && eng != eng_forced)) continue;
            const int* cl = a == AX_INNER ? candL_inner : candL_axis;
            const int ncl = a == AX_INNER ? 10 : 7;
            for (int ci = 0; ci < ncl; ++ci) {
                const int L = cl[ci];
                if (!tile_ok(ax[a][eng], a, prec, L, g)) continue;
                const PassTile save = tile[a][pi];
                tile[a][pi].eng = eng; tile[a][pi].L = L; tile[a][pi].P = ax[a][eng].M | 1;
                float ms = 0;
                for (int rep = -1; rep < 3; ++rep) {
                    if (rep == 0 && hipEventRecord(e0, stream) != hipSuccess) { st = FAME_ERR_HIP; break; }
                    hipError_t e = hipSuccess;
                    if (a == AX_INNER) {
                        KInner ki = kin(prec, 0, nullptr, LIST_INV);
                        ki.x = x; ki.pikind = 0;
                        e = fame_launch_inner(ki, 1, prec, nk, stream);
                        KInner kf = kin(prec, 0, nullptr, LIST_FWD);
                        kf.y = y; kf.pikind = 0;
                        if (e == hipSuccess) e = fame_launch_inner(kf, 0, prec, nk, stream);
                    } else if (a == AX_MID) {
                        e = fame_launch_mid(kmid(prec, 0, nullptr, LIST_INV), +1, prec, nk, stream);
                        if (e == hipSuccess) e = fame_launch_mid(kmid(prec, 0, nullptr, LIST_FWD), -1, prec, nk, stream);
                    } else {
                        e = fame_launch_outer(kout(prec, 0, nullptr, LIST_EPS), OUTER_EPS, prec, nk, stream);
                    }
                    if (e != hipSuccess) { st = FAME_ERR_HIP; break; }
                }
                if (st == FAME_OK && (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                      hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) st = FAME_ERR_HIP;
                if (st != FAME_OK) { tile[a][pi] = save; break; }
                if (ms < best_ms) { best_ms = ms; best = tile[a][pi]; }
                tile[a][pi] = save;
            }
            if (st != FAME_OK) break;
        }
        tile[a][pi] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != FAME_OK) FAIL(st, std::string("autotune: HIP error: ") + hipGetErrorString(hipGetLastError()));
    nblk_in[pi] = (g.nb * g.nc + tile[AX_INNER][pi].L - 1) / tile[AX_INNER][pi].L;
    tuned[pi] = true;
    drop_graphs();            /* captured launches carry the old tiles */
    return FAME_OK;
}

extern "C" fame_status fame_set_tiles(fame_ctx* c, int prec, const int* choice) {
    if (!c || !choice || (prec != 8 && prec != 4)) FAIL(FAME_ERR_ARG, "fame_set_tiles: bad arguments");
    const int pi = prec == 8 ? 0 : 1;
    for (int a = 0; a < 3; ++a) {
        const int eng = choice[2 * a], L = choice[2 * a + 1];
        if (eng < 0 || eng > 1 || c->ax[a][eng].n == 0 || !tile_ok(c->ax[a][eng], a, prec, L, c->g))
            FAIL(FAME_ERR_ARG, "fame_set_tiles: engine or lines per block not available for pass " + std::to_string(a));
    }
    for (int a = 0; a < 3; ++a) {
        PassTile& t = c->tile[a][pi];
        t.eng = choice[2 * a]; t.L = choice[2 * a + 1]; t.P = c->ax[a][t.eng].M | 1;
    }
    c->nblk_in[pi] = (c->g.nb * c->g.nc + c->tile[AX_INNER][pi].L - 1) / c->tile[AX_INNER][pi].L;
    c->tuned[pi] = true;
    c->drop_graphs();
    return FAME_OK;
}

extern "C" fame_status fame_autotune(fame_ctx* c, int prec, int* choice) {
    if (!c || (prec != 8 && prec != 4)) FAIL(FAME_ERR_ARG, "fame_autotune: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    CHK(c->tune(prec));
    if (choice) {
        const int pi = prec == 8 ? 0 : 1;
        for (int a = 0; a < 3; ++a) { choice[2 * a] = c->tile[a][pi].eng; choice[2 * a + 1] = c->tile[a][pi].L; }
    }
    return FAME_OK;
}

/* ======================================================================== */
/*  CG (batched)                                                             */
/* ======================================================================== */
/* Solve QBQ x = b for problems p0..p0+nprob-1 (padded layout, precision prec).
 * Stops per problem when <r,r> <= max(tol^2, reltol^2 <b,b>) or after maxit
 * (reference counter semantics, CG.cu:36-78).  iters[] (host) receives k.    */
fame_status fame_ctx::cg(int prec, const void* b, void* x, int p0, int nprob, double tol, double reltol, int maxit,
                         const int* mask_dev, int* iters, bool b64, bool defer_finish, int begun_nblk) {
    if (b64 && prec != 4) FAIL(FAME_ERR_ARG, "cg: an FP64 right-hand side needs prec 4");
    void** V = prec == 8 ? cg64 : cg32;
    const size_t n = 2 * (size_t)N, st = 2 * (size_t)N;
    const size_t es = prec == 8 ? 16 : 8;
    char* r = (char*)V[0] + es * st * p0;
    char* p = (char*)V[1] + es * st * p0;
    int nbp = NBLK_VEC;
    if (begun_nblk > 0) nbp = begun_nblk;
    else if (b64) HIPCHK(fame_cg_begin_mixed(x, r, p, b, partials, NBLK_VEC, n, st, nprob, mask_dev, stream));
    else HIPCHK(fame_cg_begin(x, r, p, b, partials, NBLK_VEC, n, st, nprob, prec, mask_dev, stream));
    HIPCHK(fame_cg_init_scal2(cgsc, partials, nbp, tol, reltol, maxit, nprob, active, mask_dev, stream));
    int done_it = 0;
    int chunk = std::max(1, std::min(maxit, cg_predict[pix(prec)] - 1));
    while (true) {
        for (int it = 0; it < chunk;) {
            const int k = std::min(chunk - it, CG_GRAPH_ITERS);
            CHK(cg_iteration(prec, x, p0, nprob, k));
            it += k;
        }
        done_it += chunk;
        /* the whole scalar state comes back, so that the last check also
         * holds the final iteration counts (no read-back after the loop) */
        HIPCHK(hipMemcpyAsync(sc_h, cgsc, sizeof(CgScal) * nprob, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        int any = 0;
        for (int i = 0; i < nprob; ++i) any |= sc_h[i].active;
        if (!any || done_it >= maxit) break;
        chunk = 2;
    }
    if (!defer_finish) HIPCHK(fame_cg_finish_x(x, p, cgsc, n, st, nprob, prec, stream));
    int mx = 0;
    long long applied = 0;   /* QBQ applications: iters - 1 per problem (masked ones make none) */
    for (int i = 0; i < nprob; ++i) {
        if (iters) iters[i] = sc_h[i].iters;
        mx = std::max(mx, sc_h[i].iters);
        applied += sc_h[i].iters - 1;
    }
    if (prec == 8) cnt_qbq64 += applied; else cnt_qbq32 += applied;
    cg_predict[pix(prec)] = std::max(2, mx - 1);
    return FAME_OK;
}

/* one CG iteration with Ap stored: QBQ p -> Ap whose first pass applies the
 * previous iteration's x, r and p updates and whose last pass sums <p,Ap> and
 * <Ap,Ap>, then alpha, beta and the flags (six kernels) */
fame_status fame_ctx::cg_step_ap(int prec, void* p, void* r, void* x, void* Ap, int p0, int nprob) {
    const long long c64 = cnt_qbq64, c32 = cnt_qbq32;
    CHK(qbq(prec, p, Ap, p0, nprob, active, p, r, x, Ap, true));
    cnt_qbq64 = c64; cnt_qbq32 = c32;
    HIPCHK(fame_cg_alpha_beta(cgsc, partials2, nblk_in[pix(prec)], nprob, active, stream));
    return FAME_OK;
}

/* one CG iteration with the residual updated by the last pass (seven kernels):
 * the first pass applies x += alpha p of the previous iteration and forms
 * p = r + beta p; the outer pass sums <p,Ap> from the real-space field; alpha;
 * the last pass stores r - alpha Ap (Ap itself is never written) and sums its
 * <r,r>; beta and the flags from that exact norm */
fame_status fame_ctx::cg_step_rc(int prec, void* p, void* r, void* x, int p0, int nprob, int npap) {
    KInner ki = kin(prec, p0, active, LIST_INV);
    ki.x = p; ki.pikind = 0;
    ki.cg_r = r; ki.cg_x = x; ki.cgsc = cgsc; ki.cg_ap = nullptr; ki.partials = partials2;
    HIPCHK(fame_launch_inner(ki, 1, prec, nprob, stream));
    HIPCHK(fame_launch_mid(kmid(prec, p0, active, LIST_INV), +1, prec, nprob, stream));
    KAxis ko = kout(prec, p0, active, LIST_EPS);
    ko.pap = pap_part;
    HIPCHK(fame_launch_outer(ko, OUTER_EPS, prec, nprob, stream));
    HIPCHK(fame_cg_alpha(cgsc, pap_part, npap, nprob, active, stream));
    HIPCHK(fame_launch_mid(kmid(prec, p0, active, LIST_FWD), -1, prec, nprob, stream));
    KInner kc = kin(prec, p0, active, LIST_FWD);
    kc.x = p; kc.y = r; kc.pikind = 0;
    kc.cg_r = r; kc.cgsc = cgsc; kc.partials = partials2;
    HIPCHK(fame_launch_inner(kc, 0, prec, nprob, stream));
    HIPCHK(fame_cg_beta(cgsc, partials2, nblk_in[pix(prec)], nprob, active, stream));
    return FAME_OK;
}

/* `count` CG iterations (CG.cu:53-78) on the masked batch, as cg_step_rc
 * when the outer pass can sum <p,Ap> (no Bluestein plan), else cg_step_ap.
 * The reference's alpha, beta and stopping rule are kept; the x and p updates
 * of iteration k are applied while iteration k+1 reads p anyway (k_cg_begin sets
 * p = r and beta = 0, so the first pass leaves p = r).  The last x update is
 * applied by k_cg_finish_x after the loop.  All scalars stay on the device, so
 * the launches are captured once per count as a graph.                      */
fame_status fame_ctx::cg_iteration(int prec, void* x, int p0, int nprob, int count) {
    void** V = prec == 8 ? cg64 : cg32;
    const size_t n = 2 * (size_t)N, st = 2 * (size_t)N;
    const size_t es = prec == 8 ? 16 : 8;
    char* r = (char*)V[0] + es * st * p0;
    char* p = (char*)V[1] + es * st * p0;
    char* Ap = (char*)V[2] + es * st * p0;
    auto launch = [&](hipStream_t s) -> fame_status {
        const hipStream_t keep = stream;
        stream = s;
        const int npap = pap_blocks(prec);
        for (int k = 0; k < count; ++k) {
            fame_status q = npap ? cg_step_rc(prec, p, r, x, p0, nprob, npap) : cg_step_ap(prec, p, r, x, Ap, p0, nprob);
            if (q != FAME_OK) { stream = keep; return q; }
        }
        stream = keep;
        return FAME_OK;
    };
    if (!use_graphs) return launch(stream);
    const auto key = std::make_tuple(prec, p0, nprob, (const void*)x, count);
    auto itg = cg_graphs.find(key);
    if (itg == cg_graphs.end()) {
        if (!cap_stream) HIPCHK(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
        fame_status q = launch(cap_stream);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(cap_stream, &gr);
        if (q != FAME_OK) { if (gr) (void)hipGraphDestroy(gr); return q; }
        HIPCHK(e);
        hipGraphExec_t ex = nullptr;
        e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        HIPCHK(e);
        itg = cg_graphs.emplace(key, ex).first;
    }
    HIPCHK(hipGraphLaunch(itg->second, stream));
    return FAME_OK;
}

extern "C" fame_status fame_cg(fame_ctx* c, int ik, const void* rhs, void* x, int maxit, double tol, int* iters) {
    if (!c || !rhs || !x) FAIL(FAME_ERR_ARG, "fame_cg: null argument");
    if (ik < 0 || ik >= c->nk) FAIL(FAME_ERR_ARG, "fame_cg: wave vector index out of range");
    if (!c->have_eps) FAIL(FAME_ERR_STATE, "fame_cg: permittivity not set");
    HIPCHK(hipSetDevice(c->device));
    const int N = c->N, Nd = c->nd[ik];
    double2* b = (double2*)c->ref64[0] + (size_t)ik * 2 * N;
    double2* y = (double2*)c->ref64[1] + (size_t)ik * 2 * N;
    HIPCHK(fame_launch_pack(rhs, b, N, Nd, +1, 8, c->stream));
    int it = 0;
    CHK(c->cg(8, b, y, ik, 1, tol, 0.0, maxit, nullptr, &it));
    HIPCHK(fame_launch_pack(y, x, N, Nd, -1, 8, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (iters) *iters = it;
    return FAME_OK;
}

/* ======================================================================== */
/*  invAr (batched): y = Sigma^-1 QBQ^-1 Sigma^-1 x                          */
/* ======================================================================== */
fame_status fame_ctx::invar(int mode, const void* x, void* y, int p0, int nprob, const int* mask_dev, fame_ls ls,
                            int* iters, size_t xstride, void* out, size_t ostride, const int* mask_host) {
    const size_t n = 2 * (size_t)N, st = 2 * (size_t)N;
    double2* b = (double2*)ref64[2] + st * p0;
    HIPCHK(fame_scale_real_mod2(b, st, x, xstride ? xstride : st, sig_inv + (size_t)p0 * N, n, N, N, nprob, 8, stream));
    std::vector<int> it(nprob, 0);
    if (mode == FAME_PREC_F64) {
        CHK(cg(8, b, y, p0, nprob, ls.tol, 0.0, ls.maxit, mask_dev, it.data()));
    } else {
        /* FP64 iterative refinement around FP32 CG solves: the FP64 residual
         * pass adds each FP32 solution into the FP64 iterate on its first pass
         * and writes the next FP32 solve's start (r = p = FP32(b - QBQ64 x),
         * x = 0) with its norm on its last, so no FP64 residual vector is kept */
        double2* xr = (double2*)y;
        float2* d32 = (float2*)cg32[4] + st * p0;
        HIPCHK(hipMemsetAsync(xr, 0, 16 * st * nprob, stream));
        /* the first FP32 solve starts from b (x = 0, r = p = FP32(b)) */
        HIPCHK(fame_cg_begin_mixed(d32, (char*)cg32[0] + 8 * st * p0, (char*)cg32[1] + 8 * st * p0, b, partials,
                                   NBLK_VEC, n, st, nprob, nullptr, stream));
        int nbp = NBLK_VEC;
        std::vector<int> msk(nprob, 1);
        if (mask_host) {
            std::copy(mask_host, mask_host + nprob, msk.begin());
        } else if (mask_dev) {
            HIPCHK(hipMemcpyAsync(msk.data(), mask_dev, sizeof(int) * nprob, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        double tol2 = ls.tol * ls.tol;
        for (int sweep = 0; sweep < refine_max_sweeps; ++sweep) {
            HIPCHK(fame_reduce_partials(partials, nbp, 1, scal, nprob, stream));
            std::vector<double> rn(nprob);
            HIPCHK(hipMemcpyAsync(rn.data(), scal, sizeof(double) * nprob, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            int any = 0;
            for (int i = 0; i < nprob; ++i) { msk[i] = msk[i] && (rn[i] > tol2); any |= msk[i]; }
            if (!any) break;
            HIPCHK(hipMemcpyAsync(mask, msk.data(), sizeof(int) * nprob, hipMemcpyHostToDevice, stream));
            std::vector<int> it32(nprob, 0);
            CHK(cg(4, b, d32, p0, nprob, ls.tol, refine_reltol, ls.maxit, mask, it32.data(), true, true, nbp));
            for (int i = 0; i < nprob; ++i) if (msk[i]) it[i] += it32[i] - 1;
            /* the next sweep's start from r = b - QBQ64 xr, with its norm */
            CHK(qbq_residual32(xr, b, d32, p0, nprob, mask, (int)std::count(msk.begin(), msk.end(), 1)));
            nbp = nblk_in[pix(8)];
        }
        for (int i = 0; i < nprob; ++i) it[i] += 1;
    }
    if (out) HIPCHK(fame_scale_real_mod2(out, ostride, y, st, sig_inv + (size_t)p0 * N, n, N, N, nprob, 8, stream, mask_dev));
    else HIPCHK(fame_scale_real_mod(y, y, sig_inv + (size_t)p0 * N, n, st, N, N, nprob, 8, stream));
    if (iters) for (int i = 0; i < nprob; ++i) iters[i] = it[i];
    return FAME_OK;
}
