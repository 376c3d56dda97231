/*
 * fame_lanczos.hip -- batched restarted Lanczos on A_r^{-1} (host orchestration).
 *
 * Follows Lanczos_Isotropic.cu:7-96 / Lanczos_decomp_gpu.cu:7-91 /
 * Lanczos_LockPurge_gpu.cu:7-48 of the reference, re-planned for one GPU and
 * many wave vectors at once:
 *  - all problems of the context advance in lock-step; a per-problem mask
 *    freezes converged ones;
 *  - alpha, beta, norms and reorthogonalisation coefficients never leave the
 *    device; T (tridiagonal) comes back once per restart;
 *  - full reorthogonalisation is classical Gram-Schmidt applied twice (the
 *    reference applies one modified Gram-Schmidt pass, one dot+axpy per column);
 *  - the exact-shift lock/purge restart is accumulated on the host as an
 *    (nstep+1)x(nstep+1) rotation matrix Q, using the Gram matrix G = U^H U for
 *    the normalisations, and applied to the basis as one U*Q product instead
 *    of sum(nstep-j) column rotations;
 *  - the tridiagonal eigensolver (LAPACKE_zpteqr in the reference) is an own
 *    cyclic-Jacobi routine, the shifted QR step an own banded explicit step
 *    (fame_tridiag.cpp).
 */
#include "fame_hip.h"
#include "fame_host.h"
#include "fame_kernels.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <thread>

using namespace fame;

extern "C" const char* fame_last_error(void);
std::string& fame_err_ref();

#define FAIL(code, msg) do { fame_err_ref() = (msg); return (code); } while (0)
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    fame_err_ref() = std::string("HIP error: ") + hipGetErrorString(e_) + " at " + __FILE__ + ":" + std::to_string(__LINE__); \
    return FAME_ERR_HIP; } } while (0)
#define CHK(x) do { fame_status s_ = (x); if (s_ != FAME_OK) return s_; } while (0)


/* partial-sum blocks per problem: 256 for single-column products (the
 * three-term step, norms), 1024 for the Gram-Schmidt products over many
 * columns, which then keep more loads in flight per CU */
#define GEMV_BLK 256
#define GEMV_BLK_GS 1024
/* Gram blocks per problem: enough to fill the GPU for small batches */
static int gram_blk(int nk) { return std::max(32, std::min(512, 2048 / std::max(nk, 1))); }

namespace {

struct LzBuffers {
    size_t n2 = 0;          /* 2N */
    int nk = 0, nstep = 0, nwant = 0;
    double2* U = nullptr;   /* nk x (nstep+1) x 2N */
    double2* W = nullptr;   /* nk x max(nwant+1, nwant) x 2N */
    double* T0d = nullptr;  /* nk x (nstep+1) */
    double* T1d = nullptr;
    double* hdot = nullptr; /* nk x 2 */
    double* h = nullptr;    /* nk x 4 */
    double* hh = nullptr;   /* nk x 2(nstep+1) */
    double* nrm = nullptr;  /* nk */
    double* invb = nullptr; /* nk */
    double* gpart = nullptr;/* nk x GEMV_BLK_GS x 2(nstep+1) ; also gram partials */
    double* G = nullptr;    /* nk x 2 m^2 */
    double* Q = nullptr;    /* nk x m x (nwant+1) */
    int* mask = nullptr;    /* nk */
    int* mask2 = nullptr;   /* nk: problems that take the second Gram-Schmidt pass */
    std::vector<double> ls_time;   /* nk: seconds inside invAr, each problem's share */
};

fame_status lz_alloc(fame_ctx* c, LzBuffers& b, int nk, int nstep, int nwant) {
    b.n2 = 2 * (size_t)c->N; b.nk = nk; b.nstep = nstep; b.nwant = nwant;
    b.ls_time.assign(nk, 0.0);
    const size_t m = nstep + 1;
    const size_t gp = std::max((size_t)GEMV_BLK_GS * 2 * m, (size_t)gram_blk(c->gram_batch > 0 ? c->gram_batch : nk) * 2 * m * m);
    void** ptr[] = {(void**)&b.U, (void**)&b.W, (void**)&b.T0d, (void**)&b.T1d, (void**)&b.hdot, (void**)&b.h,
                    (void**)&b.hh, (void**)&b.nrm, (void**)&b.invb, (void**)&b.gpart, (void**)&b.G, (void**)&b.Q,
                    (void**)&b.mask, (void**)&b.mask2};
    const std::vector<size_t> bytes = {
        16 * b.n2 * m * nk, 16 * b.n2 * (nwant + 1) * nk, sizeof(double) * m * nk, sizeof(double) * m * nk,
        sizeof(double) * 2 * nk, sizeof(double) * 4 * nk, sizeof(double) * 2 * m * nk, sizeof(double) * nk,
        sizeof(double) * nk, sizeof(double) * gp * nk, sizeof(double) * 2 * m * m * nk,
        sizeof(double) * m * (nwant + 1) * nk, sizeof(int) * nk, sizeof(int) * nk};
    /* the same shape as the last solve: keep its buffers (freeing and
     * allocating the basis, gigabytes, costs tens of ms per solve) */
    if (c->big_bytes == bytes && c->big.size() == bytes.size()) {
        for (size_t i = 0; i < bytes.size(); ++i) *ptr[i] = c->big[i];
        return FAME_OK;
    }
    for (void* p : c->big) if (p) (void)hipFree(p);
    c->big.clear();
    c->big_bytes.clear();
    for (size_t i = 0; i < bytes.size(); ++i) {
        hipError_t e = hipMalloc(ptr[i], bytes[i] ? bytes[i] : 16);
        if (e != hipSuccess) {
            *ptr[i] = nullptr;
            for (void* p : c->big) if (p) (void)hipFree(p);
            c->big.clear();
            FAIL(FAME_ERR_NOMEM, std::string("Lanczos allocation of ") + std::to_string(bytes[i]) + " bytes failed");
        }
        c->big.push_back(*ptr[i]);
    }
    c->big_bytes = bytes;
    return FAME_OK;
}

}  // namespace

/* column j of problem 0 (problem stride is (nstep+1)*2N) */
static inline double2* col(const LzBuffers& b, int j) { return b.U + (size_t)j * b.n2; }

/* U_{j+1} = invAr(U_j) for the masked problems, through contiguous staging */
static fame_status lz_invar(fame_ctx* c, LzBuffers& b, int j, int mode, const fame_ls& ls, const int* mask_dev,
                            std::vector<int>& ls_iter, std::vector<int>& es_iter, const std::vector<int>& mask_h) {
    const size_t ps = b.n2 * (b.nstep + 1);
    double2* ysolve = (double2*)c->ref64[0];
    std::vector<int> it(b.nk, 0);
    /* read straight from column j; the result goes to column j + 1 of the
     * problems still iterating only: a converged problem keeps its basis,
     * whose columns its final Ritz vectors are built from.  The solve ends
     * with a host read of the CG state, so its host time is its wall time. */
    const auto t0 = std::chrono::steady_clock::now();
    CHK(c->invar(mode, col(b, j), ysolve, 0, b.nk, mask_dev, ls, it.data(), ps, col(b, j + 1), ps, mask_h.data()));
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int active = 0;
    for (int p = 0; p < b.nk; ++p) active += mask_h[p] ? 1 : 0;
    for (int p = 0; p < b.nk; ++p)
        if (mask_h[p]) { ls_iter[p] += it[p]; es_iter[p] += 1; b.ls_time[p] += dt / active; }
    return FAME_OK;
}

/* alpha_j, subtract the three-term recurrence, beta_j, normalise (Lanczos_decomp_gpu.cu:161-174) */
static fame_status lz_three_term(fame_ctx* c, LzBuffers& b, int j, const int* mask_dev) {
    const size_t ps = b.n2 * (b.nstep + 1);
    const int ld = b.nstep + 1;
    HIPCHK(fame_gemv_h(col(b, j), col(b, j + 1), b.gpart, GEMV_BLK, 1, b.n2, b.n2, ps, ps, b.nk, mask_dev, c->stream));
    HIPCHK(fame_reduce_partials(b.gpart, GEMV_BLK, 2, b.hdot, b.nk, c->stream));
    HIPCHK(fame_lz_coeffs(b.hdot, b.T0d, b.T1d, ld, j, b.h, b.nk, mask_dev, c->stream));
    if (j >= 1) HIPCHK(fame_gemv_n(col(b, j - 1), col(b, j + 1), b.h, 2, b.n2, b.n2, ps, ps, 4, b.nk, mask_dev, c->stream));
    else HIPCHK(fame_gemv_n(col(b, j), col(b, j + 1), b.h, 1, b.n2, b.n2, ps, ps, 4, b.nk, mask_dev, c->stream));
    HIPCHK(fame_norm2_partials(col(b, j + 1), 8, b.gpart, GEMV_BLK, b.n2, ps, b.nk, c->stream));
    HIPCHK(fame_reduce_partials(b.gpart, GEMV_BLK, 1, b.nrm, b.nk, c->stream));
    HIPCHK(fame_lz_beta(b.nrm, b.T1d, ld, j, b.invb, b.nk, mask_dev, c->stream));
    HIPCHK(fame_scale_real_mod(col(b, j + 1), col(b, j + 1), b.invb, b.n2, ps, 1, 1, b.nk, 8, c->stream, mask_dev));
    return FAME_OK;
}

/* classical Gram-Schmidt against U_0..U_j (Lanczos_decomp_gpu.cu:81-87), with a
 * second pass for the problems whose first-pass coefficients h = U^H w are not
 * negligible (||h|| > 1e-8 for the unit vector w): after the three-term step h
 * is normally O(machine eps), and a second pass only pays once orthogonality
 * has started to degrade, which is when "twice is enough" is needed. */
static fame_status lz_reorth(fame_ctx* c, LzBuffers& b, int j, const int* mask_dev) {
    const size_t ps = b.n2 * (b.nstep + 1);
    const int m = j + 1;
    const int* mk = mask_dev;
    for (int rep = 0; rep < 2; ++rep) {
        HIPCHK(fame_gemv_h(col(b, 0), col(b, j + 1), b.gpart, GEMV_BLK_GS, m, b.n2, b.n2, ps, ps, b.nk, mk, c->stream));
        HIPCHK(fame_reduce_partials(b.gpart, GEMV_BLK_GS, 2 * m, b.hh, b.nk, c->stream));
        HIPCHK(fame_gemv_n(col(b, 0), col(b, j + 1), b.hh, m, b.n2, b.n2, ps, ps, 2 * m, b.nk, mk, c->stream));
        if (rep == 0) {
            HIPCHK(fame_lz_second_pass_mask(b.hh, 2 * m, 1e-16, mask_dev, b.mask2, b.nk, c->stream));
            mk = b.mask2;
        }
    }
    return FAME_OK;
}

static fame_status lz_decomp(fame_ctx* c, LzBuffers& b, bool init, int mode, const fame_ls& ls,
                             const std::vector<int>& mask_h, std::vector<int>& ls_iter, std::vector<int>& es_iter) {
    HIPCHK(hipMemcpyAsync(b.mask, mask_h.data(), sizeof(int) * b.nk, hipMemcpyHostToDevice, c->stream));
    int start = b.nwant;
    if (init) {
        /* U_0 = e_n (Lanczos_init_kernel, Lanczos_decomp_gpu.cu:93-124) */
        const size_t ps = b.n2 * (b.nstep + 1);
        const double one[2] = {1.0, 0.0};
        for (int p = 0; p < b.nk; ++p) {
            if (!mask_h[p]) continue;
            HIPCHK(hipMemsetAsync(b.U + p * ps, 0, 16 * b.n2, c->stream));
            HIPCHK(hipMemcpyAsync(b.U + p * ps + (b.n2 - 1), one, 16, hipMemcpyHostToDevice, c->stream));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        CHK(lz_invar(c, b, 0, mode, ls, b.mask, ls_iter, es_iter, mask_h));
        CHK(lz_three_term(c, b, 0, b.mask));
        start = 1;
    }
    for (int j = start; j < b.nstep; ++j) {
        CHK(lz_invar(c, b, j, mode, ls, b.mask, ls_iter, es_iter, mask_h));
        CHK(lz_three_term(c, b, j, b.mask));
        CHK(lz_reorth(c, b, j, b.mask));
    }
    return FAME_OK;
}

extern "C" fame_status fame_eigs_impl(fame_ctx* c, int precision, const fame_es* es, const fame_ls* ls, double* freq,
                                      void* evr_dev, void* E_dev, fame_profile* prof);

/* Inverse iteration on the converged Ritz pairs, in lambda space: for each
 * band y (unit norm, all problems of the batch at once)
 *   w = A_r^{-1} y,  lambda = Re <w, y> / <w, w>,  y <- w / ||w||.
 * Since A_r w = y, <w, y> = w^H A_r w, so lambda is w's Rayleigh quotient and
 * needs no extra matrix apply.  A component of y along an eigenvector of
 * eigenvalue lambda_i shrinks by lambda / lambda_i, so the high-frequency
 * components that the theta-space stopping rule leaves (they dominate the
 * field residual) are removed first; inside a cluster the ratio is ~1, so the
 * bands keep their order and identity over a few sweeps.  The solves use an
 * absolute CG tolerance of ritz_tol (well below es.tol, since the solve's
 * error enters the new vector's lambda-space residual amplified by |Lambda|).
 * Column 0 of each problem's Lanczos basis is free at this point and holds w. */
static fame_status refine_ritz(fame_ctx* c, LzBuffers& b, int precision, double* freq, std::vector<int>& ls_iter,
                               std::vector<int>& es_iter) {
    const int nk = b.nk, nwant = b.nwant;
    const size_t n2 = b.n2, ps = n2 * (b.nstep + 1), wps = n2 * (nwant + 1);
    const fame_ls ls = {c->ritz_maxit, c->ritz_tol};
    double2* ysolve = (double2*)c->ref64[0];
    std::vector<int> it(nk, 0), all(nk, 1);
    std::vector<double> hv(2 * (size_t)nk), nv(nk), sc(nk);
    for (int sweep = 0; sweep < c->ritz_sweeps; ++sweep) {
        for (int i = 0; i < nwant; ++i) {
            double2* y = b.W + (size_t)i * n2;
            std::fill(it.begin(), it.end(), 0);
            const auto t0 = std::chrono::steady_clock::now();
            CHK(c->invar(precision, y, ysolve, 0, nk, nullptr, ls, it.data(), wps, b.U, ps, all.data()));
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            for (int p = 0; p < nk; ++p) { ls_iter[p] += it[p]; es_iter[p] += 1; b.ls_time[p] += dt / nk; }
            HIPCHK(fame_gemv_h(b.U, y, b.gpart, GEMV_BLK, 1, n2, n2, ps, wps, nk, nullptr, c->stream));
            HIPCHK(fame_reduce_partials(b.gpart, GEMV_BLK, 2, b.hdot, nk, c->stream));
            HIPCHK(fame_norm2_partials(b.U, 8, b.gpart, GEMV_BLK, n2, ps, nk, c->stream));
            HIPCHK(fame_reduce_partials(b.gpart, GEMV_BLK, 1, b.nrm, nk, c->stream));
            HIPCHK(hipMemcpyAsync(hv.data(), b.hdot, sizeof(double) * 2 * nk, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(nv.data(), b.nrm, sizeof(double) * nk, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int p = 0; p < nk; ++p) {
                if (!(nv[p] > 0)) FAIL(FAME_ERR_STATE, "fame_eigs: Ritz refinement produced a zero vector");
                freq[(size_t)p * nwant + i] = std::sqrt(std::max(hv[2 * p] / nv[p], 0.0));
                sc[p] = 1.0 / std::sqrt(nv[p]);
            }
            HIPCHK(hipMemcpyAsync(b.nrm, sc.data(), sizeof(double) * nk, hipMemcpyHostToDevice, c->stream));
            HIPCHK(fame_scale_real_mod2(y, wps, b.U, ps, b.nrm, n2, 1, 1, nk, 8, c->stream));
        }
    }
    return FAME_OK;
}

extern "C" fame_status fame_set_refinement(fame_ctx* c, int sweeps, double tol, int maxit) {
    if (!c || sweeps < 0 || sweeps > 16 || !(tol > 0) || maxit < 1)
        FAIL(FAME_ERR_ARG, "fame_set_refinement: need 0 <= sweeps <= 16, tol > 0, maxit >= 1");
    c->ritz_sweeps = sweeps; c->ritz_tol = tol; c->ritz_maxit = maxit;
    return FAME_OK;
}

/* Splitting a batch over concurrent child contexts.  The problems of one batch
 * iterate in lockstep: each CG solve ends with launches in which only the
 * slowest few problems still iterate, and the Lanczos vector work and host
 * checks leave the passes' share of the GPU idle meanwhile.  Two halves of the
 * batch, each a context on its own stream, fill those gaps with each other's
 * passes (FCC 96 on one MI355X: 4.72 against 4.48 k-pts/s for 30 wave vectors,
 * 4.21 against 3.88 for 8, 3.64 against 3.20 for 2; four groups were slower
 * than two at 4, 8 and 16 wave vectors).
 * Every problem's arithmetic is independent of the batch it runs in, and the
 * children take the parent's tiles and Gram blocking, so the results are
 * bit-identical to one batch.                                              */
static int split_groups(const fame_ctx* c) {
    const int g = c->groups > 0 ? c->groups : 2;
    return std::max(1, std::min(g, c->nk));
}

/* group i of G takes the wave vectors i, i + G, i + 2G, ...: neighbouring
 * points of a path cost about the same, so the groups finish together */
static int kid_count(int nk, int G, int i) { return (nk - i + G - 1) / G; }

/* children for G groups, in step with the parent; false (and no children)
 * when the device lacks the memory for them next to the solve itself */
static fame_status sync_kids(fame_ctx* c, int G, int nstep, int nwant, bool* ok) {
    *ok = false;
    bool remake = (int)c->kids.size() != G;
    for (int i = 0; !remake && i < G; ++i) remake = c->kids[i]->maxb < kid_count(c->nk, G, i);
    if (remake) {
        for (fame_ctx* k : c->kids) (void)fame_destroy(k);
        c->kids.clear();
        c->kid_nk = 0;
        /* a child holds the per-problem work vectors of its wave vectors
         * (about 400 N bytes each) and its own permittivity; the Lanczos
         * buffers (U, W) move from the parent to the children */
        size_t freeb = 0, totalb = 0;
        HIPCHK(hipMemGetInfo(&freeb, &totalb));
        const size_t N = c->N, n2 = 2 * N;
        const size_t lz = 16 * n2 * ((size_t)nstep + 1 + (size_t)nwant + 1) * (size_t)c->nk;
        const size_t need = 400 * N * (size_t)c->nk + 100 * N * (size_t)G + lz + ((size_t)1 << 30);
        if (freeb < need) return FAME_OK;
        for (int i = 0; i < G; ++i) {
            fame_ctx* k = nullptr;
            const fame_status st = fame_create(&c->lat, kid_count(c->nk, G, i), c->device, &k);
            if (st != FAME_OK) {
                for (fame_ctx* q : c->kids) (void)fame_destroy(q);
                c->kids.clear();
                return st == FAME_ERR_NOMEM || st == FAME_ERR_HIP ? FAME_OK : st;
            }
            c->kids.push_back(k);
            /* its own stream for good (a new stream would drop its CG graphs) */
            hipStream_t s = nullptr;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
                for (fame_ctx* q : c->kids) (void)fame_destroy(q);
                c->kids.clear();
                return FAME_OK;
            }
            k->stream = s;
            k->owns_stream = true;
        }
        c->kids_eps_ver = c->kids_k_ver = -1;
    }
    if (c->kids_eps_ver != c->eps_ver) {
        std::vector<double> e(3 * (size_t)c->N);
        HIPCHK(hipMemcpyAsync(e.data(), c->eps_ref, sizeof(double) * e.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (fame_ctx* k : c->kids) CHK(fame_set_eps(k, e.data()));
        c->kids_eps_ver = c->eps_ver;
    }
    if (c->kids_k_ver != c->k_ver || c->kid_nk != c->nk) {
        for (int i = 0; i < G; ++i) {
            std::vector<double> kv;
            for (int p = i; p < c->nk; p += G) kv.insert(kv.end(), c->kvec.begin() + 3 * p, c->kvec.begin() + 3 * p + 3);
            CHK(fame_set_wave_vectors(c->kids[i], kid_count(c->nk, G, i), kv.data(), nullptr));
        }
        c->kid_nk = c->nk;
        c->kids_k_ver = c->k_ver;
    }
    for (fame_ctx* k : c->kids) {
        /* the parent's plans: a child built under another FAME_FFT setting may
         * lack the matrix-core plan that the parent's tiles name */
        k->eng_forced = c->eng_forced;
        for (int a = 0; a < 3; ++a)
            if (c->ax[a][1].n > 0 && k->ax[a][1].n == 0) CHK(k->ax[a][1].build(c->ax[a][1].n, FFT_ENGINE_MFMA));
        bool same = true;
        for (int pi = 0; pi < 2; ++pi)
            for (int a = 0; a < 3; ++a)
                same = same && k->tile[a][pi].eng == c->tile[a][pi].eng && k->tile[a][pi].L == c->tile[a][pi].L;
        if (!same) {
            for (int pi = 0; pi < 2; ++pi) {
                for (int a = 0; a < 3; ++a) k->tile[a][pi] = c->tile[a][pi];
                k->nblk_in[pi] = c->nblk_in[pi];
            }
            k->drop_graphs();
        }
        k->tuned[0] = k->tuned[1] = true;    /* the parent's choice, tuned or fixed */
        k->groups = 1;
        k->gram_batch = c->nk;
        k->out_pmul = G;
        k->use_graphs = c->use_graphs;
        k->ritz_sweeps = c->ritz_sweeps; k->ritz_tol = c->ritz_tol; k->ritz_maxit = c->ritz_maxit;
        k->refine_max_sweeps = c->refine_max_sweeps; k->refine_reltol = c->refine_reltol;
    }
    *ok = true;
    return FAME_OK;
}

extern "C" fame_status fame_eigs(fame_ctx* c, int precision, const fame_es* es, const fame_ls* ls, double* freq,
                                 void* evr_dev, void* E_dev, fame_profile* prof) {
    if (!c || !es || !ls || !freq) FAIL(FAME_ERR_ARG, "fame_eigs: null argument");
    const int G = (c->have_eps && c->nk >= 1) ? split_groups(c) : 1;
    /* one batch, or arguments that fame_eigs_impl rejects with its own message */
    if (G <= 1 || es->nwant < 1 || es->nstep <= es->nwant || es->nstep + 1 > FAME_LZ_MAXCOL || es->maxit < 1 ||
        ls->maxit < 1 || (precision != FAME_PREC_F64 && precision != FAME_PREC_F32_REFINE))
        return fame_eigs_impl(c, precision, es, ls, freq, evr_dev, E_dev, prof);
    HIPCHK(hipSetDevice(c->device));
    /* the parent's tiles, tuned on the whole batch as a single solve would */
    if (c->autotune) {
        if (!c->tuned[0]) CHK(c->tune(8));
        if (precision == FAME_PREC_F32_REFINE && !c->tuned[1]) CHK(c->tune(4));
    }
    /* the Lanczos buffers of an earlier single-batch solve are not needed by
     * the children (freed before their memory is checked) */
    HIPCHK(hipStreamSynchronize(c->stream));
    for (void* p : c->big) if (p) (void)hipFree(p);
    c->big.clear();
    c->big_bytes.clear();
    bool ok = false;
    CHK(sync_kids(c, G, es->nstep, es->nwant, &ok));
    if (!ok) return fame_eigs_impl(c, precision, es, ls, freq, evr_dev, E_dev, prof);
    HIPCHK(hipStreamSynchronize(c->stream));   /* work the caller queued on the parent */
    c->ls_last = *ls;
    const int nwant = es->nwant;
    const size_t n2 = 2 * (size_t)c->N, n3 = 3 * (size_t)c->N;
    std::vector<int> precs(G, precision);
    std::vector<fame_es> ess(G, *es);
    std::vector<fame_ls> lss(G, *ls);
    /* frequencies and profiles come back per group and are dealt out below; the
     * device outputs are written in place, each child with a problem stride of
     * G problems (out_pmul) from its first problem's slot */
    std::vector<std::vector<double>> fk(G);
    std::vector<std::vector<fame_profile>> pk(G);
    std::vector<double*> fr(G);
    std::vector<void*> ev(G), Ev(G);
    std::vector<fame_profile*> pr(G);
    long long q64 = 0, q32 = 0;
    for (int i = 0; i < G; ++i) {
        const int n = kid_count(c->nk, G, i);
        fk[i].assign((size_t)n * nwant, 0.0);
        pk[i].assign(n, fame_profile());
        fr[i] = fk[i].data();
        pr[i] = pk[i].data();
        ev[i] = evr_dev ? (void*)((double2*)evr_dev + (size_t)i * nwant * n2) : nullptr;
        Ev[i] = E_dev ? (void*)((double2*)E_dev + (size_t)i * nwant * n3) : nullptr;
        q64 -= c->kids[i]->cnt_qbq64; q32 -= c->kids[i]->cnt_qbq32;
    }
    CHK(fame_eigs_many(G, c->kids.data(), precs.data(), ess.data(), lss.data(), fr.data(), evr_dev ? ev.data() : nullptr,
                       E_dev ? Ev.data() : nullptr, pr.data()));
    /* the children's streams are their own: their fields are complete when this returns */
    for (fame_ctx* k : c->kids) HIPCHK(hipStreamSynchronize(k->stream));
    for (int i = 0; i < G; ++i)
        for (int q = 0, p = i; p < c->nk; ++q, p += G) {
            std::memcpy(freq + (size_t)p * nwant, fk[i].data() + (size_t)q * nwant, sizeof(double) * nwant);
            if (prof) prof[p] = pk[i][q];
        }
    for (fame_ctx* k : c->kids) { q64 += k->cnt_qbq64; q32 += k->cnt_qbq32; }
    c->cnt_qbq64 += q64; c->cnt_qbq32 += q32;
    return FAME_OK;
}

extern "C" fame_status fame_set_groups(fame_ctx* c, int groups) {
    if (!c || groups < 0) FAIL(FAME_ERR_ARG, "fame_set_groups: need groups >= 0 (0: automatic)");
    c->groups = groups;
    return FAME_OK;
}

/* Several structures (contexts of any grid, batch and band count) solved in
 * one call: each context runs fame_eigs from its own host thread on its own
 * non-blocking stream (created here when the context has none), so that the
 * kernels of different structures overlap on the GPU. */
extern "C" fame_status fame_eigs_many(int n, fame_ctx* const* ctxs, const int* precision, const fame_es* es,
                                      const fame_ls* ls, double* const* freq, void* const* evr_dev,
                                      void* const* E_dev, fame_profile* const* prof) {
    if (n < 1 || !ctxs || !precision || !es || !ls || !freq) FAIL(FAME_ERR_ARG, "fame_eigs_many: bad arguments");
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j)
            if (ctxs[i] == ctxs[j] || !ctxs[i]) FAIL(FAME_ERR_ARG, "fame_eigs_many: contexts must be distinct and non-null");
    std::vector<fame_status> st(n, FAME_OK);
    std::vector<std::string> err(n);
    std::vector<std::thread> th;
    th.reserve(n);
    for (int i = 0; i < n; ++i) {
        th.emplace_back([&, i]() {
            fame_ctx* c = ctxs[i];
            hipStream_t own = nullptr;
            if (!c->stream) {
                if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&own, hipStreamNonBlocking) != hipSuccess) {
                    st[i] = FAME_ERR_HIP; err[i] = "stream creation failed"; return;
                }
                fame_set_stream(c, own);
            }
            st[i] = fame_eigs(c, precision[i], &es[i], &ls[i], freq[i], evr_dev ? evr_dev[i] : nullptr,
                              E_dev ? E_dev[i] : nullptr, prof ? prof[i] : nullptr);
            if (st[i] != FAME_OK) err[i] = fame_last_error();
            if (own) {
                (void)hipStreamSynchronize(own);
                fame_set_stream(c, nullptr);
                (void)hipStreamDestroy(own);
            }
        });
    }
    for (auto& t : th) t.join();
    for (int i = 0; i < n; ++i)
        if (st[i] != FAME_OK) FAIL(st[i], "fame_eigs_many: structure " + std::to_string(i) + ": " + err[i]);
    return FAME_OK;
}

fame_status fame_restore_impl(fame_ctx* c, int nwant, const double2* ev_padded, size_t ev_pstride, void* E_dev);

extern "C" fame_status fame_eigs_impl(fame_ctx* c, int precision, const fame_es* es, const fame_ls* ls, double* freq,
                                      void* evr_dev, void* E_dev, fame_profile* prof) {
    if (!c || !es || !ls || !freq) FAIL(FAME_ERR_ARG, "fame_eigs: null argument");
    if (!c->have_eps || c->nk < 1) FAIL(FAME_ERR_STATE, "fame_eigs: permittivity and wave vectors must be set first");
    if (es->nwant < 1 || es->nstep <= es->nwant || es->nstep + 1 > FAME_LZ_MAXCOL || es->maxit < 1 || ls->maxit < 1)
        FAIL(FAME_ERR_ARG, "fame_eigs: need 1 <= nwant < nstep <= " + std::to_string(FAME_LZ_MAXCOL - 1) +
                               " and positive maxit");
    if (precision != FAME_PREC_F64 && precision != FAME_PREC_F32_REFINE) FAIL(FAME_ERR_ARG, "fame_eigs: unknown precision");
    HIPCHK(hipSetDevice(c->device));
    /* first solve of a context: tune the pass kernels on this batch shape */
    if (c->autotune) {
        if (!c->tuned[0]) CHK(c->tune(8));
        if (precision == FAME_PREC_F32_REFINE && !c->tuned[1]) CHK(c->tune(4));
    }
    auto t0 = std::chrono::steady_clock::now();
    c->ls_last = *ls;
    const int nk = c->nk, nstep = es->nstep, nwant = es->nwant, m = nstep + 1;
    LzBuffers b;
    CHK(lz_alloc(c, b, nk, nstep, nwant));
    long long q64_0 = c->cnt_qbq64, q32_0 = c->cnt_qbq32;
    std::vector<int> ls_iter(nk, 0), es_iter(nk, 0), restarts(nk, 0), conv(nk, 0);
    std::vector<int> lzmask(nk, 1);
    CHK(lz_decomp(c, b, true, precision, *ls, lzmask, ls_iter, es_iter));

    std::vector<double> T0((size_t)m * nk), T1((size_t)m * nk);
    std::vector<double> LT0((size_t)nstep * nk), Z((size_t)nstep * nstep * nk);
    std::vector<double> G((size_t)2 * m * m * nk), Qh((size_t)m * (nwant + 1) * nk);
    std::vector<double> T2(m), T3(m), cc(m), ss(m), Qf((size_t)m * m);
    for (int iter = 1; iter <= es->maxit; ++iter) {
        HIPCHK(hipMemcpyAsync(T0.data(), b.T0d, sizeof(double) * m * nk, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(T1.data(), b.T1d, sizeof(double) * m * nk, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<int> rmask(nk, 0);
        int nrest = 0;
        for (int p = 0; p < nk; ++p) {
            if (!lzmask[p]) continue;
            double* t0 = &T0[(size_t)p * m];
            double* t1 = &T1[(size_t)p * m];
            double* lt0 = &LT0[(size_t)p * nstep];
            double* z = &Z[(size_t)p * nstep * nstep];
            if (fame_tridiag_eig(nstep, t0, t1, lt0, z) != 0) FAIL(FAME_ERR_HIP, "tridiagonal eigensolver did not converge");
            restarts[p] = iter;
            int cv = 0;
            for (int i = 0; i < nwant; ++i) {     /* Lanczos_Isotropic.cu:59-68 */
                double res = t1[nstep - 1] * std::fabs(z[(size_t)(i + 1) * nstep - 1]);
                if (res < es->tol) cv++; else break;
            }
            if (cv == nwant) { conv[p] = 1; lzmask[p] = 0; continue; }
            rmask[p] = 1; nrest++;
        }
        if (!nrest) break;
        /* Gram matrices of the bases that restart */
        HIPCHK(hipMemcpyAsync(b.mask, rmask.data(), sizeof(int) * nk, hipMemcpyHostToDevice, c->stream));
        const size_t ps = b.n2 * m;
        const int gblk = gram_blk(c->gram_batch > 0 ? c->gram_batch : nk);
        HIPCHK(fame_gram(b.U, b.gpart, gblk, m, b.n2, b.n2, ps, nk, b.mask, c->stream));
        HIPCHK(fame_reduce_partials(b.gpart, gblk, 2 * m * m, b.G, nk, c->stream));
        HIPCHK(hipMemcpyAsync(G.data(), b.G, sizeof(double) * 2 * m * m * nk, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int p = 0; p < nk; ++p) {
            if (!rmask[p]) continue;
            double* t0 = &T0[(size_t)p * m];
            double* t1 = &T1[(size_t)p * m];
            const double* lt0 = &LT0[(size_t)p * nstep];
            const double* Gp = &G[(size_t)p * 2 * m * m];
            /* Q = I ; columns follow the reference's Zdrot sequence (Lanczos_LockPurge_gpu.cu:23-45) */
            std::fill(Qf.begin(), Qf.end(), 0.0);
            for (int i = 0; i < m; ++i) Qf[(size_t)i * m + i] = 1.0;
            auto rot = [&](int xcol, int ycol, double cs, double sn) {   /* x = c x + s y ; y = -s x + c y */
                double* x = &Qf[(size_t)xcol * m];
                double* y = &Qf[(size_t)ycol * m];
                for (int r = 0; r < m; ++r) { double xv = x[r], yv = y[r]; x[r] = cs * xv + sn * yv; y[r] = -sn * xv + cs * yv; }
            };
            auto gnorm = [&](int qc) {   /* ||U q|| with G = U^H U (upper triangle stored) */
                const double* q = &Qf[(size_t)qc * m];
                double s = 0.0;
                for (int r = 0; r < m; ++r) {
                    if (q[r] == 0.0) continue;
                    s += q[r] * q[r] * Gp[2 * (r * m + r)];
                    for (int t = r + 1; t < m; ++t) s += 2.0 * q[r] * q[t] * Gp[2 * (r * m + t)];
                }
                return std::sqrt(std::max(s, 0.0));
            };
            for (int j = 0; j < nstep - nwant; ++j) {
                fame_tridiag_qr_step(t0, t1, T2.data(), T3.data(), cc.data(), ss.data(), lt0[nwant + j], nstep - j);
                int jj;
                for (jj = 0; jj < nstep - 1 - j; ++jj) rot(jj + 1, jj, cc[jj], ss[jj]);
                ss[jj - 1] *= t1[jj];
                rot(jj + 1, jj, t1[jj - 1], ss[jj - 1]);
                t1[jj - 1] = gnorm(jj);
                double inv = 1.0 / t1[jj - 1];
                for (int r = 0; r < m; ++r) Qf[(size_t)jj * m + r] *= inv;
                t0[jj] = 0.0; t1[jj] = 0.0;
                ++jj;
                for (int r = 0; r < m; ++r) Qf[(size_t)jj * m + r] = 0.0;
            }
            std::memcpy(&Qh[(size_t)p * m * (nwant + 1)], Qf.data(), sizeof(double) * m * (nwant + 1));
        }
        HIPCHK(hipMemcpyAsync(b.Q, Qh.data(), sizeof(double) * m * (nwant + 1) * nk, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b.T0d, T0.data(), sizeof(double) * m * nk, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(b.T1d, T1.data(), sizeof(double) * m * nk, hipMemcpyHostToDevice, c->stream));
        const size_t wps = b.n2 * (nwant + 1);
        HIPCHK(fame_gemm_vq(b.U, b.Q, b.W, m, nwant + 1, b.n2, b.n2, ps, b.n2, wps, (size_t)m * (nwant + 1), nk, b.mask,
                            c->stream));
        for (int p = 0; p < nk; ++p)
            if (rmask[p]) HIPCHK(hipMemcpyAsync(b.U + p * ps, b.W + p * wps, 16 * wps, hipMemcpyDeviceToDevice, c->stream));
        CHK(lz_decomp(c, b, false, precision, *ls, rmask, ls_iter, es_iter));
        if (iter == es->maxit) for (int p = 0; p < nk; ++p) if (rmask[p]) restarts[p] = es->maxit + 1;
    }
    /* Ritz values and vectors: omega = sqrt(1/theta), ev = U z (Lanczos_Isotropic.cu:87-93) */
    std::vector<double> Zq((size_t)nstep * nwant * nk);
    for (int p = 0; p < nk; ++p) {
        for (int i = 0; i < nwant; ++i) freq[(size_t)p * nwant + i] = std::sqrt(1.0 / LT0[(size_t)p * nstep + i]);
        std::memcpy(&Zq[(size_t)p * nstep * nwant], &Z[(size_t)p * nstep * nstep], sizeof(double) * nstep * nwant);
    }
    HIPCHK(hipMemcpyAsync(b.Q, Zq.data(), sizeof(double) * nstep * nwant * nk, hipMemcpyHostToDevice, c->stream));
    const size_t ps = b.n2 * m, wps = b.n2 * (nwant + 1);
    HIPCHK(fame_gemm_vq(b.U, b.Q, b.W, nstep, nwant, b.n2, b.n2, ps, b.n2, wps, (size_t)nstep * nwant, nk, nullptr,
                        c->stream));
    if (c->ritz_sweeps > 0) CHK(refine_ritz(c, b, precision, freq, ls_iter, es_iter));
    if (evr_dev) {
        /* reduced eigenvectors, reference layout (2Nd) with a 2N stride per band */
        for (int p = 0; p < nk; ++p)
            for (int i = 0; i < nwant; ++i)
                HIPCHK(fame_launch_pack(b.W + p * wps + (size_t)i * b.n2,
                                        (double2*)evr_dev + ((size_t)p * c->out_pmul * nwant + i) * b.n2, c->N, c->nd[p], -1,
                                        8, c->stream));
    }
    if (E_dev) CHK(fame_restore_impl(c, nwant, b.W, wps, E_dev));
    /* Gamma rule on the frequencies (FAME_Fast_Algorithms_Isotropic.cu:61-70) */
    for (int p = 0; p < nk; ++p) {
        if (!c->prob_h[p].gamma) continue;
        double* f = freq + (size_t)p * nwant;
        for (int i = nwant - 1; i >= 2; --i) f[i] = f[i - 2];
        f[0] = 0.0;
        if (nwant > 1) f[1] = 0.0;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (prof) {
        for (int p = 0; p < nk; ++p) {
            fame_profile& q = prof[p];
            q.ls_iter = ls_iter[p]; q.es_iter = es_iter[p]; q.restarts = restarts[p]; q.nd = c->nd[p];
            q.converged = conv[p]; q.pad = 0; q.es_time = secs; q.ls_time = b.ls_time[p];
            q.qbq64 = (int)((c->cnt_qbq64 - q64_0) / nk);
            q.qbq32 = (int)((c->cnt_qbq32 - q32_0) / nk);
        }
    }
    return FAME_OK;
}

/* E_i = invB .* Qr(Sigma .* ev_i), times ||E_i|| (reference scaling), for band i
 * of every problem (FAME_Fast_Algorithms_Isotropic.cu:90-125).  ev: padded
 * reduced vectors of band i with per-problem stride ev_pstride.              */
static fame_status restore_band(fame_ctx* c, int i, int nwant, const double2* ev, size_t ev_pstride, void* E_dev) {
    const size_t N = c->N, n2 = 2 * N, N3 = 3 * N;
    const int nk = c->nk;
    double2* y = (double2*)c->ref64[0];
    std::vector<double> nr(nk);
    HIPCHK(fame_convert(y, 8, ev, 8, n2, n2, ev_pstride, nk, c->stream));
    HIPCHK(fame_scale_real_mod(y, y, c->sig, n2, n2, N, N, nk, 8, c->stream));
    CHK(c->qr_field(y, 0, nk, 0, nullptr));
    HIPCHK(fame_scale_real_mod(c->F64, c->F64, c->invB64, N3, N3, 0, N3, nk, 8, c->stream));
    HIPCHK(fame_norm2_partials(c->F64, 8, c->partials, NBLK_VEC, N3, N3, nk, c->stream));
    HIPCHK(fame_reduce_partials(c->partials, NBLK_VEC, 1, c->scal, nk, c->stream));
    HIPCHK(hipMemcpyAsync(nr.data(), c->scal, sizeof(double) * nk, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int p = 0; p < nk; ++p) nr[p] = std::sqrt(nr[p]);   /* Zdscal by the norm itself (:119-121) */
    HIPCHK(hipMemcpyAsync(c->scal, nr.data(), sizeof(double) * nk, hipMemcpyHostToDevice, c->stream));
    HIPCHK(fame_scale_real_mod(c->F64, c->F64, c->scal, N3, N3, 1, 1, nk, 8, c->stream));
    for (int p = 0; p < nk; ++p)
        CHK(c->field_to_ref(c->F64 + (size_t)p * N3, (double2*)E_dev + ((size_t)p * c->out_pmul * nwant + i) * N3, 1, N3));
    return FAME_OK;
}

/* Gamma rule on the fields: bands move up by two, E_0 = E_1 = 1/sqrt(3N)
 * (FAME_Fast_Algorithms_Isotropic.cu:71-79) */
static fame_status restore_gamma(fame_ctx* c, int nwant, void* E_dev) {
    const size_t N3 = 3 * (size_t)c->N;
    for (int p = 0; p < c->nk; ++p) {
        if (!c->prob_h[p].gamma) continue;
        double2* E = (double2*)E_dev + (size_t)p * c->out_pmul * nwant * N3;
        for (int i = nwant - 1; i >= 2; --i)
            HIPCHK(hipMemcpyAsync(E + (size_t)i * N3, E + (size_t)(i - 2) * N3, 16 * N3, hipMemcpyDeviceToDevice, c->stream));
        std::vector<double2> cst(N3, make_double2(1.0 / std::sqrt((double)N3), 0.0));
        for (int i = 0; i < std::min(2, nwant); ++i)
            HIPCHK(hipMemcpyAsync(E + (size_t)i * N3, cst.data(), 16 * N3, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return FAME_OK;
}

/* ev: padded reduced vectors with per-problem stride ev_pstride and per-band stride 2N */
fame_status fame_restore_impl(fame_ctx* c, int nwant, const double2* ev, size_t ev_pstride, void* E_dev) {
    for (int i = 0; i < nwant; ++i) CHK(restore_band(c, i, nwant, ev + (size_t)i * 2 * c->N, ev_pstride, E_dev));
    return restore_gamma(c, nwant, E_dev);
}

extern "C" fame_status fame_restore(fame_ctx* c, int nwant, const void* evr_dev, void* E_dev) {
    if (!c || !evr_dev || !E_dev || nwant < 1) FAIL(FAME_ERR_ARG, "fame_restore: bad arguments");
    if (!c->have_eps || c->nk < 1) FAIL(FAME_ERR_STATE, "fame_restore: permittivity and wave vectors must be set");
    HIPCHK(hipSetDevice(c->device));
    const size_t n2 = 2 * (size_t)c->N;
    double2* stage = (double2*)c->ref64[2];
    for (int i = 0; i < nwant; ++i) {
        for (int p = 0; p < c->nk; ++p)
            HIPCHK(fame_launch_pack((const double2*)evr_dev + ((size_t)p * nwant + i) * n2, stage + (size_t)p * n2, c->N,
                                    c->nd[p], +1, 8, c->stream));
        CHK(restore_band(c, i, nwant, stage, n2, E_dev));
    }
    CHK(restore_gamma(c, nwant, E_dev));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FAME_OK;
}
