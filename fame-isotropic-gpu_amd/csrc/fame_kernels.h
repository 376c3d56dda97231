/*
 * fame_kernels.h -- launch interface between the host solver and the gfx950
 * kernels (not part of the public C-ABI; see include/fame_hip.h for that).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "fame_device.hpp"

enum { OUTER_EPS = 0, OUTER_QR = 1, OUTER_QRS = 2 };

#define MAX_LINES_INNER 32   /* lines (b, c) per block of the inner-axis kernels */

/* stage lists an axis plan provides (per precision):
 * FWD  forward transform, INV inverse transform (both unnormalised),
 * EPS  INV, kernel elementwise step, FWD (outer pass of QBQ),
 * FI   FWD, kernel elementwise step, INV (curl-curl inner pass).             */
enum { LIST_FWD = 0, LIST_INV = 1, LIST_EPS = 2, LIST_FI = 3, NLIST = 4 };

/* per-problem CG scalars, device resident */
struct CgScal {
    double rr;        /* current <r,r>                                         */
    double rr0;       /* <b,b> of the solve                                    */
    double pap_re, pap_im;
    double alpha_re, alpha_im;
    double beta;
    double tol2;      /* absolute tolerance squared                           */
    int iters;        /* reference-style counter k (starts at 1)              */
    int maxit;
    int active;       /* 1 while iterating                                      */
    int pend;         /* 1 when x still lacks alpha p of the last iteration    */
};

struct KInner {
    fame::GridDev g;
    const fame::ProbDev* prob;   /* device, one per problem                    */
    fame::LinePlanDev plan;      /* stage list along the inner axis            */
    const int* active;           /* device flags or nullptr                    */
    const void* x;               /* EXPAND: reduced input (padded, 2N/problem) */
    void* y;                     /* !EXPAND: reduced output                    */
    void* f;                     /* field, 3N per problem                      */
    const void* dotvec;          /* !EXPAND: optional d for partial conj(d).y  */
    double* partials;            /* [problem][block][2]                        */
    size_t xstride, fstride;
    int L, P;                    /* lines per block, LDS line pitch            */
    int pikind;                  /* 0 Pi_Qr, 1 Pi_Pr                           */
    /* EXPAND inside CG (cg_r != nullptr): x is the search direction p, and
     * before it is expanded each element is updated in place,
     *   x_sol += alpha p  (when pend),   p = r + beta p,
     * with alpha, beta, pend of the problem's CgScal: the previous iteration's
     * solution and direction updates ride on this pass's reads.             */
    void* cg_r;
    void* cg_x;
    const CgScal* cgsc;
    /* CG residual update on the same pass: with cg_ap set, r -= alpha Ap of
     * the previous iteration (when pend) is applied before p = r + beta p,
     * and <r,r> goes to slot 5 of the partials.
     * !EXPAND with cg_r set: the pass updates r -= alpha y in place instead of
     * storing y (alpha of this iteration, already in cgsc), and puts <r,r> of
     * the updated residual in slot 3 of 6 doubles per block                 */
    const void* cg_ap;
    /* EXPAND, refinement (fx32 set): before it is expanded, x += fx32 + alpha
     * fp32 (the FP32 solve's solution and its pending update, alpha and pend
     * of cgsc, FP32 complex vectors with problem stride xstride) is formed in
     * FP32, added in FP64 and stored back to x                               */
    const void* fx32;
    const void* fp32;
    /* !EXPAND, refinement residual (rb set, complex double like y): instead
     * of y the pass stores the next FP32 solve's start, r32 = p32 = rb - y
     * rounded to FP32 and x32 = 0 (FP32 complex vectors, problem stride
     * xstride), and the block sums of |r32|^2 into rpart[problem][block]   */
    const void* rb;
    void* r32;
    void* p32;
    void* x32;
    double* rpart;
    /* !EXPAND inside CG: <y, y> into slot 2 of 6 doubles per block (the
     * expanding pass puts <r, r> in slot 5; with dotvec, conj(dotvec).y goes
     * to slots 0 and 1), from which the CG takes <r, r> of the next residual
     * without a pass over r                                                  */
    int cgdots;
};

struct KAxis {
    fame::GridDev g;
    const fame::ProbDev* prob;
    fame::LinePlanDev plan;
    const int* active;
    void* f;
    const void* invB;            /* outer EPS mode: 3N reals, internal layout  */
    /* outer EPS mode inside CG (optional): per block, the sum over its lines of
     * conj(v) (scale eps^-1 v) of the real-space values v, i.e. the block's
     * share of <p, QBQ p> when every line transform is an unnormalised DFT;
     * layout [problem][component][gridDim.x], blocks past the last tile write 0 */
    double* pap;
    size_t fstride;
    int L, P;                    /* lines per block, LDS line pitch            */
    unsigned mL;                 /* udiv magic for / L                         */
};

/* inner-axis kernels keep a table of na complex doubles after their lines:
 * its offset in line entries (rounded to 16 bytes for FP32 entries) and the
 * entries it adds to the LDS size                                            */
__host__ __device__ inline int inner_tab_offset(int entries, int prec) {
    return prec == 8 ? entries : ((entries + 1) & ~1);
}
__host__ __device__ inline int inner_tab_extra(int na, int prec) { return prec == 8 ? na : 2 * na + 1; }
/* behind it, per line, the inner twiddle exp(2 i pi phi a / na) split as
 * a = 16 ah + al: 16 entries for al and ceil(na / 16) for ah, in the kernel's
 * precision.  Entries (kernel precision) of both tables for L lines:         */
__host__ __device__ inline int inner_tw_per_line(int na) { return 16 + (na + 15) / 16; }
__host__ __device__ inline int inner_lds_extra(int na, int prec, int L) {
    return inner_tab_extra(na, prec) + 1 + L * inner_tw_per_line(na);
}

/* LDS bytes of a line-transform launch: tables + nl lines + extra entries */
size_t fame_line_lds(const fame::LinePlanDev& pl, int prec, int nl, int P, int extra);

hipError_t fame_launch_inner(const KInner& a, int expand, int prec, int nprob, hipStream_t s);
hipError_t fame_launch_mid(const KAxis& a, int dir, int prec, int nprob, hipStream_t s);
hipError_t fame_launch_outer(const KAxis& a, int mode, int prec, int nprob, hipStream_t s);
hipError_t fame_launch_xz_transpose(const void* in, void* out, int nx, int ny, int nz, int dir, size_t stride,
                                    int ncomp, int nprob, int prec, hipStream_t s);
hipError_t fame_launch_pack(const void* in, void* out, int N, int Nd, int dir, int prec, hipStream_t s);

/* ---- vector / reduction kernels (fame_blas.hip) ------------------------- */

/* x = 0, r = p = b and the partials of <r,r> (stride: problem stride in elements) */
hipError_t fame_cg_begin(void* x, void* r, void* p, const void* b, double* partials, int nblk,
                         size_t n, size_t stride, int nprob, int prec, const int* mask, hipStream_t s);
hipError_t fame_cg_init_scal(CgScal* sc, const double* partials, int nblk, double tol, int maxit, int nprob,
                             int* active, hipStream_t s);
/* CG iteration with Ap stored (any plan): alpha from the summed <r,r>, <r,r>
 * of the next residual by one step of |r - alpha Ap|^2 = |alpha|^2 |Ap|^2 -
 * |r|^2 (conjugacy: <r,Ap> = <p,Ap>), beta and the active flags, from the
 * 6-double partials of the expanding and contracting passes */
/* CG iteration with the residual updated by the contracting pass (when the
 * outer pass sums <p,Ap>): alpha = <r,r> / <p,Ap> from the outer pass's 3 npap
 * block sums per problem, launched before the contracting pass; then beta and
 * the active flags from the exact <r,r> of the updated residual (slot 3 of the
 * inner partials) */
hipError_t fame_cg_alpha(CgScal* sc, const double* pap, int npap, int nprob, const int* active, hipStream_t s);
hipError_t fame_cg_beta(CgScal* sc, const double* partials, int nblk, int nprob, int* active, hipStream_t s);
hipError_t fame_cg_alpha_beta(CgScal* sc, const double* partials, int nblk, int nprob, int* active, hipStream_t s);
hipError_t fame_cg_finish_x(void* x, const void* p, const CgScal* sc, size_t n, size_t stride, int nprob, int prec,
                            hipStream_t s);
/* an FP32 solve inside FP64 refinement starting from the FP64 right-hand
 * side: x32 = 0, r32 = p32 = b64 rounded, with the <r,r> partials */
hipError_t fame_cg_begin_mixed(void* x32, void* r32, void* p32, const void* b64, double* partials, int nblk, size_t n,
                               size_t stride, int nprob, const int* mask, hipStream_t s);

/* generic helpers (complex double unless prec says float) */
hipError_t fame_scale_real(void* y, const void* x, const double* sreal, size_t n, size_t xstride, size_t sstride,
                           int nprob, int prec, hipStream_t s);
/* mask (device, optional): problems with mask 0 are left untouched */
hipError_t fame_convert(void* dst, int dprec, const void* src, int sprec, size_t n, size_t dstride, size_t sstride,
                        int nprob, hipStream_t s, const int* mask = nullptr);
hipError_t fame_norm2_partials(const void* x, int prec, double* partials, int nblk, size_t n, size_t stride, int nprob,
                               hipStream_t s);
hipError_t fame_reduce_partials(const double* partials, int nblk, int ncomp, double* out, int nprob, hipStream_t s);

/* Lanczos block kernels: V columns of length n (complex double), column stride cstride,
 * problem stride pstride */
hipError_t fame_gemv_h(const void* V, const void* w, double* partials, int nblk, int m, size_t n, size_t cstride,
                       size_t pstride, size_t wstride, int nprob, const int* mask, hipStream_t s);
hipError_t fame_gemv_n(const void* V, void* w, const double* h, int m, size_t n, size_t cstride, size_t pstride,
                       size_t wstride, size_t hstride, int nprob, const int* mask, hipStream_t s);
/* most basis columns (nstep + 1) the Gram and U*Q kernels take */
#define FAME_LZ_MAXCOL 128
hipError_t fame_gram(const void* V, double* partials, int nblk, int m, size_t n, size_t cstride, size_t pstride,
                     int nprob, const int* mask, hipStream_t s);
hipError_t fame_gemm_vq(const void* V, const double* Q, void* W, int m, int q, size_t n, size_t cstride,
                        size_t pstride, size_t wcstride, size_t wpstride, size_t qstride, int nprob, const int* mask,
                        hipStream_t s);
hipError_t fame_lanczos_step(void* V, int j, const double* alpha_beta_prev, double* partials, int nblk, size_t n,
                             size_t cstride, size_t pstride, int nprob, const int* mask, hipStream_t s);
hipError_t fame_launch_sigma(const fame::GridDev& g, const fame::ProbDev* prob, double* sig, double* sig_inv, int nprob,
                             hipStream_t s);

/* Lanczos scalar helpers (fame_blas.hip) */
hipError_t fame_lz_coeffs(const double* hdot, double* T0d, const double* T1d, int ld, int j, double* hout,
                          int nprob, const int* mask, hipStream_t s);
hipError_t fame_lz_second_pass_mask(const double* h, int n, double tau2, const int* in, int* out, int nprob,
                                    hipStream_t s);
hipError_t fame_lz_beta(const double* nrm2, double* T1d, int ld, int j, double* invb, int nprob, const int* mask,
                        hipStream_t s);
hipError_t fame_scale_real_mod(void* y, const void* x, const double* sreal, size_t n, size_t xstride, size_t sstride,
                               size_t smod, int nprob, int prec, hipStream_t s, const int* mask = nullptr);
/* y = s x with separate problem strides for y and x */
hipError_t fame_scale_real_mod2(void* y, size_t ystride, const void* x, size_t xstride, const double* sreal, size_t n,
                                size_t sstride, size_t smod, int nprob, int prec, hipStream_t s,
                                const int* mask = nullptr);
hipError_t fame_cg_init_scal2(CgScal* sc, const double* partials, int nblk, double tol, double reltol, int maxit,
                              int nprob, int* active, const int* mask, hipStream_t s);
