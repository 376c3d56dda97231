/*
 * fame_host.h -- internal host-side state of a solver context (not public).
 */
#pragma once
#include <map>
#include <tuple>
#include <vector>
#include <hip/hip_runtime.h>
#include "fame_hip.h"
#include "fame_kernels.h"

#define NBLK_VEC 1024   /* blocks per problem of the vector/reduction kernels */
#define CG_GRAPH_ITERS 4   /* CG iterations per captured graph (fewer launch gaps between them) */

enum { FFT_ENGINE_AUTO = 0, FFT_ENGINE_STOCKHAM = 1, FFT_ENGINE_MFMA = 2 };
enum { AX_INNER = 0, AX_MID = 1, AX_OUTER = 2 };

struct PassTile { int eng = 0, L = 0, P = 0; };

struct AxisPlan {
    int n = 0, M = 0, blue = 0, mfma = 0;
    std::vector<int> radices;
    fame::LinePlanDev lists[2][NLIST];   /* [0] FP64, [1] FP32; LIST_* stage lists */
    int cap[2] = {0, 0};                 /* elements a 256-thread block transforms at once */
    std::vector<void*> bufs;
    fame_status build(int n, int engine);
    ~AxisPlan();
};

struct fame_ctx {
    int device = 0;
    hipStream_t stream = 0;
    fame_lattice lat;
    fame::GridDev g;
    int N = 0, maxb = 0;
    /* plans per axis (AX_*) and engine (0 Stockham/Bluestein, 1 direct DFT on the
     * matrix cores; ax[a][1].n == 0 when not available for that length) */
    AxisPlan ax[3][2];
    PassTile tile[3][2];          /* chosen engine and lines per block, [axis][precision index] */
    int nblk_in[2] = {0, 0};      /* blocks of the inner-axis kernels per problem */
    int eng_forced = -1;          /* FAME_FFT: -1 auto, else the only engine allowed */
    bool tuned[2] = {false, false};
    bool autotune = true;
    /* permittivity, kernel layout */
    double* invB64 = nullptr; float* invB32 = nullptr; double* eps_int = nullptr; double* eps_ref = nullptr;
    bool have_eps = false;
    /* wave vectors */
    int nk = 0;
    std::vector<fame::ProbDev> prob_h;
    std::vector<double> kvec;    /* the wave vectors as given (3 per problem) */
    std::vector<int> nd;
    fame::ProbDev* prob_d = nullptr;
    double* sig = nullptr;       /* nk x N  |Lambda|          */
    double* sig_inv = nullptr;   /* nk x N  1/|Lambda| (0 at the removed Gamma mode) */
    /* work */
    double2* F64 = nullptr; float2* F32 = nullptr;        /* fields, 3N per problem        */
    void* cg64[4] = {nullptr, nullptr, nullptr, nullptr};   /* r, p, Ap, spare (2N/problem) */
    void* cg32[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  /* r, p, Ap, rhs32, sol32 */
    void* ref64[3] = {nullptr, nullptr, nullptr};
    double* partials = nullptr;   /* [problem][NBLK_VEC][<=4]   */
    double* partials2 = nullptr;  /* [problem][inner blocks][2, or 6 inside CG] */
    double* pap_part = nullptr;   /* [problem][3][outer blocks]: <p,QBQ p> block sums of the outer pass */
    CgScal* cgsc = nullptr;
    int* active = nullptr;
    int* mask = nullptr;
    double* scal = nullptr;       /* 64 doubles per problem of small device scalars */
    CgScal* sc_h = nullptr;       /* pinned host copy of cgsc */
    std::vector<void*> allocs;
    std::vector<void*> big;       /* Lanczos basis etc., reallocated per solve shape */
    std::vector<size_t> big_bytes;  /* their sizes (the shape they were allocated for) */
    /* tuning / state */
    int cg_predict[2] = {40, 40};
    /* inverse-iteration sweeps applied by fame_eigs to the converged Ritz pairs
     * (fame_set_refinement; 0 = the reference's algorithm), their absolute CG
     * tolerance and iteration limit */
    int ritz_sweeps = 0;
    double ritz_tol = 1e-14;
    int ritz_maxit = 2000;
    int refine_max_sweeps = 8;
    double refine_reltol = 1e-6;
    fame_ls ls_last = {0, 0.0};
    /* fame_eigs on a large batch: the wave vectors are split over child
     * contexts of the same structure (kids[i] holds wave vectors i, i + G,
     * i + 2G, ... of the parent's kid_nk), solved concurrently as by
     * fame_eigs_many, with the parent's pass tiles; groups: 0 auto, 1 one
     * batch, g >= 2 that many.  eps_ver / k_ver count fame_set_eps /
     * fame_set_wave_vectors calls, so the children are brought up to date only
     * when the parent changed.  A child sizes its Gram partial sums for
     * gram_batch problems (the parent's batch), so that its arithmetic is the
     * parent's, and writes its device outputs (reduced vectors, fields) with a
     * problem stride of out_pmul problems, into its slots of the parent's. */
    int groups = 0;
    bool owns_stream = false;     /* a child's stream, destroyed with it */
    std::vector<fame_ctx*> kids;
    int kid_nk = 0;
    int eps_ver = 0, k_ver = 0, kids_eps_ver = -1, kids_k_ver = -1;
    int gram_batch = 0;
    int out_pmul = 1;
    long long cnt_qbq64 = 0, cnt_qbq32 = 0;
    int last_iters = 0;
    /* one CG iteration captured as a hipGraph per (precision, first problem,
     * problem count, solution vector); replayed chunk times between the host
     * checks of the active flags.  FAME_GRAPH=0 launches eagerly instead.   */
    bool use_graphs = true;
    hipStream_t cap_stream = nullptr;
    std::map<std::tuple<int, int, int, const void*, int>, hipGraphExec_t> cg_graphs;
    void drop_graphs();
    fame_status cg_iteration(int prec, void* x, int p0, int nprob, int count = 1);
    fame_status cg_step_ap(int prec, void* p, void* r, void* x, void* Ap, int p0, int nprob);
    fame_status cg_step_rc(int prec, void* p, void* r, void* x, int p0, int nprob, int npap);

    KInner kin(int prec, int p0, const int* act, int list) const;
    KAxis kmid(int prec, int p0, const int* act, int list) const;
    KAxis kout(int prec, int p0, const int* act, int list) const;
    fame_status qbq(int prec, const void* x, void* y, int p0, int nprob, const int* act, const void* dotvec,
                    void* cg_r = nullptr, void* cg_x = nullptr, const void* cg_ap = nullptr,
                    bool cgdots = false);
    fame_status qr_field(const void* x, int p0, int nprob, int pikind, const int* act);
    fame_status qrs_field(void* y, int p0, int nprob, int pikind, const int* act);
    fame_status field_to_ref(const void* fint, void* fref, int nprob, size_t dst_stride);
    fame_status field_from_ref(const void* fref, void* fint, int nprob);
    /* b64: b is complex double while prec is 4 (rounded to FP32 at the start);
     * defer_finish: the last pending x += alpha p is left to the caller
     * (qbq_residual32 applies it while adding x into the FP64 iterate);
     * begun_nblk > 0: x, r, p were set up by qbq_residual32 (or the caller),
     * with the <r,r> partials in partials, begun_nblk per problem */
    fame_status cg(int prec, const void* b, void* x, int p0, int nprob, double tol, double reltol, int maxit,
                   const int* mask_dev, int* iters, bool b64 = false, bool defer_finish = false, int begun_nblk = 0);
    /* refinement residual: xr += x32 + alpha p32 (the FP32 solve's result and
     * its pending update) on the first pass, then QBQ64 xr, whose last pass
     * stores the next FP32 solve's start r = p = FP32(b - QBQ64 xr), x = 0, and
     * its norm partials (partials[problem][inner block]) */
    fame_status qbq_residual32(const void* xr, const void* b, void* x32, int p0, int nprob, const int* act, int nact);
    /* x: problem stride xstride (0: 2N); out (optional): the result is written
     * there, problem stride ostride, for the problems of mask_dev only, and y
     * serves as the contiguous solve buffer; mask_host: a host copy of
     * mask_dev, which saves reading it back */
    fame_status invar(int mode, const void* x, void* y, int p0, int nprob, const int* mask_dev, fame_ls ls, int* iters,
                      size_t xstride = 0, void* out = nullptr, size_t ostride = 0, const int* mask_host = nullptr);
    fame_status tune(int prec);
    const AxisPlan& plan(int axis, int prec) const { return ax[axis][tile[axis][prec == 8 ? 0 : 1].eng]; }
    /* outer-pass blocks per (problem, component) of precision prec, when the
     * outer pass can sum <p, QBQ p> from the real-space field: every current
     * line transform is an unnormalised DFT (no Bluestein plan), so that the
     * forward chain is the exact adjoint of the inverse one and the outer
     * pass's 1/N is the whole normalisation; 0 otherwise */
    int pap_blocks(int prec) const;
    ~fame_ctx();
};

void fame_phases(const fame_lattice& L, const double* k, fame::ProbDev* p);
unsigned fame_magic32(int d);
