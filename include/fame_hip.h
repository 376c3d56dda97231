/*
 * fame_hip.h -- C-ABI of the MI355X (gfx950) null-space-free Maxwell band solver.
 *
 * Plain C types only (pointers, sizes, status codes); no torch or HIP types in
 * the signatures.  Vector arguments named *_dev are device pointers; layouts are
 * the reference's:
 *   reduced vectors  [x1; x2], 2*Nd complex doubles (interleaved re,im),
 *                    mode i <-> Fourier index N-Nd+i
 *   fields           [Ex; Ey; Ez], 3*N complex doubles, each x-fastest
 *   permittivity     3*N doubles on the Yee edges, same layout (B_eps)
 *
 * Every entry point below replaces one reference function of the
 * FAME_Main_Code -> Lanczos_Isotropic -> invAr -> CG -> QBQ -> Qr/Qrs path; the
 * reference interface is cited next to it (file:line under the reference tree).
 * No entry point aborts: errors are returned as fame_status, and
 * fame_last_error() gives the message.
 */
#ifndef FAME_HIP_H
#define FAME_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FAME_OK = 0,
    FAME_ERR_ARG = 1,          /* invalid argument or size                       */
    FAME_ERR_HIP = 2,          /* HIP runtime error (message via fame_last_error) */
    FAME_ERR_NOMEM = 3,        /* device allocation failed                       */
    FAME_ERR_STATE = 4,        /* call order (e.g. eigs before set_eps)          */
    FAME_ERR_UNSUPPORTED = 5   /* e.g. FFT length beyond the supported maximum   */
} fame_status;

/* Discretisation of one structure.  Fields mirror PAR.mesh / PAR.lattice of
 * include/FAME_Internal_Common.h:62-127 (grid_nums, mesh_len, lattice_vec_a,
 * lattice_constant.m1..m4, t1..t4); `general` is the Simple/General dispatch of
 * FAME_Main/FAME_Main_Code.cu:156-171.                                          */
typedef struct {
    int    grid_nums[3];
    int    general;
    double mesh_len[3];
    double lattice_vec_a[9];
    int    m1, m2, m3, m4;
    double t1[3], t2[3], t3[3], t4[3];
} fame_lattice;

/* ES / LS of include/FAME_Internal_Common.h:160-170 */
typedef struct { int nwant, nstep, maxit; double tol; } fame_es;
typedef struct { int maxit; double tol; } fame_ls;

/* PROFILE of include/FAME_Internal_Common.h:205-211, per wave vector, plus the
 * counters of the mixed-precision solver                                        */
typedef struct {
    int    ls_iter;        /* CG iterations summed over all invAr (reference counting) */
    int    es_iter;        /* invAr applies                                          */
    int    restarts;       /* Lanczos_Isotropic return value                         */
    int    nd;             /* Nd (N-1 at Gamma)                                      */
    int    converged;      /* 1 if the Lanczos convergence test passed               */
    int    qbq64;          /* FP64 QBQ applications, mean per wave vector of the batch
                              (or of its group, fame_set_groups) solved with it      */
    int    qbq32;          /* FP32 QBQ applications (refinement inner solves), same   */
    int    pad;
    double es_time;        /* seconds, eigensolve of the whole batch (or group)      */
    double ls_time;        /* seconds inside invAr (CG) solves, this wave vector's share
                              of the batch's solves (the reference's Profile.ls_time,
                              FAME_Matrix_Vector_Production_Isotropic_invAr.cu:29-43) */
} fame_profile;

enum { FAME_PREC_F64 = 0, FAME_PREC_F32_REFINE = 1 };

enum {
    FAME_OP_QR = 0,    /* FAME_Main/FAME_Matrix_Vector_Production_Qr.cu:8,43     2Nd -> 3N */
    FAME_OP_QRS = 1,   /* FAME_Main/FAME_Matrix_Vector_Production_Qrs.cu:7,31    3N -> 2Nd */
    FAME_OP_PR = 2,    /* FAME_Main/FAME_Matrix_Vector_Production_Pr.cu:11,45    2Nd -> 3N */
    FAME_OP_PRS = 3,   /* FAME_Main/FAME_Matrix_Vector_Production_Prs.cu:9,33    3N -> 2Nd */
    FAME_OP_QBQ = 4,   /* FAME_Main/FAME_Matrix_Vector_Production_Isotropic_QBQ.cu:8,38  2Nd -> 2Nd */
    FAME_OP_INVAR = 5, /* FAME_Main/FAME_Matrix_Vector_Production_Isotropic_invAr.cu:8   2Nd -> 2Nd */
    FAME_OP_INVAR_MIXED = 6, /* invAr with FP32 CG inside FP64 iterative refinement (no reference counterpart) */
    FAME_OP_QBQ32 = 7  /* QBQ evaluated in complex FP32 (inner operator of the mixed solver), FP64 in/out */
};

typedef struct fame_ctx fame_ctx;

/* Create a solver context for one structure and up to max_batch wave vectors.
 * Replaces FAME_Create_cublas + FAME_Create_Buffer (FAME_Main/FAME_Create_cublas.cu:4,
 * FAME_Main/FAME_Create_Buffer.cu:4): FFT plans, twiddle tables, work buffers. */
fame_status fame_create(const fame_lattice* lat, int max_batch, int device, fame_ctx** ctx);
/* Replaces FAME_Destroy_Main (FAME_Main/FAME_Destroy_Main.cu:4) */
fame_status fame_destroy(fame_ctx* ctx);

/* Permittivity on the Yee edges, host array of 3N doubles (B_eps layout).
 * Replaces FAME_Matrix_B_Isotropic (FAME_Main/FAME_Matrix_B_Isotropic.cu:4).     */
fame_status fame_set_eps(fame_ctx* ctx, const double* eps3N);

/* Wave vectors (host, 3*nk doubles, units of 2*pi as PAR.recip_lattice.WaveVector).
 * nd_out (host, nk ints, may be NULL) receives Nd per wave vector.
 * Replaces FAME_Matrix_Lambdas (FAME_Main/FAME_Matrix_Lambdas.cu:40): the Lambda,
 * Pi and D tables are evaluated on the fly by the kernels from these phases.   */
fame_status fame_set_wave_vectors(fame_ctx* ctx, int nk, const double* k, int* nd_out);

/* Apply one operator for wave vector ik on device vectors (reference layouts). */
fame_status fame_apply(fame_ctx* ctx, int op, int ik, const void* x_dev, void* y_dev);

/* CG on QBQ y = rhs for wave vector ik (FAME_Main/CG.cu:5,90): x0 = 0, absolute
 * tolerance, returns the reference's iteration counter k in *iters.           */
fame_status fame_cg(fame_ctx* ctx, int ik, const void* rhs_dev, void* x_dev, int maxit, double tol, int* iters);

/* Restarted Lanczos on A_r^{-1} for all wave vectors of the context in one
 * batch, then eigenvector restoration and the Gamma rule.
 * Replaces FAME_Fast_Algorithms_Isotropic + Lanczos_Isotropic
 * (FAME_Main/FAME_Fast_Algorithms_Isotropic.cu:20, FAME_Main/Lanczos_Isotropic.cu:7).
 *   precision : FAME_PREC_F64 (reference arithmetic) or FAME_PREC_F32_REFINE
 *               (complex FP32 FFT/CG inside FP64 iterative refinement)
 *   freq      : host, nk*nwant doubles, omega = sqrt(1/theta) (0 for Gamma zero modes)
 *   evr_dev   : optional device buffer of nk*nwant*2N complex doubles receiving the
 *               Ritz vectors of A_r (band b of wave vector k at offset (k*nwant+b)*2N,
 *               2Nd values in the reference layout, Lanczos order); may be NULL
 *   E_dev     : optional device buffer of nk*nwant*3N complex doubles receiving the
 *               restored E fields (reference scaling and Gamma rule); may be NULL
 *   prof      : optional host array of nk profiles                              */
fame_status fame_eigs(fame_ctx* ctx, int precision, const fame_es* es, const fame_ls* ls, double* freq,
                      void* evr_dev, void* E_dev, fame_profile* prof);

/* Inverse-iteration sweeps that fame_eigs applies to the converged Ritz pairs
 * before restoration (default 0: the reference's algorithm, whose stopping rule
 * bounds the Ritz residual of A_r^{-1}, not of A_r).  Each sweep solves
 * w = A_r^{-1} y for every band (one batched invAr per band, absolute CG
 * tolerance tol, at most maxit iterations, in the solve's precision), takes
 * lambda = Re<w,y>/<w,w> and y = w/||w||; it removes the high-frequency error
 * that dominates the field residual ||C^H C E - lambda B E|| / ||E||.  The extra
 * solves are counted in the profile.  No reference counterpart.             */
fame_status fame_set_refinement(fame_ctx* ctx, int sweeps, double tol, int maxit);

/* How fame_eigs runs a batch: groups = 1 as one batch; g >= 2 split into g
 * groups (wave vectors i, i + g, i + 2g, ... in group i, so that neighbouring
 * points of a path, of similar cost, go to different groups), each solved by
 * a child context of the same structure on its own stream, concurrently (as
 * fame_eigs_many does for several structures), so that one group's CG tails
 * and vector work overlap another's passes; 0 (default) splits every batch of
 * two or more wave vectors in two.  The children use the parent's pass tiles and
 * arithmetic, so the results are bit-identical to one batch.  They hold the
 * per-problem work vectors of their wave vectors (about 400 N bytes each)
 * until the context is destroyed; when the device lacks that memory,
 * fame_eigs runs one batch.  Replaces the reference's one-wave-vector-at-a-time
 * loop in FAME_Main_Code (FAME_Main_Code.cu:79-131) as a batching policy; no
 * reference counterpart.                                                    */
fame_status fame_set_groups(fame_ctx* ctx, int groups);

/* Restored E fields (device, nk*nwant*3N, reference scaling and Gamma rule, as
 * fame_eigs' E_dev) recomputed from the reduced Ritz vectors evr_dev (device,
 * nk*nwant*2N, as fame_eigs' evr_dev) with the permittivity and wave vectors
 * currently set.  Replaces Eigen_Restoration_Isotropic
 * (FAME_Main/FAME_Fast_Algorithms_Isotropic.cu:90) as a separate step, so that a
 * caller (the adjoint) keeps 2N instead of 3N complex values per band.      */
fame_status fame_restore(fame_ctx* ctx, int nwant, const void* evr_dev, void* E_dev);

/* n structures in one call (uneven grids, batch sizes and band counts): arrays
 * of n contexts and per-structure arguments as for fame_eigs (evr_dev, E_dev
 * and prof may be NULL, or hold NULL entries).  The contexts run concurrently,
 * each on its own stream.  Replaces the reference's one-structure-per-run
 * FAME_Main_Code loop (FAME_Main/FAME_Main_Code.cu:110) for batches of
 * structures; no direct counterpart.                                         */
fame_status fame_eigs_many(int n, fame_ctx* const* ctxs, const int* precision, const fame_es* es, const fame_ls* ls,
                           double* const* freq, void* const* evr_dev, void* const* E_dev, fame_profile* const* prof);

/* Relative residuals ||C^H C E - omega^2 B E|| / ||E|| of fields produced by
 * fame_eigs (device, nk*nwant*3N) -> res (host, nk*nwant).  GPU counterpart of
 * Check_Residual (FAME_Main/FAME_Main_Code.cu:383), reported relative.       */
fame_status fame_residuals(fame_ctx* ctx, int nwant, const double* freq, const void* ev_dev, double* res);

/* GPU counterpart of Check_Eigendecomp (FAME_Main/FAME_Main_Code.cu:217-381)
 * for wave vector ik of the context: with a random field x (real and imaginary
 * parts uniform in [0, 1), drawn from `seed` instead of the reference's time
 * seed), err[i] = ||C_i x_i - T Lambda_i T* x_i|| for i = x, y, z and
 * err[3] = ||C x - P Sigma Q^H x||, absolute 2-norms as the reference prints
 * them (EigDecomp_test_C1..C3, SVD_test_C).  C_i and C are the sparse forward
 * differences and curl of FAME_Matrix_Curl applied as a stencil; the other
 * side is this library's factorised operators.  Allocates 4 x 3N complex
 * doubles of scratch for the call.                                           */
fame_status fame_check_eigendecomp(fame_ctx* ctx, int ik, unsigned long long seed, double err[4]);

/* The sparse curl C (3N x 3N) of FAME_Matrix_Curl (FAME_Main/FAME_Matrix_Curl.cpp:
 * 90-443) for one wave vector k (3 doubles): 12N entries (host arrays rows,
 * cols, and vals as interleaved complex doubles), in no particular order; the
 * same rule as the stencil of fame_check_eigendecomp.                        */
fame_status fame_curl_coo(const fame_lattice* lat, const double k[3], int* rows, int* cols, double* vals);

/* fame_residuals without any device allocation, for callers that cannot
 * allocate (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The fields are staged
 * in a file-scope device array of 8M complex doubles (128 MiB per device, held
 * until the process ends), so 3N may not exceed 8,388,608 (N <= 2,796,202;
 * larger grids return FAME_ERR_UNSUPPORTED); wave vectors are processed in
 * chunks that fit.  Same arguments and results as fame_residuals.  Every call
 * on one device shares that array: calls that overlap on one GPU (other host
 * threads or streams) race.                                                  */
fame_status fame_residuals_noalloc(fame_ctx* ctx, int nwant, const double* freq, const void* ev_dev, double* res);

/* Vector-Jacobian product of the eigenvalues lambda_kj = omega_kj^2 with respect
 * to the permittivity of every Yee edge, by Hellmann-Feynman:
 *   grad_e = sum_kj w_kj d lambda_kj / d eps_e,
 *   d lambda / d eps_e = -lambda |E_e|^2 / (E^H B E)
 * for fields E_dev as produced by fame_eigs (nk*nwant*3N).  Inside a degenerate
 * cluster (relative gap <= 1e-8) the weights are replaced by their mean, i.e. the
 * derivative of the cluster trace.  weights: host nk*nwant; grad_dev: device 3N
 * doubles (reference layout), overwritten.  No reference counterpart: this is
 * the adjoint the band solver adds for inverse design.                        */
fame_status fame_eig_grad(fame_ctx* ctx, int nwant, const double* freq, const void* E_dev, const double* weights,
                          void* grad_dev);

/* Average durations in microseconds (hipEvent timing on the context stream) of
 * the five QBQ passes and of a whole QBQ over all wave vectors of the context,
 * us6 = {inner expand, middle inverse, outer eps, middle forward, inner contract,
 * whole QBQ}; prec = 8 (FP64) or 4 (FP32).  Measurement helper, no reference
 * counterpart.                                                                 */
fame_status fame_time_qbq_passes(fame_ctx* ctx, int prec, int reps, double* us6);

/* Average durations in microseconds (hipEvent timing on the context stream) of
 * the kernels of one fused CG iteration with every wave vector active:
 * us8 = {inner expand with the x and p updates, middle inverse, outer eps with
 * the <p,Ap> sums, alpha, middle forward, inner contract with the r update and
 * <r,r> sums, beta, whole iteration}; prec = 8 or 4.  FAME_ERR_UNSUPPORTED when
 * an axis takes a Bluestein plan (that iteration form stores Ap).  Measurement
 * helper, no reference counterpart.                                          */
fame_status fame_time_cg_passes(fame_ctx* ctx, int prec, int reps, double* us8);

/* Autotuner: time every feasible (line-transform engine, lines per block) of the
 * inner, middle and outer pass kernels of precision prec (8 or 4) on the
 * current batch and keep the fastest.  fame_eigs runs it once per context and
 * precision unless FAME_AUTOTUNE=0.  choice (host, 6 ints, may be NULL)
 * receives {engine, lines} for the inner, middle and outer passes (engine 0
 * Stockham/Bluestein, 1 direct DFT on the matrix cores).  Replaces the fixed
 * cuFFT plans of FAME_Create_cublas (FAME_Main/FAME_Create_cublas.cu:16-25).  */
fame_status fame_autotune(fame_ctx* ctx, int prec, int* choice);

/* Set the pass tiles of precision prec explicitly, in the format of
 * fame_autotune's choice (e.g. to reproduce a tuned configuration); marks the
 * precision as tuned.  FAME_ERR_ARG when an engine or line count is not
 * available for this grid.  No reference counterpart.                       */
fame_status fame_set_tiles(fame_ctx* ctx, int prec, const int* choice);

/* Stream used by all launches (a hipStream_t passed as void*; NULL = default). */
fame_status fame_set_stream(fame_ctx* ctx, void* stream);

/* Problem sizes: N and the Stockham/Bluestein lengths of the three axes. */
fame_status fame_query(const fame_ctx* ctx, int* N, int* nk, int* fft_len3);

/* Host routines of the Lanczos restart (no device needed).
 * fame_tridiag_eig: all eigenpairs of the symmetric tridiagonal (d[0..n-1],
 * e[0..n-2]); w receives the eigenvalues in descending order, z (n x n,
 * column-major) the eigenvectors.  Replaces LAPACKE_zpteqr(compz='I')
 * (FAME_Main/Lanczos_Isotropic.cu:48-53).
 * fame_tridiag_qr_step: one explicit shifted QR step T - mu I = QR, T <- RQ + mu I,
 * in place on (diag, off); the rotation annihilating T(i+1,i) against the pivot
 * a = T(i,i) is c[i] = a/r, s[i] = -T(i+1,i)/r, r = hypot(a, T(i+1,i)); the new
 * off-diagonal is the mean of RQ's sub- and super-diagonal.  ru and rw are
 * scratch of n doubles.  Replaces GVqrrq_g (FAME_Main/Lanczos_LockPurge_gpu.cu:95,245).
 * Both return FAME_ERR_ARG for n < 1 or missing arrays.                        */
int fame_tridiag_eig(int n, const double* d, const double* e, double* w, double* z);
int fame_tridiag_qr_step(double* diag, double* off, double* ru, double* rw, double* c, double* s, double mu, int n);

/* Message of the last error on this thread. */
const char* fame_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
