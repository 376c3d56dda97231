/*
 * fame_oracle.h -- FP64 CPU restatement of the FAME-Isotropic-GPU eigensolver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path.  It may be loaded only by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg -- never by the product library.
 *
 * Every routine cites the reference file:line whose behaviour it restates.
 * Layout conventions are the reference's:
 *   - real-space fields: component-stacked [Ex;Ey;Ez], each x-fastest,
 *     r = x + Nx*(y + Ny*z)            (FAME_Material_Locate_Index.cpp:175)
 *   - Fourier index, Simple lattices: x-fastest, same as real space
 *                                      (FAME_Matrix_Lambdas.cu:96,116,132)
 *   - Fourier index, General lattices: z-fastest, f = (mx*Ny+my)*Nz+mz
 *                                      (FAME_Matrix_Lambdas.cu:317,345,371)
 *   - reduced vectors [x1;x2], each of length Nd, mode i <-> f = N-Nd+i
 *                                      (FAME_Matrix_Vector_Production_Qr.cu:33-35)
 * Complex arrays are interleaved (re,im) doubles.
 */
#ifndef FAME_ORACLE_H
#define FAME_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int    N[3];        /* grid_nums                                        */
    int    general;     /* 0 = Simple (SC/primitive ortho/tetra), 1 = General */
    double h[3];        /* mesh_len                                         */
    double a[9];        /* lattice_vec_a, computational frame, a1|a2|a3      */
    int    m1, m2, m3, m4;
    double t1[3], t2[3], t3[3], t4[3];
} orc_lattice;

typedef struct { int nwant, nstep, maxit; double tol; } orc_es;
typedef struct { int maxit; double tol; } orc_ls;

typedef struct orc_tables orc_tables;

/* ---- FFT (own mixed-radix Stockham + Bluestein) -------------------------- */
/* in-place unnormalised DFT of `howmany` contiguous lines of length n;
 * sign = -1: exp(-2 pi i jk/n) (cuFFT FORWARD), +1: exp(+...) (INVERSE)     */
void orc_fft(int n, int sign, double* data, int howmany);
void orc_set_threads(int nthreads);
int  orc_get_threads(void);

/* ---- per-k tables (FAME_Matrix_Lambdas.cu:40-554) ------------------------- */
orc_tables* orc_tables_create(const orc_lattice* lat, const double k[3]);
void        orc_tables_destroy(orc_tables* t);
int         orc_tables_nd(const orc_tables* t);
/* which: 0 Lambda_x, 1 Lambda_y, 2 Lambda_z (Nd complex each), 3 Lambda_q_sqrt (Nd real),
 *        4 Pi_Qr, 5 Pi_Pr, 6 Pi_Qrs, 7 Pi_Prs (6Nd complex)                  */
void        orc_tables_get(const orc_tables* t, int which, double* out);

/* ---- operators (FAME_Matrix_Vector_Production_*.cu) ------------------------ */
void orc_T(const orc_tables* t, const double* u, double* y);       /* one component, Fourier -> real */
void orc_Tstar(const orc_tables* t, const double* v, double* y);   /* one component, real -> Fourier */
void orc_qr (const orc_tables* t, const double* x, double* y);     /* 2Nd -> 3N */
void orc_qrs(const orc_tables* t, const double* v, double* y);     /* 3N -> 2Nd (v untouched) */
void orc_pr (const orc_tables* t, const double* x, double* y);
void orc_prs(const orc_tables* t, const double* v, double* y);
void orc_qbq(const orc_tables* t, const double* invB, const double* x, double* y);
/* C E and C^H C E via the Fourier factorisation C_i = T Lambda_i T*          */
void orc_curl(const orc_tables* t, const double* E, double* CE);
void orc_curlh(const orc_tables* t, const double* F, double* CHF);
int  orc_cg (const orc_tables* t, const double* invB, const double* rhs, double* x,
             int maxit, double tol, double* final_res);
int  orc_invar(const orc_tables* t, const double* invB, const double* x, double* y,
               int maxit, double tol);

/* ---- explicit sparse curl (FAME_Matrix_Curl.cpp:90-443) -------------------- */
/* rows/cols: 12N ints, vals: 12N complex; returns nnz (=12N)                  */
int orc_curl_coo(const orc_lattice* lat, const double k[3], int* rows, int* cols, double* vals);

/* ---- small dense pieces ---------------------------------------------------- */
/* symmetric tridiagonal eigen-decomposition, eigenvalues DESCENDING (zpteqr order),
 * z: n x n column-major eigenvectors.  Returns 0 on success.                  */
int  orc_tridiag_eig(int n, const double* d, const double* e, double* w, double* z);
/* one exact-shift QR step of Lanczos_LockPurge_gpu.cu:95-190 (in place)       */
void orc_gvqrrq(double* T0, double* T1, double* T2, double* T3, double* c, double* s,
                double shift, int n);

/* ---- full solve (FAME_Fast_Algorithms_Isotropic.cu + Lanczos_Isotropic.cu) - */
typedef struct {
    int ls_iter;    /* total CG iterations (Profile.ls_iter)                  */
    int es_iter;    /* invAr applies (Profile.es_iter)                         */
    int restarts;   /* Lanczos_Isotropic return value                         */
    int nd;
} orc_profile;

/* eps: 3N permittivity on Yee edges (B_eps layout).  freq: nwant.
 * ev: optional 3N x nwant complex E fields (reference scaling quirk kept).
 * Returns 0.                                                                  */
int orc_eigs(const orc_lattice* lat, const double* eps, const double k[3],
             const orc_es* es, const orc_ls* ls, double* freq, double* ev,
             orc_profile* prof);
/* relative residual ||C^H C E - w^2 B E|| / ||E|| for one field (3N complex)  */
double orc_rel_residual(const orc_tables* t, const double* eps, const double* E, double omega);

#ifdef __cplusplus
}
#endif
#endif
