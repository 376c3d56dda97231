/*
 * fame_main.h -- source- and binary-level drop-in for the reference's top entry points.
 *
 * The structs carry the same field names, types and order as the reference's
 * call-surface types (include/FAME_Internal_Common.h:29-211 in the reference
 * tree), so their layouts are identical (tests/test_layout.py compares every
 * offset with a table compiled from the reference header) and a caller that
 * fills a PAR for the reference can pass it to this library's FAME_Main_Code
 * unchanged.  FAME_Main_Code has C++ linkage in C++, as the reference declares
 * it (include/FAME_Main_Code.cuh:4), and a C-linkage symbol of the same name for
 * C and FFI callers.  Device handles (CULIB_HANDLES, LAMBDAS_CUDA, MTX_B, ...)
 * do not appear: the context of include/fame_hip.h owns everything on the device.
 */
#ifndef FAME_MAIN_H
#define FAME_MAIN_H

#include "fame_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Brillouin-zone vertices of the lattice, by letter, in units of 2 pi
 * (FAME_Internal_Common.h:29-59; filled by FAME_Parameter_Brillouin_Zone_Point.cpp) */
typedef struct {
    double A[3], B[3], C[3], D[3], E[3], F[3], G[3], H[3], I[3], J[3], K[3], L[3], M[3];
    double N[3], O[3], P[3], Q[3], R[3], S[3], T[3], U[3], V[3], W[3], X[3], Y[3], Z[3];
    double s[3], t[3], k[3];
} VERTEX;

typedef struct { int grid_nums[3]; double edge_len[3]; double mesh_len[3]; } MESH;

typedef struct {
    double a, b, c, alpha, beta, gamma;
    double theta_1, theta_2, theta_3;
    int m1, m2, m3, m4;
    double *t1, *t2, *t3, *t4;
    char flag[6];
    double length_a1, length_a2, length_a3;
} LATTICE_CONSTANT;

typedef struct {
    char data_name[50];
    char material_type[50];
    int* Binout;                 /* 3 * material_num * N, x-fastest per Yee set */
    int material_num;
    int* sphere_num;
    double* sphere_centers;
    double* sphere_radius;
    int* cylinder_num;
    double* cylinder_top_centers;
    double* cylinder_bot_centers;
    double* cylinder_radius;
    int num_ele_permitt_in;
    double* ele_permitt_in;
    double ele_permitt_out;
    int flag_radius_adjustment;
    double flag_radius;
} MATERIAL;

typedef struct {
    int part_num;
    int Wave_vec_num;
    double* WaveVector;          /* 3 * Wave_vec_num, computational frame, units of 2 pi */
    char path_string[50];
    double reciprocal_lattice_vector_b[9];
    VERTEX vertex;
} RECIP_LATTICE;

typedef struct {
    char lattice_type[50];
    int Permutation[3];
    double lattice_vec_a[9];
    double lattice_vec_a_orig[9];
    double Omega[9];
    LATTICE_CONSTANT lattice_constant;
} LATTICE;

typedef struct { int maxit; double tol; } LS;
typedef struct { int nwant; int nstep; int maxit; double tol; } ES;

typedef struct {
    int device, printf_user_option, printf_parameter, create_parameter, create_B_inout, create_wave_vector,
        save_eigen_vector, grid_nums_max, mem_size, radius_adjustment;
    double radius;
} FLAG;

typedef struct {
    FLAG flag;
    MESH mesh;
    LATTICE lattice;
    MATERIAL material;
    RECIP_LATTICE recip_lattice;
    LS ls;
    ES es;
} PAR;

typedef struct {
    int idx;
    int* ls_iter;
    int* es_iter;
    double* ls_time;
    double* es_time;
} PROFILE;

/* Replaces FAME_Main_Code (FAME_Main/FAME_Main_Code.cu:37): all wave vectors of
 * Par.recip_lattice, in batches sized to the free device memory, FP64 (or the
 * mixed-precision solver when FAME_PRECISION=mixed is set in the environment).
 * Writes Data_Omega_array.txt (and Data_EV_kk_bb.txt per band when
 * Par.flag.save_eigen_vector is set) and fills Profile like the reference
 * (ls_time: seconds inside the invAr solves, es_time: seconds of the eigensolve,
 * per wave vector); a frequency whose absolute residual (reference field
 * scaling) exceeds 1e-10 is negated (FAME_Main_Code.cu:383-421).  Returns 0, or
 * a fame_status on failure.  Declared below with the linkage of the includer. */
int fame_main_code_run(const PAR* Par, PROFILE* Profile);

/* Writes Data_Profile.txt for rows 0..Profile.idx, in the format of the
 * reference's FAME_Create_Profile_txt (FAME_Tools/FAME_Profile.cpp:27-39),
 * which its Main.cpp calls after FAME_Main_Code.  Returns 0, or -1.        */
int FAME_Create_Profile_txt(PROFILE Profile);

/* Replaces Lanczos_Isotropic (FAME_Main/Lanczos_Isotropic.cu:7) for the wave
 * vectors already set in ctx: Freq_array (host, nk*nwant) receives omega =
 * sqrt(1/theta), ev_dev (optional, device) the Ritz vectors with a 2N stride per
 * band.  flag_CompType ("Simple"/"General") must match the context's lattice.
 * Returns the largest restart count, or -status on failure.                  */
int Lanczos_Isotropic(double* Freq_array, void* ev_dev, fame_ctx* ctx, ES es, LS ls, int Nx, int Ny, int Nz,
                      const char* flag_CompType, PROFILE* Profile);

#ifdef __cplusplus
}  /* extern "C" */
/* C++ linkage, as include/FAME_Main_Code.cuh:4 of the reference declares it */
int FAME_Main_Code(PAR Par, PROFILE* Profile);
#else
/* C linkage alias of the same entry point, for C and FFI callers */
int FAME_Main_Code(PAR Par, PROFILE* Profile);
#endif
#endif
