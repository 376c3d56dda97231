/*
 * fame_main.h -- source-level drop-in for the reference's top entry points.
 *
 * The structs carry the same field names and meanings as the reference's
 * call-surface types (include/FAME_Internal_Common.h:62-211 in the reference
 * tree) so that a caller that fills a PAR for the reference can call this
 * library's FAME_Main_Code unchanged.  Device handles (CULIB_HANDLES,
 * LAMBDAS_CUDA, MTX_B, ...) do not appear: the context of include/fame_hip.h
 * owns everything on the device.
 */
#ifndef FAME_MAIN_H
#define FAME_MAIN_H

#include "fame_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

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
 * Par.flag.save_eigen_vector is set) and fills Profile like the reference; a frequency
 * whose absolute residual (reference field scaling) exceeds 1e-10 is negated
 * (FAME_Main_Code.cu:383-421).  Returns 0, or a fame_status on failure.      */
int FAME_Main_Code(PAR Par, PROFILE* Profile);

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
}
#endif
#endif
