/*
 * ref_driver.cpp -- a thin C-ABI around the reference's own HOST code
 * (FAME_Preprocessing, FAME_Matrix_Curl), compiled from the reference sources
 * where they lie by oracle/build_ref.sh into oracle/_ref/.  Test
 * infrastructure only: it lets tests/ compare this project's restated input
 * generation and sparse curl with the reference's, bit for bit.  Nothing of it
 * is shipped or linked by the product library.
 */
#include "FAME_Internal_Common.h"
#include "FAME_Preprocessing.h"
#include "FAME_Malloc_mtx_C.h"
#include "FAME_Matrix_Curl.h"
#include <unistd.h>
#include <fcntl.h>

static PAR g_par;
static int g_have = 0;

extern "C" {

typedef struct {
    int grid_nums[3];
    double edge_len[3], mesh_len[3];
    double lattice_vec_a[9], lattice_vec_a_orig[9], Omega[9];
    int Permutation[3];
    double length_a[3], theta[3];
    int m1, m2, m3, m4;
    double t1[3], t2[3], t3[3], t4[3];
    char flag[6];
    double recip_b[9];
    int material_num;
    int wave_vec_num;
    double a, b, c, alpha, beta, gamma;
} ref_out;

/* Run the reference's FAME_Preprocessing in `workdir` (which holds Popt.txt and
 * FAME_Material_Data/<name>.txt).  Its console output goes to /dev/null. */
int ref_preprocess(const char* workdir, ref_out* o) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd))) return -1;
    if (chdir(workdir) != 0) return -2;
    fflush(stdout);
    int saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    dup2(devnull, 1);
    memset(&g_par, 0, sizeof(g_par));
    FAME_Preprocessing(&g_par);
    fflush(stdout);
    dup2(saved, 1); close(saved); close(devnull);
    if (chdir(cwd) != 0) return -3;
    g_have = 1;
    LATTICE* L = &g_par.lattice;
    memcpy(o->grid_nums, g_par.mesh.grid_nums, sizeof(o->grid_nums));
    memcpy(o->edge_len, g_par.mesh.edge_len, sizeof(o->edge_len));
    memcpy(o->mesh_len, g_par.mesh.mesh_len, sizeof(o->mesh_len));
    memcpy(o->lattice_vec_a, L->lattice_vec_a, sizeof(o->lattice_vec_a));
    memcpy(o->lattice_vec_a_orig, L->lattice_vec_a_orig, sizeof(o->lattice_vec_a_orig));
    memcpy(o->Omega, L->Omega, sizeof(o->Omega));
    memcpy(o->Permutation, L->Permutation, sizeof(o->Permutation));
    o->length_a[0] = L->lattice_constant.length_a1;
    o->length_a[1] = L->lattice_constant.length_a2;
    o->length_a[2] = L->lattice_constant.length_a3;
    o->theta[0] = L->lattice_constant.theta_1;
    o->theta[1] = L->lattice_constant.theta_2;
    o->theta[2] = L->lattice_constant.theta_3;
    o->m1 = L->lattice_constant.m1; o->m2 = L->lattice_constant.m2;
    o->m3 = L->lattice_constant.m3; o->m4 = L->lattice_constant.m4;
    memcpy(o->t1, L->lattice_constant.t1, 3 * sizeof(double));
    memcpy(o->t2, L->lattice_constant.t2, 3 * sizeof(double));
    memcpy(o->t3, L->lattice_constant.t3, 3 * sizeof(double));
    memcpy(o->t4, L->lattice_constant.t4, 3 * sizeof(double));
    memcpy(o->flag, L->lattice_constant.flag, sizeof(o->flag));
    memcpy(o->recip_b, g_par.recip_lattice.reciprocal_lattice_vector_b, sizeof(o->recip_b));
    o->material_num = g_par.material.material_num;
    o->wave_vec_num = g_par.recip_lattice.Wave_vec_num;
    o->a = L->lattice_constant.a; o->b = L->lattice_constant.b; o->c = L->lattice_constant.c;
    o->alpha = L->lattice_constant.alpha; o->beta = L->lattice_constant.beta; o->gamma = L->lattice_constant.gamma;
    return 0;
}

/* Binout of the last preprocessing: 3 * material_num * N ints */
int ref_binout(int* out) {
    if (!g_have) return -1;
    int N = g_par.mesh.grid_nums[0] * g_par.mesh.grid_nums[1] * g_par.mesh.grid_nums[2];
    memcpy(out, g_par.material.Binout, sizeof(int) * 3 * N * g_par.material.material_num);
    return 0;
}

/* WaveVector of the last preprocessing: 3 * Wave_vec_num doubles */
int ref_wave_vectors(double* out) {
    if (!g_have) return -1;
    memcpy(out, g_par.recip_lattice.WaveVector, sizeof(double) * 3 * g_par.recip_lattice.Wave_vec_num);
    return 0;
}

/* The reference's sparse curl for the last preprocessing at wave vector k:
 * 12N (row, col, complex value) triplets. */
int ref_curl(const double* k, int* rows, int* cols, double* vals) {
    if (!g_have) return -1;
    int N = g_par.mesh.grid_nums[0] * g_par.mesh.grid_nums[1] * g_par.mesh.grid_nums[2];
    MTX_C C;
    FAME_Malloc_mtx_C(&C, N);
    double kk[3] = {k[0], k[1], k[2]};
    FAME_Matrix_Curl(&C, kk, g_par.mesh.grid_nums, g_par.mesh.edge_len, g_par.mesh.mesh_len, g_par.lattice);
    memcpy(rows, C.C_r, sizeof(int) * 12 * N);
    memcpy(cols, C.C_c, sizeof(int) * 12 * N);
    memcpy(vals, C.C_v, sizeof(double) * 2 * 12 * N);
    free(C.C1_r); free(C.C1_c); free(C.C1_v); free(C.C2_r); free(C.C2_c); free(C.C2_v);
    free(C.C3_r); free(C.C3_c); free(C.C3_v); free(C.C_r); free(C.C_c); free(C.C_v);
    return 12 * N;
}

}
