/*
 * fame_geometry.h -- host-side input generation for the band solver:
 * lattice vectors in the computational frame, grid adjustment, skew boundary
 * points, reciprocal lattice, Brillouin-zone paths and the permittivity of
 * every Yee edge.  This is the C-ABI restatement of the reference's
 * FAME_Preprocessing stage; its outputs are exactly the inputs that
 * FAME_Main_Code consumes, so a PAR built here reproduces the reference's
 * discretisation bit for bit.
 *
 * Replaces (reference interface, file:line):
 *   FAME_Parameter_Lattice_Constants_Format   FAME_Preprocessing/FAME_Parameter_Lattice_Constants_Format.cpp:4
 *   FAME_Parameter_Lattice_Vector (+QRp)      FAME_Preprocessing/FAME_Parameter_Lattice_Vector.cpp:7,382
 *   grid_nums_adjust                          FAME_Preprocessing/FAME_Parameter_Generator.cpp:106
 *   FAME_Parameter_Boundary_Point             FAME_Preprocessing/FAME_Parameter_Boundary_Point.cpp:4
 *   FAME_Parameter_Recip_Lattice_Vector       FAME_Preprocessing/FAME_Parameter_Recip_Lattice_Vector.cpp:4
 *   FAME_Parameter_Brillouin_Zone_Point/_Path FAME_Preprocessing/FAME_Parameter_Brillouin_Zone_Path.cpp:9
 *   FAME_Material_Locate_Index                FAME_Preprocessing/FAME_Material_Locate_Index.cpp:101
 *   FAME_Matrix_B_Isotropic (eps per edge)    FAME_Main/FAME_Matrix_B_Isotropic.cu:4
 */
#ifndef FAME_GEOMETRY_H
#define FAME_GEOMETRY_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    char   lattice_type[50];   /* "simple_cubic", "face_centered_cubic", ... */
    double a, b, c, alpha, beta, gamma;    /* as read from the material file */
    int    grid_nums[3];       /* Nx is used; Ny, Nz are derived             */
} fame_lattice_in;

typedef struct {
    char   lattice_type[50];
    int    general;            /* 0 Simple, 1 General (FAME_Main_Code.cu:158-160) */
    int    grid_nums[3];
    double edge_len[3];
    double mesh_len[3];
    double a, b, c, alpha, beta, gamma;    /* after FAME_Parameter_Lattice_Constants_Format */
    double lattice_vec_a[9];   /* computational frame                        */
    double lattice_vec_a_orig[9];
    double Omega[9];
    int    Permutation[3];
    double length_a[3];
    double theta[3];
    int    m1, m2, m3, m4;
    double t1[3], t2[3], t3[3], t4[3];
    char   flag[6];
    double recip_b[9];         /* reciprocal_lattice_vector_b = inv(a_orig)^T */
} fame_lattice_out;

/* Geometry of the inclusions in fractional coordinates of lattice_vec_a_orig */
typedef struct {
    int     material_num;
    const int*    sphere_num;        /* material_num */
    const double* sphere_centers;    /* 3 * total spheres */
    const double* sphere_radius;
    const int*    cylinder_num;      /* material_num */
    const double* cylinder_top_centers;
    const double* cylinder_bot_centers;
    const double* cylinder_radius;
} fame_material_in;

/* returns 0 on success, -1 for an unknown lattice type */
int fame_lattice_build(const fame_lattice_in* in, fame_lattice_out* out);

/* idx3N: for each Yee edge (component-stacked, x-fastest), the index of the
 * first material containing it, or -1 (FAME_Material_Locate_Index.cpp:171-183) */
int fame_material_locate(const fame_lattice_out* lat, const fame_material_in* mat, int* idx3N);

/* eps per Yee edge from material indices (FAME_Matrix_B_Isotropic.cu:33-62);
 * eps_in has material_num entries.  Returns -1 if fewer eps values than
 * materials are given (the reference falls back to an interactive prompt). */
int fame_material_eps(int N, int material_num, const int* idx3N, const double* eps_in,
                      int num_eps_in, double eps_out, double* eps3N);

/* Brillouin-zone vertex (letter) in the ORIGINAL frame, e.g. 'G','X','W','L' */
int fame_bz_vertex(const fame_lattice_out* lat, char name, double v[3]);
/* Map an original-frame wave vector to the computational frame (BZ_Path.cpp:80-90) */
void fame_wave_vec_rotate(const fame_lattice_out* lat, const double v[3], double k[3]);
/* Number of points of the reference path rule for a path string ("GXWKGLUWLK|UX")
 * and part_num (BZ_Path.cpp:283-298). */
int fame_bz_path_count(const char* path, int part_num);
/* Fill 3*count wave vectors (computational frame) along a path string; returns count
 * or -1 for an unknown vertex letter (BZ_Path.cpp:9-93). */
int fame_bz_path(const fame_lattice_out* lat, const char* path, int part_num, double* wave_vectors);
/* The reference's default path string for a lattice type (BZ_Path.cpp:166-267) */
const char* fame_bz_default_path(const fame_lattice_out* lat);

#ifdef __cplusplus
}
#endif
#endif
