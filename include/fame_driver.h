/*
 * fame_driver.h -- the reference's input files and its top-level program on
 * top of the solver: Popt.txt and material-file readers, the grid auto-sizing
 * of Popt's "Grid nums Max" flag, and a Main.cpp-style run that reads a
 * directory, builds the PAR and calls FAME_Main_Code.
 *
 * Replaces (reference interface, file:line):
 *   FAME_Set_User_Option          FAME_Preprocessing/FAME_Set_User_Option.cpp:4
 *   FAME_Set_User_Define          FAME_Preprocessing/FAME_Set_User_Define.cpp:4
 *   grid_nums_adjust (flag_Max)   FAME_Preprocessing/FAME_Parameter_Generator.cpp:106-148
 *   FAME_Preprocessing            FAME_Preprocessing/FAME_Preprocessing.cpp:13
 *   main                          Main.cpp:8-49
 */
#ifndef FAME_DRIVER_H
#define FAME_DRIVER_H

#include "fame_geometry.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FAME_POPT_MAX_EPS 64

/* The fields of Popt.txt, by the reference's section titles. */
typedef struct {
    int    grid_nums[3];                      /* # Grid number                        */
    int    part_num;                          /* # Part number                        */
    char   data_name[50];                     /* # Material data name                 */
    char   material_type[50];                 /* # Material type                      */
    int    num_ele_permitt_in;                /* # Permittivity (inner material)      */
    double ele_permitt_in[FAME_POPT_MAX_EPS];
    double ele_permitt_out;                   /* # Permittivity (outer material)      */
    int    nwant;                             /* # Desired eigenpair number           */
    int    nstep;                             /* # Dimension of Krylov subspace ...   */
    double es_tol;                            /* # Tolerance of Lanczos               */
    int    es_maxit;                          /* # Maximum restart number of Lanczos  */
    double ls_tol;                            /* # Tolerance of linear solver         */
    int    ls_maxit;                          /* # Maximum iteration number of ...    */
    int    device, printf_user_option, printf_parameter, create_parameter, create_B_inout,
           create_wave_vector, save_eigen_vector;
    int    grid_nums_max, mem_size;           /* # Grid nums Max: flag, memory (GB)    */
    int    radius_adjustment;                 /* # Sphere/Cylinder radius adjustment   */
    double radius;
} fame_popt;

/* Reads Popt.txt.  Sections may come in any order; unknown lines are skipped,
 * as by the reference.  Returns 0, -1 if the file cannot be opened, -2 if a
 * section's values cannot be read, -3 for more than FAME_POPT_MAX_EPS
 * permittivities. */
int fame_read_popt(const char* path, fame_popt* popt);

/* The fields of a FAME_Material_Data/<name>.txt file.  Arrays are allocated by
 * fame_read_material and released by fame_material_file_free. */
typedef struct {
    char    data_name[50];
    char    lattice_type[50];
    double  a, b, c, alpha, beta, gamma;
    int     material_num;
    int*    sphere_num;               /* material_num */
    double* sphere_centers;           /* 3 * total spheres, fractional */
    double* sphere_radius;
    int*    cylinder_num;             /* material_num */
    double* cylinder_top_centers;     /* 3 * total cylinders */
    double* cylinder_bot_centers;
    double* cylinder_radius;
} fame_material_file;

/* Reads a material data file.  With radius_adjustment set, every sphere and
 * cylinder radius is `radius` (the reference's Popt option).  Returns 0, -1 if
 * the file cannot be opened, -2 on a malformed section. */
int fame_read_material(const char* path, int radius_adjustment, double radius, fame_material_file* mat);
void fame_material_file_free(fame_material_file* mat);

/* Nx after the reference's memory-based grid auto-sizing for a lattice whose
 * constants are given in `in` (grid_nums[0] the starting Nx): Nx moves in
 * steps of 4 to the largest value whose Nx*Ny*Nz fits the point budget of
 * mem_size GB with nwant/nstep Lanczos vectors (Ny, Nz from the grid rule).
 * Returns the new Nx, or -1 for an unknown lattice type or if no Nx >= 4 fits. */
int fame_grid_nums_max(const fame_lattice_in* in, int nwant, int nstep, int mem_size);

/* The reference program: reads <dir>/Popt.txt and
 * <dir>/FAME_Material_Data/<name>.txt, builds the lattice, grid (auto-sized when
 * the Popt flag asks), Yee-edge material mask and the default Brillouin-zone
 * path with Popt's part number, runs FAME_Main_Code and writes
 * Data_Profile.txt, all output files going to <dir>.  Returns 0, or a negative
 * reader error / positive fame_status. */
int fame_run(const char* dir);

#ifdef __cplusplus
}
#endif
#endif
