/*
 * fame_driver.cpp -- the reference's input files and top-level program on top
 * of the solver (include/fame_driver.h): Popt.txt and material-file readers
 * (FAME_Set_User_Option.cpp:4-143, FAME_Set_User_Define.cpp:4-123), the
 * memory-based grid auto-sizing (FAME_Parameter_Generator.cpp:106-148) and the
 * Main.cpp:8-49 run: read, preprocess, FAME_Main_Code, Data_Profile.txt.
 *
 * The readers follow the reference's file semantics: a section is a line that
 * equals its title, its values are read with scanf conversions from the lines
 * that follow, sections may come in any order and other lines are skipped.
 */
#include "fame_driver.h"
#include "fame_main.h"

#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

/* one line without its line end and trailing blanks; false at end of file */
bool read_title(FILE* fp, std::string& line) {
    char buf[512];
    if (!std::fgets(buf, sizeof(buf), fp)) return false;
    size_t n = std::strlen(buf);
    while (n > 0 && (buf[n - 1] == '\n' || buf[n - 1] == '\r' || buf[n - 1] == ' ' || buf[n - 1] == '\t')) --n;
    line.assign(buf, n);
    return true;
}

bool scan_ints(FILE* fp, int* v, int n) {
    for (int i = 0; i < n; ++i)
        if (std::fscanf(fp, "%d", &v[i]) != 1) return false;
    return true;
}
bool scan_doubles(FILE* fp, double* v, int n) {
    for (int i = 0; i < n; ++i)
        if (std::fscanf(fp, "%lf", &v[i]) != 1) return false;
    return true;
}
bool scan_word(FILE* fp, char* dst, size_t cap) {
    char buf[256];
    if (std::fscanf(fp, "%255s", buf) != 1) return false;
    const size_t n = std::strlen(buf);
    if (n >= cap) return false;              /* the reference's fields hold 49 characters */
    std::memcpy(dst, buf, n + 1);
    return true;
}

}  // namespace

extern "C" int fame_read_popt(const char* path, fame_popt* P) {
    if (!path || !P) return -2;
    std::memset(P, 0, sizeof(*P));
    FILE* fp = std::fopen(path, "r");
    if (!fp) return -1;
    std::string t;
    bool ok = true;
    while (ok && read_title(fp, t)) {
        if (t == "# Grid number") ok = scan_ints(fp, P->grid_nums, 3);
        else if (t == "# Part number") ok = scan_ints(fp, &P->part_num, 1);
        else if (t == "# Material data name") ok = scan_word(fp, P->data_name, sizeof(P->data_name));
        else if (t == "# Material type") ok = scan_word(fp, P->material_type, sizeof(P->material_type));
        else if (t == "# Permittivity (inner material)") {
            ok = scan_ints(fp, &P->num_ele_permitt_in, 1);
            if (ok && (P->num_ele_permitt_in < 0 || P->num_ele_permitt_in > FAME_POPT_MAX_EPS)) { std::fclose(fp); return -3; }
            ok = ok && scan_doubles(fp, P->ele_permitt_in, P->num_ele_permitt_in);
        }
        else if (t == "# Permittivity (outer material)") ok = scan_doubles(fp, &P->ele_permitt_out, 1);
        else if (t == "# Desired eigenpair number") ok = scan_ints(fp, &P->nwant, 1);
        else if (t == "# Dimension of Krylov subspace for Lanczos") ok = scan_ints(fp, &P->nstep, 1);
        else if (t == "# Tolerance of Lanczos") ok = scan_doubles(fp, &P->es_tol, 1);
        else if (t == "# Maximum restart number of Lanczos") ok = scan_ints(fp, &P->es_maxit, 1);
        else if (t == "# Tolerance of linear solver") ok = scan_doubles(fp, &P->ls_tol, 1);
        else if (t == "# Maximum iteration number of linear solver") ok = scan_ints(fp, &P->ls_maxit, 1);
        else if (t == "# Device") ok = scan_ints(fp, &P->device, 1);
        else if (t == "# Printf User Option") ok = scan_ints(fp, &P->printf_user_option, 1);
        else if (t == "# Printf Parameter") ok = scan_ints(fp, &P->printf_parameter, 1);
        else if (t == "# Create Parameter") ok = scan_ints(fp, &P->create_parameter, 1);
        else if (t == "# Create B_inout") ok = scan_ints(fp, &P->create_B_inout, 1);
        else if (t == "# Create Wave Vector") ok = scan_ints(fp, &P->create_wave_vector, 1);
        else if (t == "# Save Eigen Vector") ok = scan_ints(fp, &P->save_eigen_vector, 1);
        else if (t == "# Grid nums Max") ok = scan_ints(fp, &P->grid_nums_max, 1) && scan_ints(fp, &P->mem_size, 1);
        else if (t == "# Sphere/Cylinder radius adjustment")
            ok = scan_ints(fp, &P->radius_adjustment, 1) && scan_doubles(fp, &P->radius, 1);
    }
    std::fclose(fp);
    return ok ? 0 : -2;
}

extern "C" void fame_material_file_free(fame_material_file* M) {
    if (!M) return;
    std::free(M->sphere_num); std::free(M->sphere_centers); std::free(M->sphere_radius);
    std::free(M->cylinder_num); std::free(M->cylinder_top_centers); std::free(M->cylinder_bot_centers);
    std::free(M->cylinder_radius);
    std::memset(M, 0, sizeof(*M));
}

extern "C" int fame_read_material(const char* path, int radius_adjustment, double radius, fame_material_file* M) {
    if (!path || !M) return -2;
    std::memset(M, 0, sizeof(*M));
    FILE* fp = std::fopen(path, "r");
    if (!fp) return -1;
    std::string t;
    bool ok = true;
    int ns = 0, nc = 0;
    auto alloc_d = [](int n) { return (double*)std::calloc(n > 0 ? n : 1, sizeof(double)); };
    while (ok && read_title(fp, t)) {
        if (t == "# Data name") ok = scan_word(fp, M->data_name, sizeof(M->data_name));
        else if (t == "# Material number") {
            ok = scan_ints(fp, &M->material_num, 1) && M->material_num >= 0 && M->material_num < 1000000;
            if (ok) {
                std::free(M->sphere_num); std::free(M->cylinder_num);
                M->sphere_num = (int*)std::calloc(M->material_num + 1, sizeof(int));
                M->cylinder_num = (int*)std::calloc(M->material_num + 1, sizeof(int));
            }
        }
        else if (t == "# Lattice type") ok = scan_word(fp, M->lattice_type, sizeof(M->lattice_type));
        else if (t == "# Lattice constant") {
            double v[6];
            ok = scan_doubles(fp, v, 6);
            M->a = v[0]; M->b = v[1]; M->c = v[2]; M->alpha = v[3]; M->beta = v[4]; M->gamma = v[5];
        }
        else if (t == "# Sphere number") {
            ok = M->sphere_num && scan_ints(fp, M->sphere_num, M->material_num);
            ns = 0;
            for (int i = 0; ok && i < M->material_num; ++i) { ok = M->sphere_num[i] >= 0; ns += M->sphere_num[i]; }
            if (ok) {
                std::free(M->sphere_centers); std::free(M->sphere_radius);
                M->sphere_centers = alloc_d(3 * ns); M->sphere_radius = alloc_d(ns);
            }
        }
        else if (t == "# Sphere centers") ok = M->sphere_centers && scan_doubles(fp, M->sphere_centers, 3 * ns);
        else if (t == "# Sphere radius") {
            ok = M->sphere_radius != nullptr;
            if (ok && radius_adjustment) for (int i = 0; i < ns; ++i) M->sphere_radius[i] = radius;
            else if (ok) ok = scan_doubles(fp, M->sphere_radius, ns);
        }
        else if (t == "# Cylinder number") {
            ok = M->cylinder_num && scan_ints(fp, M->cylinder_num, M->material_num);
            nc = 0;
            for (int i = 0; ok && i < M->material_num; ++i) { ok = M->cylinder_num[i] >= 0; nc += M->cylinder_num[i]; }
            if (ok) {
                std::free(M->cylinder_top_centers); std::free(M->cylinder_bot_centers); std::free(M->cylinder_radius);
                M->cylinder_top_centers = alloc_d(3 * nc); M->cylinder_bot_centers = alloc_d(3 * nc);
                M->cylinder_radius = alloc_d(nc);
            }
        }
        else if (t == "# Cylinder top centers") ok = M->cylinder_top_centers && scan_doubles(fp, M->cylinder_top_centers, 3 * nc);
        else if (t == "# Cylinder bot centers") ok = M->cylinder_bot_centers && scan_doubles(fp, M->cylinder_bot_centers, 3 * nc);
        else if (t == "# Cylinder radius") {
            ok = M->cylinder_radius != nullptr;
            if (ok && radius_adjustment) for (int i = 0; i < nc; ++i) M->cylinder_radius[i] = radius;
            else if (ok) ok = scan_doubles(fp, M->cylinder_radius, nc);
        }
    }
    std::fclose(fp);
    if (!ok) { fame_material_file_free(M); return -2; }
    /* files without cylinders (or spheres) still get their count arrays */
    if (!M->sphere_num) M->sphere_num = (int*)std::calloc(M->material_num + 1, sizeof(int));
    if (!M->cylinder_num) M->cylinder_num = (int*)std::calloc(M->material_num + 1, sizeof(int));
    return 0;
}

/* Grid auto-sizing (FAME_Parameter_Generator.cpp:106-148): a budget of grid
 * points from the memory size in GB and the bytes the solver keeps per point
 * for nwant + nstep + 1 Lanczos vectors, then a walk of Nx in steps of 4:
 * upwards while the grid fits, one step back at the first that does not;
 * downwards from a start that does not fit, to the first that does. */
extern "C" int fame_grid_nums_max(const fame_lattice_in* in, int nwant, int nstep, int mem_size) {
    if (!in || nwant < 1 || nstep < 1 || mem_size < 1 || mem_size > 2047) return -1;
    const size_t per_point = 768 + 32 * (size_t)(nwant + nstep + 1) + 128;
    const size_t a = (size_t)std::floor((double)(mem_size * 1024 * 1024) / (double)per_point);
    const size_t b = (size_t)std::floor((double)(nstep * nstep * 16) / (double)per_point);
    if (a * 1024 <= b) return -1;
    const long long budget = (long long)(a * 1024 - b);
    fame_lattice_in q = *in;
    fame_lattice_out o;
    auto points = [&](int nx) -> long long {
        q.grid_nums[0] = nx;
        if (fame_lattice_build(&q, &o) != 0) return -1;
        return (long long)o.grid_nums[0] * o.grid_nums[1] * o.grid_nums[2];
    };
    int nx = in->grid_nums[0];
    bool down = false;
    for (int guard = 0; guard < 100000; ++guard) {
        if (nx < 4) return -1;
        const long long n = points(nx);
        if (n < 0) return -1;
        if (n <= budget) {
            if (down) return nx;
            nx += 4;
        } else {
            nx -= 4;
            down = true;
        }
    }
    return -1;
}

namespace {

struct CwdGuard {
    std::string old;
    bool ok = false;
    explicit CwdGuard(const char* dir) {
        char buf[4096];
        if (getcwd(buf, sizeof(buf))) old = buf;
        ok = !old.empty() && chdir(dir) == 0;
    }
    ~CwdGuard() {
        if (ok && chdir(old.c_str()) != 0) std::fprintf(stderr, "fame_run: cannot return to %s\n", old.c_str());
    }
};

}  // namespace

extern "C" int fame_run(const char* dir) {
    if (!dir) return -2;
    CwdGuard cwd(dir);
    if (!cwd.ok) return -1;
    const auto t0 = std::chrono::steady_clock::now();
    fame_popt P;
    int rc = fame_read_popt("Popt.txt", &P);
    if (rc != 0) { std::fprintf(stderr, "fame_run: cannot read %s/Popt.txt (%d)\n", dir, rc); return rc; }
    const std::string mpath = std::string("./FAME_Material_Data/") + P.data_name + ".txt";
    fame_material_file M;
    rc = fame_read_material(mpath.c_str(), P.radius_adjustment, P.radius, &M);
    if (rc != 0) { std::fprintf(stderr, "fame_run: cannot read material data %s (%d)\n", mpath.c_str(), rc); return rc; }

    fame_lattice_in li;
    std::memset(&li, 0, sizeof(li));
    std::snprintf(li.lattice_type, sizeof(li.lattice_type), "%s", M.lattice_type);
    li.a = M.a; li.b = M.b; li.c = M.c; li.alpha = M.alpha; li.beta = M.beta; li.gamma = M.gamma;
    for (int i = 0; i < 3; ++i) li.grid_nums[i] = P.grid_nums[i];
    if (P.grid_nums_max) {
        const int nx = fame_grid_nums_max(&li, P.nwant, P.nstep, P.mem_size);
        if (nx < 0) { fame_material_file_free(&M); std::fprintf(stderr, "fame_run: grid auto-sizing failed\n"); return FAME_ERR_ARG; }
        li.grid_nums[0] = nx;
    }
    fame_lattice_out lo;
    if (fame_lattice_build(&li, &lo) != 0) {
        fame_material_file_free(&M);
        std::fprintf(stderr, "fame_run: unknown lattice type %s\n", M.lattice_type);
        return FAME_ERR_ARG;
    }
    std::printf("Grid_nums = [ %3d %3d %3d ]\n", lo.grid_nums[0], lo.grid_nums[1], lo.grid_nums[2]);

    /* wave vectors: the default path of the lattice with Popt's part number */
    const char* path = fame_bz_default_path(&lo);
    if (!path || P.part_num < 2) {
        fame_material_file_free(&M);
        std::fprintf(stderr, "fame_run: no Brillouin-zone path for %s\n", lo.lattice_type);
        return FAME_ERR_ARG;
    }
    const int nk = fame_bz_path_count(path, P.part_num);
    std::vector<double> W(3 * (size_t)nk);
    if (fame_bz_path(&lo, path, P.part_num, W.data()) != nk) { fame_material_file_free(&M); return FAME_ERR_ARG; }

    /* Yee-edge material mask as the reference's Binout (first material containing the point) */
    const int N = lo.grid_nums[0] * lo.grid_nums[1] * lo.grid_nums[2];
    const int nm = M.material_num;
    fame_material_in mi = {nm, M.sphere_num, M.sphere_centers, M.sphere_radius, M.cylinder_num,
                           M.cylinder_top_centers, M.cylinder_bot_centers, M.cylinder_radius};
    std::vector<int> idx(3 * (size_t)N), binout(3 * (size_t)nm * N, 0);
    if (fame_material_locate(&lo, &mi, idx.data()) != 0) { fame_material_file_free(&M); return FAME_ERR_ARG; }
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < N; ++i) {
            const int m = idx[(size_t)c * N + i];
            if (m >= 0) binout[(size_t)c * nm * N + (size_t)m * N + i] = 1;
        }

    PAR par;
    std::memset(&par, 0, sizeof(par));
    par.flag.device = P.device; par.flag.printf_user_option = P.printf_user_option;
    par.flag.printf_parameter = P.printf_parameter; par.flag.create_parameter = P.create_parameter;
    par.flag.create_B_inout = P.create_B_inout; par.flag.create_wave_vector = P.create_wave_vector;
    par.flag.save_eigen_vector = P.save_eigen_vector; par.flag.grid_nums_max = P.grid_nums_max;
    par.flag.mem_size = P.mem_size; par.flag.radius_adjustment = P.radius_adjustment; par.flag.radius = P.radius;
    for (int i = 0; i < 3; ++i) {
        par.mesh.grid_nums[i] = lo.grid_nums[i];
        par.mesh.edge_len[i] = lo.edge_len[i];
        par.mesh.mesh_len[i] = lo.mesh_len[i];
    }
    LATTICE& L = par.lattice;
    std::snprintf(L.lattice_type, sizeof(L.lattice_type), "%s", lo.lattice_type);
    for (int i = 0; i < 3; ++i) L.Permutation[i] = lo.Permutation[i];
    std::memcpy(L.lattice_vec_a, lo.lattice_vec_a, sizeof(L.lattice_vec_a));
    std::memcpy(L.lattice_vec_a_orig, lo.lattice_vec_a_orig, sizeof(L.lattice_vec_a_orig));
    std::memcpy(L.Omega, lo.Omega, sizeof(L.Omega));
    LATTICE_CONSTANT& C = L.lattice_constant;
    C.a = lo.a; C.b = lo.b; C.c = lo.c; C.alpha = lo.alpha; C.beta = lo.beta; C.gamma = lo.gamma;
    C.theta_1 = lo.theta[0]; C.theta_2 = lo.theta[1]; C.theta_3 = lo.theta[2];
    C.m1 = lo.m1; C.m2 = lo.m2; C.m3 = lo.m3; C.m4 = lo.m4;
    double t1[3], t2[3], t3[3], t4[3];
    std::memcpy(t1, lo.t1, sizeof(t1)); std::memcpy(t2, lo.t2, sizeof(t2));
    std::memcpy(t3, lo.t3, sizeof(t3)); std::memcpy(t4, lo.t4, sizeof(t4));
    C.t1 = t1; C.t2 = t2; C.t3 = t3; C.t4 = t4;
    std::memcpy(C.flag, lo.flag, sizeof(C.flag));
    C.length_a1 = lo.length_a[0]; C.length_a2 = lo.length_a[1]; C.length_a3 = lo.length_a[2];
    MATERIAL& Mt = par.material;
    std::snprintf(Mt.data_name, sizeof(Mt.data_name), "%s", M.data_name[0] ? M.data_name : P.data_name);
    std::snprintf(Mt.material_type, sizeof(Mt.material_type), "%s", P.material_type);
    Mt.Binout = binout.data();
    Mt.material_num = nm;
    Mt.sphere_num = M.sphere_num; Mt.sphere_centers = M.sphere_centers; Mt.sphere_radius = M.sphere_radius;
    Mt.cylinder_num = M.cylinder_num; Mt.cylinder_top_centers = M.cylinder_top_centers;
    Mt.cylinder_bot_centers = M.cylinder_bot_centers; Mt.cylinder_radius = M.cylinder_radius;
    Mt.num_ele_permitt_in = P.num_ele_permitt_in;
    Mt.ele_permitt_in = P.ele_permitt_in;
    Mt.ele_permitt_out = P.ele_permitt_out;
    Mt.flag_radius_adjustment = P.radius_adjustment; Mt.flag_radius = P.radius;
    RECIP_LATTICE& R = par.recip_lattice;
    R.part_num = P.part_num; R.Wave_vec_num = nk; R.WaveVector = W.data();
    std::snprintf(R.path_string, sizeof(R.path_string), "%s", path);
    std::memcpy(R.reciprocal_lattice_vector_b, lo.recip_b, sizeof(R.reciprocal_lattice_vector_b));
    double* vtx = &R.vertex.A[0];
    const char letters[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZstk";
    for (int i = 0; i < 29; ++i)
        if (fame_bz_vertex(&lo, letters[i], vtx + 3 * i) != 0) vtx[3 * i] = vtx[3 * i + 1] = vtx[3 * i + 2] = 0.0;
    par.ls.maxit = P.ls_maxit; par.ls.tol = P.ls_tol;
    par.es.nwant = P.nwant; par.es.nstep = P.nstep; par.es.maxit = P.es_maxit; par.es.tol = P.es_tol;

    /* FAME_Set_Profile (FAME_Tools/FAME_Profile.cpp:3-10) */
    std::vector<int> ls_iter(nk, 0), es_iter(nk, 0);
    std::vector<double> ls_time(nk, 0.0), es_time(nk, 0.0);
    PROFILE prof = {0, ls_iter.data(), es_iter.data(), ls_time.data(), es_time.data()};
    const double t_pre = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const auto t1c = std::chrono::steady_clock::now();
    rc = fame_main_code_run(&par, &prof);
    const double t_main = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1c).count();
    if (rc == 0) FAME_Create_Profile_txt(prof);
    std::printf("preprocessing %.2f s, FAME_Main_Code %.2f s, total %.2f s\n", t_pre, t_main, t_pre + t_main);
    fame_material_file_free(&M);
    return rc;
}
