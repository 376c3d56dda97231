/*
 * fame_geometry.cpp -- input generation for the band solver (host C++).
 *
 * Restates the reference's preprocessing so that the lattice, grid, skew
 * boundary points, wave vectors and Yee-edge permittivity handed to the solver
 * are the reference's, bit for bit.  Arithmetic is written in the same order
 * as the reference so that rounding ties (round(), floor(), <= r^2) fall the
 * same way; each block cites the reference lines it follows.
 */
#include "fame_geometry.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

namespace {

const double PI = 3.141592653589793;   /* include/FAME_Internal_Common.h:24 */

bool is(const char* a, const char* b) { return std::strcmp(a, b) == 0; }

double inner3(const double* a, const double* b) {
    double s = 0;
    for (int i = 0; i < 3; ++i) s += a[i] * b[i];
    return s;
}
double norm_n(const double* x, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i] * x[i];
    return std::sqrt(s);
}

/* column-major 3x3 inverse, FAME_Tools/inv3.cpp:8-23 */
void inv3(const double* A, double* r) {
    double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[3] * (A[1] * A[8] - A[7] * A[2]) + A[6] * (A[1] * A[5] - A[4] * A[2]);
    double invdet = 1 / det;
    r[0] = (A[4] * A[8] - A[5] * A[7]) * invdet;
    r[1] = -(A[1] * A[8] - A[2] * A[7]) * invdet;
    r[2] = (A[1] * A[5] - A[2] * A[4]) * invdet;
    r[3] = -(A[3] * A[8] - A[6] * A[5]) * invdet;
    r[4] = (A[0] * A[8] - A[6] * A[2]) * invdet;
    r[5] = -(A[0] * A[5] - A[2] * A[3]) * invdet;
    r[6] = (A[3] * A[7] - A[6] * A[4]) * invdet;
    r[7] = -(A[0] * A[7] - A[1] * A[6]) * invdet;
    r[8] = (A[0] * A[4] - A[1] * A[3]) * invdet;
}
/* transpose of the inverse, FAME_Tools/inv3.cpp:25-41 */
void inv3_T(const double* A, double* r) {
    double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[3] * (A[1] * A[8] - A[7] * A[2]) + A[6] * (A[1] * A[5] - A[4] * A[2]);
    double invdet = 1 / det;
    r[0] = (A[4] * A[8] - A[5] * A[7]) * invdet;
    r[1] = -(A[3] * A[8] - A[6] * A[5]) * invdet;
    r[2] = (A[3] * A[7] - A[6] * A[4]) * invdet;
    r[3] = -(A[1] * A[8] - A[2] * A[7]) * invdet;
    r[4] = (A[0] * A[8] - A[6] * A[2]) * invdet;
    r[5] = -(A[0] * A[7] - A[1] * A[6]) * invdet;
    r[6] = (A[1] * A[5] - A[2] * A[4]) * invdet;
    r[7] = -(A[0] * A[5] - A[2] * A[3]) * invdet;
    r[8] = (A[0] * A[4] - A[1] * A[3]) * invdet;
}
/* column-major product, FAME_Tools/mtx_prod.cpp:10-22 */
void mtx_prod(double* ans, const double* M1, const double* M2, int m, int n, int p) {
    for (int i = 0; i < m; i++)
        for (int j = 0; j < p; j++) {
            ans[j * m + i] = 0.0;
            for (int k = 0; k < n; k++) ans[j * m + i] += M1[k * m + i] * M2[j * n + k];
        }
}

/* FAME_Parameter_Lattice_Vector.cpp:366-380 */
void householder(double* w, double* c, const double* b, int size) {
    double normb = norm_n(b, size);
    if (std::fabs(b[0]) < 1e-14) c[0] = normb;
    else c[0] = -b[0] / std::fabs(b[0]) * normb;
    double coe = std::sqrt(2 * normb * (normb + std::fabs(b[0])));
    std::memcpy(w, b, size * sizeof(double));
    w[0] = w[0] - c[0];
    for (int ii = 0; ii < size; ii++) w[ii] = w[ii] / coe;
}

/* FAME_Parameter_Lattice_Vector.cpp:382-445 */
void QRp(double* v, int* perm) {
    double len[3], w[3], c, temp;
    len[0] = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    len[1] = v[3] * v[3] + v[4] * v[4] + v[5] * v[5];
    len[2] = v[6] * v[6] + v[7] * v[7] + v[8] * v[8];
    int mx = 0;
    if (len[1] > len[0] && len[1] > len[2]) mx = 1;
    else if (len[2] > len[0] && len[2] > len[1]) mx = 2;
    if (mx != 0) {
        for (int ii = 0; ii < 3; ii++) { temp = v[mx * 3 + ii]; v[mx * 3 + ii] = v[ii]; v[ii] = temp; }
        temp = len[0]; len[0] = len[mx]; len[mx] = temp;
        perm[0] = mx + 1;
        perm[mx] = 1;
    }
    householder(w, &c, &v[0], 3);
    v[0] = c; v[1] = 0.0; v[2] = 0.0;
    temp = inner3(w, &v[3]);
    for (int ii = 0; ii < 3; ii++) v[ii + 3] = v[ii + 3] - 2 * temp * w[ii];
    temp = inner3(w, &v[6]);
    for (int ii = 0; ii < 3; ii++) v[ii + 6] = v[ii + 6] - 2 * temp * w[ii];
    if (len[1] - v[3] * v[3] < len[2] - v[6] * v[6]) {
        int it = perm[1]; perm[1] = perm[2]; perm[2] = it;
        for (int ii = 0; ii < 3; ii++) { temp = v[3 + ii]; v[3 + ii] = v[6 + ii]; v[6 + ii] = temp; }
    }
    householder(w, &c, &v[4], 2);
    v[4] = c; v[5] = 0.0;
    temp = w[0] * v[7] + w[1] * v[8];
    for (int ii = 0; ii < 2; ii++) v[ii + 7] = v[ii + 7] - 2 * temp * w[ii];
}

bool is_simple(const char* t) {
    return is(t, "simple_cubic") || is(t, "primitive_orthorhombic") || is(t, "primitive_tetragonal");
}

}  // namespace

extern "C" int fame_lattice_build(const fame_lattice_in* in, fame_lattice_out* o) {
    std::memset(o, 0, sizeof(*o));
    const char* t = in->lattice_type;
    std::snprintf(o->lattice_type, sizeof(o->lattice_type), "%s", t);
    double a = in->a, b = in->b, c = in->c, al = in->alpha, be = in->beta, ga = in->gamma;
    /* FAME_Parameter_Lattice_Constants_Format.cpp:4-62 */
    if (is(t, "simple_cubic") || is(t, "face_centered_cubic") || is(t, "body_centered_cubic")) {
        b = a; c = a; al = PI / 2; be = PI / 2; ga = PI / 2;
    } else if (is(t, "hexagonal")) {
        b = a; al = PI / 2; be = PI / 2; ga = 2 * PI / 3;
    } else if (is(t, "rhombohedral")) {
        b = a; c = a; be = al; ga = al;
    } else if (is(t, "primitive_tetragonal") || is(t, "body_centered_tetragonal")) {
        b = a; al = PI / 2; be = PI / 2; ga = PI / 2;
    } else if (is(t, "primitive_orthorhombic") || is(t, "a_base_centered_orthorhombic") ||
               is(t, "c_base_centered_orthorhombic") || is(t, "face_centered_orthorhombic") ||
               is(t, "body_centered_orthorhombic")) {
        al = PI / 2; be = PI / 2; ga = PI / 2;
    } else if (is(t, "primitive_monoclinic") || is(t, "base_centered_monoclinic")) {
        be = PI / 2; ga = PI / 2;
    } else if (is(t, "triclinic")) {
    } else {
        return -1;
    }
    o->a = a; o->b = b; o->c = c; o->alpha = al; o->beta = be; o->gamma = ga;

    /* FAME_Parameter_Lattice_Vector.cpp:13-253 */
    double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0}, a3[3] = {0, 0, 0};
    auto set = [&](double x0, double x1, double x2, double y0, double y1, double y2, double z0, double z1, double z2) {
        a1[0] = x0; a1[1] = x1; a1[2] = x2; a2[0] = y0; a2[1] = y1; a2[2] = y2; a3[0] = z0; a3[1] = z1; a3[2] = z2;
    };
    if (is(t, "simple_cubic")) set(a, 0, 0, 0, a, 0, 0, 0, a);
    else if (is(t, "face_centered_cubic")) { double h = 0.5 * a; set(0, h, h, h, 0, h, h, h, 0); }
    else if (is(t, "body_centered_cubic")) { double h = 0.5 * a; set(-1 * h, h, h, h, -1 * h, h, h, h, -1 * h); }
    else if (is(t, "primitive_tetragonal")) set(a, 0, 0, 0, a, 0, 0, 0, c);
    else if (is(t, "body_centered_tetragonal")) { double h = 0.5 * a, hc = 0.5 * c; set(-h, h, hc, h, -h, hc, h, h, -hc); }
    else if (is(t, "primitive_orthorhombic")) set(a, 0, 0, 0, b, 0, 0, 0, c);
    else if (is(t, "face_centered_orthorhombic")) { double x = 0.5 * a, y = 0.5 * b, z = 0.5 * c; set(0, y, z, x, 0, z, x, y, 0); }
    else if (is(t, "body_centered_orthorhombic")) { double x = 0.5 * a, y = 0.5 * b, z = 0.5 * c; set(-x, y, z, x, -y, z, x, y, -z); }
    else if (is(t, "a_base_centered_orthorhombic")) set(a, 0, 0, 0, 0.5 * b, -0.5 * c, 0, 0.5 * b, 0.5 * c);
    else if (is(t, "c_base_centered_orthorhombic")) set(0.5 * a, -0.5 * b, 0, 0.5 * a, 0.5 * b, 0, 0, 0, c);
    else if (is(t, "hexagonal")) set(a, 0, 0, -0.5 * a, std::sqrt(3) * 0.5 * a, 0, 0, 0, c);
    else if (is(t, "rhombohedral"))
        set(a * std::cos(al / 2), -a * std::sin(al / 2), 0, a * std::cos(al / 2), a * std::sin(al / 2), 0,
            a * std::cos(al) / std::cos(al / 2), 0, a * std::sqrt(1.0 - (std::pow(std::cos(al) / std::cos(al / 2), 2))));
    else if (is(t, "primitive_monoclinic")) set(a, 0, 0, 0, b, 0, 0, c * std::cos(al), c * std::sin(al));
    else if (is(t, "base_centered_monoclinic")) {
        double x = 0.5 * a, y = 0.5 * b;
        set(x, y, 0, -x, y, 0, 0, c * std::cos(al), c * std::sin(al));
    } else if (is(t, "triclinic"))
        set(a, 0, 0, b * std::cos(ga), b * std::sin(ga), 0, c * std::cos(be),
            c * (std::cos(al) - std::cos(be) * std::cos(ga)) / std::sin(ga),
            c * (std::sqrt(1 - std::pow(std::cos(al), 2) - std::pow(std::cos(be), 2) - std::pow(std::cos(ga), 2) +
                           2 * std::cos(al) * std::cos(be) * std::cos(ga)) / std::sin(ga)));
    double* L = o->lattice_vec_a;
    for (int i = 0; i < 3; ++i) { L[i] = a1[i]; L[i + 3] = a2[i]; L[i + 6] = a3[i]; o->Permutation[i] = i + 1; }

    /* :264-347 */
    std::memcpy(o->lattice_vec_a_orig, L, sizeof(double) * 9);
    o->general = is_simple(t) ? 0 : 1;
    if (o->general) QRp(L, o->Permutation);
    double l1 = 0, l2 = 0, l3 = 0;
    for (int i = 0; i < 3; ++i) { l1 += std::pow(L[i], 2); l2 += std::pow(L[i + 3], 2); l3 += std::pow(L[i + 6], 2); }
    l1 = std::sqrt(l1); l2 = std::sqrt(l2); l3 = std::sqrt(l3);
    double th1 = 0, th2 = 0, th3 = 0;
    for (int i = 0; i < 3; ++i) {
        th1 += L[i + 3] * L[i + 6] / (l2 * l3);
        th2 += L[i] * L[i + 6] / (l1 * l3);
        th3 += L[i] * L[i + 3] / (l1 * l2);
    }
    th1 = std::acos(th1); th2 = std::acos(th2); th3 = std::acos(th3);
    if (std::fabs(th1 - PI / 2) < 1e-14) th1 = PI / 2;
    if (std::fabs(th2 - PI / 2) < 1e-14) th2 = PI / 2;
    if (std::fabs(th3 - PI / 2) < 1e-14) th3 = PI / 2;
    if (is(t, "base_centered_monoclinic")) {
        if (th1 >= PI / 2) th1 = PI - th1;
        if (th2 >= PI / 2) th2 = PI - th2;
        if (th3 >= PI / 2) th3 = PI - th3;
    }
    o->length_a[0] = l1; o->length_a[1] = l2; o->length_a[2] = l3;
    o->theta[0] = th1; o->theta[1] = th2; o->theta[2] = th3;
    double ll2 = (std::cos(th1) - std::cos(th3) * std::cos(th2)) / std::sin(th3);
    double ll3 = std::sqrt(std::pow(std::sin(th2), 2) - std::pow(ll2, 2));
    L[0] = l1; L[1] = 0.0; L[2] = 0.0;
    L[3] = l2 * std::cos(th3); L[4] = l2 * std::sin(th3); L[5] = 0.0;
    L[6] = l3 * std::cos(th2); L[7] = l3 * ll2; L[8] = l3 * ll3;
    /* :349-361 */
    double P[9], invA[9];
    for (int r = 0; r < 3; ++r)
        for (int i = 0; i < 3; ++i) P[3 * r + i] = o->lattice_vec_a_orig[(o->Permutation[r] - 1) * 3 + i];
    inv3(L, invA);
    mtx_prod(o->Omega, P, invA, 3, 3, 3);

    /* FAME_Parameter_Generator.cpp:108-114, grid_nums_adjust :172-179 */
    o->edge_len[0] = L[0]; o->edge_len[1] = L[4]; o->edge_len[2] = L[8];
    int* g = o->grid_nums;
    g[0] = in->grid_nums[0];
    double* h = o->mesh_len;
    h[0] = o->edge_len[0] / g[0];
    g[1] = (int)std::round(o->edge_len[1] / h[0]);
    h[1] = o->edge_len[1] / g[1];
    g[2] = (int)std::round(o->edge_len[2] / h[0]);
    h[2] = o->edge_len[2] / g[2];

    /* FAME_Parameter_Boundary_Point.cpp:4-288 */
    double R3R8, R10R3 = 0.0;
    if ((th3 > 0) && (th3 <= PI / 2)) R3R8 = L[3];
    else R3R8 = L[0] + L[3];
    double *t1 = o->t1, *t2 = o->t2, *t3 = o->t3, *t4 = o->t4;
    auto lin = [&](double* dst, double ca1, double ca2) { for (int i = 0; i < 3; ++i) dst[i] = ca1 * L[i] + ca2 * L[i + 3]; };
    const bool acute = (th3 > 0) && (th3 <= PI / 2);
    const bool obtuse = (th3 > PI / 2) && (th3 < PI);
    if ((L[6] > 0) && (L[7] > 0)) {
        R10R3 = L[7];
        lin(t1, 0, 0); lin(t2, -1, 0);
        if (acute && L[3] <= L[6]) { std::strcpy(o->flag, "1-i"); lin(t3, -1, -1); lin(t4, 0, -1); }
        else if (acute && L[3] > L[6]) { std::strcpy(o->flag, "1-ii"); lin(t3, 0, -1); lin(t4, 1, -1); }
        else if (obtuse && -1 * L[3] <= L[0] - L[6]) { std::strcpy(o->flag, "1-iii"); lin(t3, -1, -1); lin(t4, 0, -1); }
        else if (obtuse && -1 * L[3] > L[0] - L[6]) { std::strcpy(o->flag, "1-iv"); lin(t3, -2, -1); lin(t4, -1, -1); }
    }
    if ((L[6] < 0) && (L[7] > 0)) {
        R10R3 = L[7];
        lin(t1, 1, 0); lin(t2, 0, 0);
        if (acute && L[3] <= L[0] + L[6]) { std::strcpy(o->flag, "2-i"); lin(t3, 0, -1); lin(t4, 1, -1); }
        else if (acute && L[3] > L[0] + L[6]) { std::strcpy(o->flag, "2-ii"); lin(t3, 1, -1); lin(t4, 2, -1); }
        else if (obtuse && -1 * L[3] <= -1 * L[6]) { std::strcpy(o->flag, "2-iii"); lin(t3, 0, -1); lin(t4, 1, -1); }
        else if (obtuse && -1 * L[3] > -1 * L[6]) { std::strcpy(o->flag, "2-iv"); lin(t3, -1, -1); lin(t4, 0, -1); }
    }
    if ((L[6] < 0) && (L[7] < 0)) {
        R10R3 = L[4] + L[7];
        lin(t3, 0, 0); lin(t4, 1, 0);
        if (acute && L[3] <= -1 * L[6]) { std::strcpy(o->flag, "3-i"); lin(t1, 1, 1); lin(t2, 0, 1); }
        /* case 3-ii keeps the reference's t1 = a2 + a2 (:184) */
        else if (acute && L[3] > -1 * L[6]) { std::strcpy(o->flag, "3-ii"); lin(t1, 0, 2); lin(t2, -1, 1); }
        else if (obtuse && -1 * L[3] <= L[0] + L[6]) { std::strcpy(o->flag, "3-iii"); lin(t1, 1, 1); lin(t2, 0, 1); }
        else if (obtuse && -1 * L[3] > L[0] + L[6]) { std::strcpy(o->flag, "3-iv"); lin(t1, 2, 1); lin(t2, 1, 1); }
    }
    if ((L[6] > 0) && (L[7] < 0)) {
        R10R3 = L[4] + L[7];
        lin(t3, -1, 0); lin(t4, 0, 0);
        if (acute && L[3] <= L[0] - L[6]) { std::strcpy(o->flag, "4-i"); lin(t1, 0, 1); lin(t2, -1, 1); }
        else if (acute && L[3] > L[0] - L[6]) { std::strcpy(o->flag, "4-ii"); lin(t1, -1, 1); lin(t2, -2, 1); }
        else if (obtuse && -1 * L[3] <= L[6]) { std::strcpy(o->flag, "4-iii"); lin(t1, 0, 1); lin(t2, -1, 1); }
        else if (obtuse && -1 * L[3] > L[6]) { std::strcpy(o->flag, "4-iv"); lin(t1, 1, 1); lin(t2, 0, 1); }
    }
    o->m1 = (int)std::round(R3R8 / h[0]);
    o->m2 = (int)std::round((L[6] + t1[0]) / h[0]);
    o->m3 = (int)std::round(R10R3 / h[1]);
    o->m4 = (o->m2 - o->m1) % g[0];
    if (o->m4 < 0) o->m4 += g[0];

    /* FAME_Parameter_Recip_Lattice_Vector.cpp:4-8 */
    inv3_T(o->lattice_vec_a_orig, o->recip_b);
    return 0;
}

/* ---------------------------------------------------------------------------- */
/*  Yee-edge material mask  (FAME_Material_Locate_Index.cpp, FAME_Material_Handle.cpp) */
/* ---------------------------------------------------------------------------- */
namespace {

int sphere_handle(double px, double py, double pz, const double* A, const fame_material_in* M, int kind) {
    int sum = 0;
    for (int i = 0; i < kind; i++) sum += M->sphere_num[i];
    for (int i = sum; i < sum + M->sphere_num[kind]; i++)
        for (int sx = -1; sx <= 1; sx++)
            for (int sy = -1; sy <= 1; sy++)
                for (int sz = -1; sz <= 1; sz++) {
                    const double* c = &M->sphere_centers[i * 3];
                    double cx = (c[0] + sx) * A[0] + (c[1] + sy) * A[3] + (c[2] + sz) * A[6];
                    double cy = (c[0] + sx) * A[1] + (c[1] + sy) * A[4] + (c[2] + sz) * A[7];
                    double cz = (c[0] + sx) * A[2] + (c[1] + sy) * A[5] + (c[2] + sz) * A[8];
                    double d = std::pow(cx - px, 2) + std::pow(cy - py, 2) + std::pow(cz - pz, 2);
                    if (d <= std::pow(M->sphere_radius[i], 2)) return 1;
                }
    return 0;
}

int cylinder_handle(double px, double py, double pz, const double* A, const fame_material_in* M, int kind) {
    int sum = 0;
    for (int i = 0; i < kind; i++) sum += M->cylinder_num[i];
    for (int i = sum; i < sum + M->cylinder_num[kind]; i++) {
        bool hit = false;
        for (int sx = -1; sx <= 1 && !hit; sx++)
            for (int sy = -1; sy <= 1 && !hit; sy++)
                for (int sz = -1; sz <= 1; sz++) {
                    const double* T = &M->cylinder_top_centers[i * 3];
                    const double* B = &M->cylinder_bot_centers[i * 3];
                    double ctx = (T[0] + sx) * A[0] + (T[1] + sy) * A[3] + (T[2] + sz) * A[6];
                    double cty = (T[0] + sx) * A[1] + (T[1] + sy) * A[4] + (T[2] + sz) * A[7];
                    double ctz = (T[0] + sx) * A[2] + (T[1] + sy) * A[5] + (T[2] + sz) * A[8];
                    double cbx = (B[0] + sx) * A[0] + (B[1] + sy) * A[3] + (B[2] + sz) * A[6];
                    double cby = (B[0] + sx) * A[1] + (B[1] + sy) * A[4] + (B[2] + sz) * A[7];
                    double cbz = (B[0] + sx) * A[2] + (B[1] + sy) * A[5] + (B[2] + sz) * A[8];
                    double tt = ((px - ctx) * (cbx - ctx) + (py - cty) * (cby - cty) + (pz - ctz) * (cbz - ctz)) /
                                (std::pow(cbx - ctx, 2) + std::pow(cby - cty, 2) + std::pow(cbz - ctz, 2));
                    if (0.0 <= tt && tt <= 1.0) {
                        double d = std::pow(ctx + tt * (cbx - ctx) - px, 2) + std::pow(cty + tt * (cby - cty) - py, 2) +
                                   std::pow(ctz + tt * (cbz - ctz) - pz, 2);
                        if (d <= std::pow(M->cylinder_radius[i], 2)) { hit = true; break; }
                    }
                }
        /* the reference's `break` leaves only the innermost loop (Material_Handle.cpp:63-67),
         * so a hit is final for this material either way */
        if (hit) return 1;
    }
    return 0;
}

}  // namespace

extern "C" int fame_material_locate(const fame_lattice_out* lat, const fame_material_in* mat, int* idx3N) {
    const int Nx = lat->grid_nums[0], Ny = lat->grid_nums[1], Nz = lat->grid_nums[2];
    const int n = Nx * Ny * Nz;
    const double* Om = lat->Omega;
    const double* A = lat->lattice_vec_a;
    /* FAME_Material_Locate_Index.cpp:131-168 */
    double AP[9], invAP[9], Aorig[9];
    int count = 0;
    for (int i = 0; i < 3; i++) {
        AP[i] = Om[count] * A[0] + Om[count + 1] * A[1] + Om[count + 2] * A[2];
        AP[i + 3] = Om[count] * A[3] + Om[count + 1] * A[4] + Om[count + 2] * A[5];
        AP[i + 6] = Om[count] * A[6] + Om[count + 1] * A[7] + Om[count + 2] * A[8];
        count += 3;
    }
    inv3(AP, invAP);
    int invP[3];
    const int* P = lat->Permutation;
    if (P[0] == 2 && P[1] == 3 && P[2] == 1) { invP[0] = 3; invP[1] = 1; invP[2] = 2; }
    else if (P[0] == 3 && P[1] == 1 && P[2] == 2) { invP[0] = 2; invP[1] = 3; invP[2] = 1; }
    else { invP[0] = P[0]; invP[1] = P[1]; invP[2] = P[2]; }
    for (int r = 0; r < 3; ++r)
        for (int i = 0; i < 3; ++i) Aorig[3 * r + i] = AP[(invP[r] - 1) * 3 + i];
    const double* h = lat->mesh_len;
    for (int comp = 0; comp < 3; ++comp) {
        double sx = comp == 0 ? 0.5 : 0.0, sy = comp == 1 ? 0.5 : 0.0, sz = comp == 2 ? 0.5 : 0.0;
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) {
                    int idx = k * Nx * Ny + j * Nx + i;
                    double pt[3] = {(i + sx) * h[0], (j + sy) * h[1], (k + sz) * h[2]};   /* FAME_Matrix_Grid.cpp:51-53 */
                    double po[3], coef[3];
                    po[0] = pt[0] * Om[0] + pt[1] * Om[1] + pt[2] * Om[2];
                    po[1] = pt[0] * Om[3] + pt[1] * Om[4] + pt[2] * Om[5];
                    po[2] = pt[0] * Om[6] + pt[1] * Om[7] + pt[2] * Om[8];
                    coef[0] = po[0] * invAP[0] + po[1] * invAP[3] + po[2] * invAP[6];
                    coef[1] = po[0] * invAP[1] + po[1] * invAP[4] + po[2] * invAP[7];
                    coef[2] = po[0] * invAP[2] + po[1] * invAP[5] + po[2] * invAP[8];
                    double s1 = std::floor(coef[0]), s2 = std::floor(coef[1]), s3 = std::floor(coef[2]);
                    po[0] = po[0] - (s1 * AP[0] + s2 * AP[3] + s3 * AP[6]);
                    po[1] = po[1] - (s1 * AP[1] + s2 * AP[4] + s3 * AP[7]);
                    po[2] = po[2] - (s1 * AP[2] + s2 * AP[5] + s3 * AP[8]);
                    int found = -1;
                    for (int m = 0; m < mat->material_num; m++) {
                        if (sphere_handle(po[0], po[1], po[2], Aorig, mat, m) ||
                            cylinder_handle(po[0], po[1], po[2], Aorig, mat, m)) { found = m; break; }
                    }
                    idx3N[comp * n + idx] = found;
                }
    }
    return 0;
}

extern "C" int fame_material_eps(int N, int material_num, const int* idx3N, const double* eps_in,
                                 int num_eps_in, double eps_out, double* eps3N) {
    if (material_num > num_eps_in) return -1;
    for (int i = 0; i < 3 * N; ++i) {
        int m = idx3N[i];
        eps3N[i] = m >= 0 ? eps_in[m] : eps_out;
    }
    return 0;
}

/* ---------------------------------------------------------------------------- */
/*  Brillouin zone  (FAME_Parameter_Brillouin_Zone_Point.cpp / _Path.cpp)        */
/* ---------------------------------------------------------------------------- */
namespace {

/* High-symmetry points of the Brillouin zones of the 14 Bravais lattices, in
 * units of the reciprocal basis vectors, after Setyawan & Curtarolo, Comput.
 * Mater. Sci. 49 (2010) 299-312, as the reference tabulates them
 * (FAME_Parameter_Brillouin_Zone_Point.cpp:12-345, including its variant
 * conditions).  Each coordinate is a constant plus or minus at most one of the
 * variant's parameters, written as a short expression ("0.5", "eta",
 * "1-eta", "0.5+zeta", "-mu"); evaluated as c0 + s * p, one rounding, which is
 * the reference's arithmetic.                                                */
struct BzParams { double eta, zeta, delta, phi, mu, nu, psi, rho, omega; };
struct BzPoint { char name; const char* x; const char* y; const char* z; };
struct BzSet { const BzPoint* pts; int n; };

double bz_param(const BzParams& P, const char* nm, int len) {
    auto eq = [&](const char* w) { return (int)std::strlen(w) == len && std::strncmp(nm, w, len) == 0; };
    if (eq("eta")) return P.eta;
    if (eq("zeta")) return P.zeta;
    if (eq("delta")) return P.delta;
    if (eq("phi")) return P.phi;
    if (eq("mu")) return P.mu;
    if (eq("nu")) return P.nu;
    if (eq("psi")) return P.psi;
    if (eq("rho")) return P.rho;
    if (eq("omega")) return P.omega;
    return NAN;
}

/* "[c0][+|-][param]" */
double bz_coord(const BzParams& P, const char* e) {
    const char* q = e;
    double c0 = 0.0;
    bool have_c = false;
    if (*q == '-' && (q[1] >= '0' && q[1] <= '9')) { c0 = std::strtod(q, (char**)&q); have_c = true; }
    else if (*q >= '0' && *q <= '9') { c0 = std::strtod(q, (char**)&q); have_c = true; }
    if (!*q) return c0;
    double sg = 1.0;
    if (*q == '+') ++q;
    else if (*q == '-') { sg = -1.0; ++q; }
    const double p = bz_param(P, q, (int)std::strlen(q));
    if (!have_c) return sg * p;
    return sg > 0 ? c0 + p : c0 - p;
}

#define BZN(a) (int)(sizeof(a) / sizeof(a[0]))
const BzPoint BZ_CUB[] = {{'G', "0", "0", "0"}, {'M', "0.5", "0.5", "0"}, {'R', "0.5", "0.5", "0.5"}, {'X', "0", "0.5", "0"}};
const BzPoint BZ_FCC[] = {{'G', "0", "0", "0"}, {'K', "0.375", "0.375", "0.75"}, {'L', "0.5", "0.5", "0.5"},
                          {'U', "0.625", "0.25", "0.625"}, {'W', "0.5", "0.25", "0.75"}, {'X', "0.5", "0", "0.5"}};
const BzPoint BZ_BCC[] = {{'G', "0", "0", "0"}, {'N', "0", "0", "0.5"}, {'H', "0.5", "-0.5", "0.5"},
                          {'P', "0.25", "0.25", "0.25"}};
const BzPoint BZ_TET[] = {{'G', "0", "0", "0"}, {'A', "0.5", "0.5", "0.5"}, {'M', "0.5", "0.5", "0"},
                          {'R', "0", "0.5", "0.5"}, {'X', "0", "0.5", "0"}, {'Z', "0", "0", "0.5"}};
const BzPoint BZ_BCT1[] = {{'G', "0", "0", "0"}, {'X', "0", "0", "0.5"}, {'M', "-0.5", "0.5", "0.5"},
                           {'N', "0", "0.5", "0"}, {'Z', "eta", "eta", "-eta"}, {'B', "-eta", "1-eta", "eta"},
                           {'P', "0.25", "0.25", "0.25"}};
const BzPoint BZ_BCT2[] = {{'G', "0", "0", "0"}, {'X', "0", "0", "0.5"}, {'Y', "-zeta", "zeta", "0.5"},
                           {'A', "0.5", "0.5", "-zeta"}, {'s', "-eta", "eta", "eta"}, {'t', "eta", "1-eta", "-eta"},
                           {'N', "0", "0.5", "0"}, {'P', "0.25", "0.25", "0.25"}, {'Z', "0.5", "0.5", "-0.5"}};
const BzPoint BZ_ORC[] = {{'G', "0", "0", "0"}, {'R', "0.5", "0.5", "0.5"}, {'S', "0.5", "0.5", "0"},
                          {'T', "0", "0.5", "0.5"}, {'U', "0.5", "0", "0.5"}, {'X', "0.5", "0", "0"},
                          {'Y', "0", "0.5", "0"}, {'Z', "0", "0", "0.5"}};
const BzPoint BZ_ORCF1[] = {{'G', "0", "0", "0"}, {'L', "0.5", "0.5", "0.5"}, {'A', "0.5", "0.5+zeta", "zeta"},
                            {'B', "0.5", "0.5-zeta", "1-zeta"}, {'T', "1", "0.5", "0.5"}, {'Y', "0.5", "0", "0.5"},
                            {'X', "0", "eta", "eta"}, {'J', "1", "1-eta", "1-eta"}, {'Z', "0.5", "0.5", "0"}};
const BzPoint BZ_ORCF2[] = {{'G', "0", "0", "0"}, {'L', "0.5", "0.5", "0.5"}, {'C', "0.5", "0.5-eta", "1-eta"},
                            {'E', "0.5", "0.5+eta", "eta"}, {'D', "0.5-delta", "0.5", "1-delta"},
                            {'F', "0.5+delta", "0.5", "delta"}, {'H', "1-phi", "0.5-phi", "0.5"},
                            {'I', "phi", "0.5+phi", "0.5"}, {'X', "0", "0.5", "0.5"}, {'Y', "0.5", "0", "0.5"},
                            {'Z', "0.5", "0.5", "0"}};
const BzPoint BZ_ORCI[] = {{'G', "0", "0", "0"}, {'L', "-mu", "mu", "0.5-delta"}, {'I', "mu", "-mu", "0.5+delta"},
                           {'J', "0.5-delta", "0.5+delta", "-mu"}, {'R', "0", "0.5", "0"}, {'S', "0.5", "0", "0"},
                           {'T', "0", "0", "0.5"}, {'W', "0.25", "0.25", "0.25"}, {'X', "-zeta", "zeta", "zeta"},
                           {'A', "zeta", "1-zeta", "-zeta"}, {'Y', "eta", "-eta", "eta"}, {'B', "1-eta", "eta", "-eta"},
                           {'Z', "0.5", "0.5", "-0.5"}};
const BzPoint BZ_ORCC[] = {{'G', "0", "0", "0"}, {'R', "0", "0.5", "0.5"}, {'S', "0", "0.5", "0"},
                           {'T', "-0.5", "0.5", "0.5"}, {'A', "zeta", "zeta", "0.5"}, {'B', "-zeta", "1-zeta", "0.5"},
                           {'X', "zeta", "zeta", "0"}, {'C', "-zeta", "1-zeta", "0"}, {'Y', "-0.5", "0.5", "0"},
                           {'Z', "0", "0", "0.5"}};
const BzPoint BZ_RHL1[] = {{'G', "0", "0", "0"}, {'F', "0.5", "0.5", "0"}, {'B', "eta", "0.5", "1-eta"},
                           {'C', "0.5", "1-eta", "-1+eta"}, {'L', "0.5", "0", "0"}, {'I', "0", "0", "-0.5"},
                           {'P', "eta", "nu", "nu"}, {'R', "1-nu", "1-nu", "1-eta"}, {'S', "nu", "nu", "-1+eta"},
                           {'Q', "1-nu", "nu", "0"}, {'X', "nu", "0", "-nu"}, {'Z', "0.5", "0.5", "0.5"}};
const BzPoint BZ_RHL2[] = {{'G', "0", "0", "0"}, {'F', "0.5", "-0.5", "0"}, {'L', "0.5", "0", "0"},
                           {'Z', "0.5", "-0.5", "0.5"}, {'P', "1-nu", "-nu", "1-nu"}, {'R', "nu", "-1+nu", "-1+nu"},
                           {'Q', "eta", "eta", "eta"}, {'S', "1-eta", "-eta", "-eta"}};
const BzPoint BZ_MCL[] = {{'G', "0", "0", "0"}, {'A', "0.5", "0.5", "0"}, {'C', "0", "0.5", "0.5"},
                          {'D', "0.5", "0", "0.5"}, {'F', "0.5", "0", "-0.5"}, {'E', "0.5", "0.5", "0.5"},
                          {'H', "0", "eta", "1-nu"}, {'I', "0", "1-eta", "nu"}, {'J', "0", "eta", "-nu"},
                          {'M', "0.5", "eta", "1-nu"}, {'N', "0.5", "1-eta", "nu"}, {'O', "0.5", "eta", "-nu"},
                          {'X', "0", "0.5", "0"}, {'Y', "0", "0", "0.5"}, {'B', "0", "0", "-0.5"}, {'Z', "0.5", "0", "0"}};
const BzPoint BZ_MCLC1[] = {{'G', "0", "0", "0"}, {'N', "0.5", "0", "0"}, {'S', "0", "-0.5", "0"},
                            {'F', "1-zeta", "1-zeta", "1-eta"}, {'J', "zeta", "zeta", "eta"},
                            {'K', "-zeta", "-zeta", "1-eta"}, {'O', "1-zeta", "-zeta", "1-eta"},
                            {'I', "phi", "1-phi", "0.5"}, {'R', "1-phi", "-1+phi", "0.5"}, {'L', "0.5", "0.5", "0.5"},
                            {'M', "0.5", "0", "0.5"}, {'X', "1-psi", "-1+psi", "0"}, {'T', "psi", "1-psi", "0"},
                            {'U', "-1+psi", "-psi", "0"}, {'Y', "0.5", "0.5", "0"}, {'V', "-0.5", "-0.5", "0"},
                            {'Z', "0", "0", "0.5"}};
const BzPoint BZ_MCLC3[] = {{'G', "0", "0", "0"}, {'F', "1-phi", "1-phi", "1-psi"}, {'J', "phi", "-1+phi", "psi"},
                            {'K', "1-phi", "-phi", "1-psi"}, {'H', "zeta", "zeta", "eta"},
                            {'P', "1-zeta", "-zeta", "1-eta"}, {'Q', "-zeta", "-zeta", "1-eta"},
                            {'I', "0.5", "-0.5", "0.5"}, {'M', "0.5", "0", "0.5"}, {'N', "0.5", "0", "0"},
                            {'S', "0", "-0.5", "0"}, {'X', "0.5", "-0.5", "0"}, {'Y', "mu", "mu", "delta"},
                            {'V', "1-mu", "-mu", "-delta"}, {'W', "-mu", "-mu", "-delta"}, {'A', "mu", "-1+mu", "delta"},
                            {'Z', "0", "0", "0.5"}};
const BzPoint BZ_MCLC5[] = {{'G', "0", "0", "0"}, {'F', "nu", "nu", "omega"}, {'J', "1-nu", "1-nu", "1-omega"},
                            {'K', "nu", "-1+nu", "omega"}, {'H', "zeta", "zeta", "eta"},
                            {'P', "1-zeta", "-zeta", "1-eta"}, {'Q', "-zeta", "-zeta", "1-eta"},
                            {'I', "rho", "1-rho", "0.5"}, {'R', "1-rho", "-1+rho", "0.5"}, {'L', "0.5", "0.5", "0.5"},
                            {'M', "0.5", "0", "0.5"}, {'N', "0.5", "0", "0"}, {'S', "0", "-0.5", "0"},
                            {'X', "0.5", "-0.5", "0"}, {'Y', "mu", "mu", "delta"}, {'V', "1-mu", "-mu", "-delta"},
                            {'W', "-mu", "-mu", "-delta"}, {'A', "mu", "-1+mu", "delta"}, {'Z', "0", "0", "0.5"}};
const BzPoint BZ_TRI1[] = {{'G', "0", "0", "0"}, {'L', "0.5", "0.5", "0"}, {'M', "0", "0.5", "0.5"},
                           {'N', "0.5", "0", "0.5"}, {'R', "0.5", "0.5", "0.5"}, {'X', "0.5", "0", "0"},
                           {'Y', "0", "0.5", "0"}, {'Z', "0", "0", "0.5"}};
const BzPoint BZ_TRI2[] = {{'G', "0", "0", "0"}, {'L', "0.5", "-0.5", "0"}, {'M', "0", "0", "0.5"},
                           {'N', "-0.5", "-0.5", "0.5"}, {'R', "0", "-0.5", "0.5"}, {'X', "0", "-0.5", "0"},
                           {'Y', "0.5", "0", "0"}, {'Z', "-0.5", "0", "0.5"}};

/* the point set and parameters of a lattice (variant conditions as in the
 * reference); {nullptr, 0} where the reference defines none (a = c for BCT,
 * theta_1 = pi/2 for the rhombohedral lattice, b2.b3 = 0 for triclinic) */
BzSet bz_set(const fame_lattice_out* L, BzParams& P) {
    std::memset(&P, 0, sizeof(P));
    const char* t = L->lattice_type;
    const double a = L->a, b = L->b, c = L->c, al = L->alpha;
    const double* B = L->recip_b;
    if (is(t, "simple_cubic")) return {BZ_CUB, BZN(BZ_CUB)};
    if (is(t, "face_centered_cubic")) return {BZ_FCC, BZN(BZ_FCC)};
    if (is(t, "body_centered_cubic")) return {BZ_BCC, BZN(BZ_BCC)};
    if (is(t, "primitive_tetragonal")) return {BZ_TET, BZN(BZ_TET)};
    if (is(t, "body_centered_tetragonal")) {
        if (c < a) { P.eta = (1.0 + std::pow(c / a, 2)) / 4.0; return {BZ_BCT1, BZN(BZ_BCT1)}; }
        if (c > a) {
            P.eta = (1.0 + std::pow(a / c, 2)) / 4.0;
            P.zeta = 0.5 * std::pow(a, 2) / std::pow(c, 2);
            return {BZ_BCT2, BZN(BZ_BCT2)};
        }
        return {nullptr, 0};
    }
    if (is(t, "primitive_orthorhombic")) return {BZ_ORC, BZN(BZ_ORC)};
    if (is(t, "face_centered_orthorhombic")) {
        if ((1.0 / std::pow(a, 2)) >= (1.0 / std::pow(b, 2) + 1.0 / std::pow(c, 2))) {
            P.eta = (1.0 + std::pow(a / b, 2) + std::pow(a / c, 2)) / 4.0;
            P.zeta = (1.0 + std::pow(a / b, 2) - std::pow(a / c, 2)) / 4.0;
            return {BZ_ORCF1, BZN(BZ_ORCF1)};
        }
        P.eta = (1.0 + std::pow(a / b, 2) - std::pow(a / c, 2)) / 4.0;
        P.delta = (1.0 + std::pow(b / a, 2) - std::pow(b / c, 2)) / 4.0;
        P.phi = (1.0 + std::pow(c / b, 2) - std::pow(c / a, 2)) / 4.0;
        return {BZ_ORCF2, BZN(BZ_ORCF2)};
    }
    if (is(t, "body_centered_orthorhombic")) {
        P.zeta = (1.0 + std::pow(a / c, 2)) / 4.0;
        P.eta = (1.0 + std::pow(b / c, 2)) / 4.0;
        P.delta = (std::pow(b, 2) - std::pow(a, 2)) / (4.0 * std::pow(c, 2));
        P.mu = (std::pow(a, 2) + std::pow(b, 2)) / (4.0 * std::pow(c, 2));
        return {BZ_ORCI, BZN(BZ_ORCI)};
    }
    if (is(t, "c_base_centered_orthorhombic")) {
        P.zeta = (1.0 + std::pow(a / b, 2)) / 4.0;
        return {BZ_ORCC, BZN(BZ_ORCC)};
    }
    if (is(t, "hexagonal")) return {nullptr, -1};      /* handled by the caller (k point) */
    if (is(t, "rhombohedral")) {
        const double th = L->theta[0];
        if (th < PI / 2) {
            P.eta = (1.0 + 4.0 * std::cos(th)) / (2.0 + 4.0 * std::cos(th));
            P.nu = 3.0 / 4.0 - P.eta / 2.0;
            return {BZ_RHL1, BZN(BZ_RHL1)};
        }
        if (th > PI / 2) {
            P.eta = 1.0 / (2.0 * std::pow(std::tan(th / 2), 2));
            P.nu = 3.0 / 4.0 - P.eta / 2.0;
            return {BZ_RHL2, BZN(BZ_RHL2)};
        }
        return {nullptr, 0};
    }
    if (is(t, "primitive_monoclinic")) {
        P.eta = (1.0 - b * (std::cos(al) / c)) / (2.0 * std::pow(std::sin(al), 2));
        P.nu = 1.0 / 2.0 - P.eta * c * (std::cos(al) / b);
        return {BZ_MCL, BZN(BZ_MCL)};
    }
    if (is(t, "base_centered_monoclinic")) {
        const double b12 = B[0] * B[3] + B[1] * B[4] + B[2] * B[5];
        if (b12 <= 0) {
            P.zeta = (2.0 - b * std::cos(al) / c) / (4.0 * std::pow(std::sin(al), 2));
            P.eta = 1.0 / 2.0 + 2.0 * P.zeta * c * std::cos(al) / b;
            P.psi = 3.0 / 4.0 - std::pow(a, 2) / (4.0 * std::pow(b, 2) * std::pow(std::sin(al), 2));
            P.phi = P.psi + (3.0 / 4.0 - P.psi) * b * std::cos(al) / c;
            return {BZ_MCLC1, BZN(BZ_MCLC1)};
        }
        const double q = b * std::cos(al) / c + std::pow(b, 2) * std::pow(std::sin(al), 2) / std::pow(a, 2);
        if (q <= 1.0) {
            P.mu = (1.0 + std::pow(b / a, 2)) / 4.0;
            P.delta = b * c * std::cos(al) / (2.0 * std::pow(a, 2));
            P.zeta = P.mu - 1.0 / 4.0 + (1.0 - b * std::cos(al) / c) / (4.0 * std::pow(std::sin(al), 2));
            P.eta = 1.0 / 2.0 + 2.0 * P.zeta * c * std::cos(al) / b;
            P.phi = 1.0 + P.zeta - 2.0 * P.mu;
            P.psi = P.eta - 2.0 * P.delta;
            return {BZ_MCLC3, BZN(BZ_MCLC3)};
        }
        P.zeta = (std::pow(b / a, 2) + (1.0 - b * std::cos(al) / c) / std::pow(std::sin(al), 2)) / 4.0;
        P.eta = 1.0 / 2.0 + 2.0 * P.zeta * c * std::cos(al) / b;
        P.mu = P.eta / 2.0 + std::pow(b, 2) / (4.0 * std::pow(a, 2)) - b * c * std::cos(al) / (2.0 * std::pow(a, 2));
        P.nu = 2.0 * P.mu - P.zeta;
        P.omega = ((4.0 * P.nu - 1.0 - std::pow(b, 2) * std::pow(std::sin(al), 2) / std::pow(a, 2)) * c) /
                  (2.0 * b * std::cos(al));
        P.delta = P.zeta * c * std::cos(al) / b + P.omega / 2.0 - 1.0 / 4.0;
        P.rho = 1.0 - P.zeta * std::pow(a / b, 2);
        return {BZ_MCLC5, BZN(BZ_MCLC5)};
    }
    if (is(t, "triclinic")) {
        const double b23 = B[3] * B[6] + B[4] * B[7] + B[5] * B[8];
        if (b23 < 0) return {BZ_TRI1, BZN(BZ_TRI1)};
        if (b23 > 0) return {BZ_TRI2, BZN(BZ_TRI2)};
        return {nullptr, 0};
    }
    return {nullptr, 0};
}

}  // namespace

extern "C" int fame_bz_vertex(const fame_lattice_out* lat, char name, double v[3]) {
    const double* B = lat->recip_b;
    auto MV = [&](double V1, double V2, double V3) {
        v[0] = B[0] * V1 + B[3] * V2 + B[6] * V3;
        v[1] = B[1] * V1 + B[4] * V2 + B[7] * V3;
        v[2] = B[2] * V1 + B[5] * V2 + B[8] * V3;
        return 0;
    };
    if (is(lat->lattice_type, "hexagonal")) {
        /* :163-172; the auxiliary point k has an irrational coordinate */
        switch (name) {
            case 'G': return MV(0.0, 0.0, 0.0);
            case 'A': return MV(0.0, 0.0, 0.5);
            case 'H': return MV(1 / 3., 1 / 3., 0.5);
            case 'K': return MV(1 / 3., 1 / 3., 0.0);
            case 'L': return MV(0.5, 0.0, 0.5);
            case 'M': return MV(0.5, 0.0, 0.0);
            case 'k': return MV((0.5 * std::sqrt(2) - 1 / 3.), -1 / 3., 0.0);
        }
        return -1;
    }
    BzParams P;
    const BzSet set = bz_set(lat, P);
    for (int i = 0; i < set.n; ++i)
        if (set.pts[i].name == name)
            return MV(bz_coord(P, set.pts[i].x), bz_coord(P, set.pts[i].y), bz_coord(P, set.pts[i].z));
    return -1;
}

extern "C" void fame_wave_vec_rotate(const fame_lattice_out* lat, const double V[3], double k[3]) {
    const double* Om = lat->Omega;   /* BZ_Path.cpp:80-90 */
    k[0] = Om[0] * V[0] + Om[3] * V[1] + Om[6] * V[2];
    k[1] = Om[1] * V[0] + Om[4] * V[1] + Om[7] * V[2];
    k[2] = Om[2] * V[0] + Om[5] * V[1] + Om[8] * V[2];
}

/* The reference's default path per lattice type and variant
 * (FAME_Parameter_Brillouin_Zone_Path.cpp:166-267), with its own variant tests:
 * the rhombohedral angle from a and c, the face-centred orthorhombic test with
 * strict inequalities, and the base-centred monoclinic test in single
 * precision (cosf, sinf) against the angle between b1 and b2.  nullptr where
 * the reference sets no path. */
extern "C" const char* fame_bz_default_path(const fame_lattice_out* lat) {
    const char* t = lat->lattice_type;
    const double a = lat->a, b = lat->b, c = lat->c, al = lat->alpha;
    if (is(t, "simple_cubic")) return "GXMGRX|MR";
    if (is(t, "face_centered_cubic")) return "GXWKGLUWLK|UX";
    if (is(t, "body_centered_cubic")) return "GHNGPH|PN";
    if (is(t, "hexagonal")) return "GMKGALHA|LM|KH";
    if (is(t, "rhombohedral")) {
        const double th = 2 * std::asin(3 * a / (2 * std::sqrt(std::pow(c, 2) + 3 * std::pow(a, 2))));
        return th <= PI / 2 ? "GLC|BZGX|QFRZ|LP" : "GPZQGFRSLZ";
    }
    if (is(t, "primitive_tetragonal")) return "GXMGZRAZ|XR|MA";
    if (is(t, "body_centered_tetragonal")) {
        if (c < a) return "GXMGZPNBM|XP";
        if (c > a) return "GXYsGZtNPAZ|XP";
        return nullptr;
    }
    if (is(t, "primitive_orthorhombic")) return "GXSYGZURTZ|YT|UX|SR";
    if (is(t, "c_base_centered_orthorhombic")) return "GXSRAZGYCBTY|ZT";
    if (is(t, "face_centered_orthorhombic")) {
        const double q1 = 1 / std::pow(a, 2), q2 = 1 / std::pow(b, 2) + 1 / std::pow(c, 2);
        if (q1 > q2) return "GYTZGXBY|TJ|XAZ|LG";
        if (q1 < q2) return "GYCDXGZFHC|EZ|XI|HY|LG";
        return "GYTZGXBY|XAZ|LG";
    }
    if (is(t, "body_centered_orthorhombic")) return "GXLTWRAZGYSW|IY|BZ";
    if (is(t, "primitive_monoclinic")) return "GYHCENAXI|MDZ|YD";
    if (is(t, "base_centered_monoclinic")) {
        const double q = b * cosf(al) / c + std::pow(b, 2) * std::pow(sinf(al), 2) / std::pow(a, 2);
        const double* B = lat->recip_b;
        double dot = 0, n1 = 0, n2 = 0;
        for (int i = 0; i < 3; i++) { dot += B[i] * B[3 + i]; n1 += std::pow(B[i], 2); n2 += std::pow(B[i + 3], 2); }
        n1 = std::sqrt(n1); n2 = std::sqrt(n2);
        const double kg = std::acos(dot / (n1 * n2));
        if (kg > PI / 2) return "GYFLI|RZJ|YT|XGN|MG";
        if (kg == PI / 2) return "GYFLI|RZJ|NGM";
        if (q < 1) return "GYFHZIJ|PVXGN|MG";
        if (q == 1) return "GYFHZI|PVXGN|MG";
        return "GYFLI|RZHJ|PVXGN|MG";
    }
    if (is(t, "triclinic")) return "XGY|LGZ|NGM|RG";
    return nullptr;
}

/* confirm_wave_vec_array_size, BZ_Path.cpp:270-289 */
extern "C" int fame_bz_path_count(const char* p, int part_num) {
    int count = 0, total = 0, n = (int)std::strlen(p);
    for (int i = 0; i < n; i++) {
        if (p[i] == '|') { total += (i - count - 1) * (part_num - 1) + 1; count = i + 1; }
        else if (i == n - 1) total += (i - count) * (part_num - 1) + 1;
    }
    return total;
}

/* FAME_Parameter_Brillouin_Zone_Path.cpp:9-93 */
extern "C" int fame_bz_path(const fame_lattice_out* lat, const char* p, int part_num, double* W) {
    int n = (int)std::strlen(p), count = 0;
    for (int i = 0; i < n - 1; i++) {
        if (p[i] != '|' && p[i + 1] != '|') {
            double s[3], e[3];
            if (fame_bz_vertex(lat, p[i], s) != 0 || fame_bz_vertex(lat, p[i + 1], e) != 0) return -1;
            double px = (e[0] - s[0]) / (double)(part_num - 1);
            double py = (e[1] - s[1]) / (double)(part_num - 1);
            double pz = (e[2] - s[2]) / (double)(part_num - 1);
            int last = (i == n - 2 || p[i + 2] == '|') ? part_num : part_num - 1;
            for (int j = 0; j < last; j++) {
                W[count * 3 + 0] = s[0] + px * j;
                W[count * 3 + 1] = s[1] + py * j;
                W[count * 3 + 2] = s[2] + pz * j;
                count++;
            }
        }
    }
    for (int i = 0; i < count; i++) {
        double V[3] = {W[3 * i], W[3 * i + 1], W[3 * i + 2]};
        fame_wave_vec_rotate(lat, V, &W[3 * i]);
    }
    return count;
}
