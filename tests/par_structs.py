"""ctypes mirrors of the reference's call-surface structs (PAR, PROFILE, ...),
laid out as include/FAME_Internal_Common.h:29-211 of the reference tree lays
them out; tests/test_layout.py checks every offset against the table compiled
from that header (tests/golden/par_layout.json)."""
import ctypes as C

dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)


class MESH(C.Structure):
    _fields_ = [("grid_nums", C.c_int * 3), ("edge_len", C.c_double * 3), ("mesh_len", C.c_double * 3)]


class LATTICE_CONSTANT(C.Structure):
    _fields_ = [("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("alpha", C.c_double),
                ("beta", C.c_double), ("gamma", C.c_double), ("theta_1", C.c_double), ("theta_2", C.c_double),
                ("theta_3", C.c_double), ("m1", C.c_int), ("m2", C.c_int), ("m3", C.c_int), ("m4", C.c_int),
                ("t1", dp), ("t2", dp), ("t3", dp), ("t4", dp), ("flag", C.c_char * 6),
                ("length_a1", C.c_double), ("length_a2", C.c_double), ("length_a3", C.c_double)]


class MATERIAL(C.Structure):
    _fields_ = [("data_name", C.c_char * 50), ("material_type", C.c_char * 50), ("Binout", ip),
                ("material_num", C.c_int), ("sphere_num", ip), ("sphere_centers", dp), ("sphere_radius", dp),
                ("cylinder_num", ip), ("cylinder_top_centers", dp), ("cylinder_bot_centers", dp),
                ("cylinder_radius", dp), ("num_ele_permitt_in", C.c_int), ("ele_permitt_in", dp),
                ("ele_permitt_out", C.c_double), ("flag_radius_adjustment", C.c_int), ("flag_radius", C.c_double)]


class VERTEX(C.Structure):
    _fields_ = [(n, C.c_double * 3) for n in list("ABCDEFGHIJKLMNOPQRSTUVWXYZ") + ["s", "t", "k"]]


class RECIP_LATTICE(C.Structure):
    _fields_ = [("part_num", C.c_int), ("Wave_vec_num", C.c_int), ("WaveVector", dp), ("path_string", C.c_char * 50),
                ("reciprocal_lattice_vector_b", C.c_double * 9), ("vertex", VERTEX)]


class LATTICE(C.Structure):
    _fields_ = [("lattice_type", C.c_char * 50), ("Permutation", C.c_int * 3), ("lattice_vec_a", C.c_double * 9),
                ("lattice_vec_a_orig", C.c_double * 9), ("Omega", C.c_double * 9),
                ("lattice_constant", LATTICE_CONSTANT)]


class LS(C.Structure):
    _fields_ = [("maxit", C.c_int), ("tol", C.c_double)]


class ES(C.Structure):
    _fields_ = [("nwant", C.c_int), ("nstep", C.c_int), ("maxit", C.c_int), ("tol", C.c_double)]


class FLAG(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("device", "printf_user_option", "printf_parameter", "create_parameter",
                                       "create_B_inout", "create_wave_vector", "save_eigen_vector",
                                       "grid_nums_max", "mem_size", "radius_adjustment")] + [("radius", C.c_double)]


class PAR(C.Structure):
    _fields_ = [("flag", FLAG), ("mesh", MESH), ("lattice", LATTICE), ("material", MATERIAL),
                ("recip_lattice", RECIP_LATTICE), ("ls", LS), ("es", ES)]


class PROFILE(C.Structure):
    _fields_ = [("idx", C.c_int), ("ls_iter", ip), ("es_iter", ip), ("ls_time", dp), ("es_time", dp)]
