"""The reference's input files and top-level program (include/fame_driver.h), on
the CPU: Popt.txt and material-file readers against the reference's own files
and against files written here, the grid auto-sizing of Popt's "Grid nums Max"
against the reference's compiled preprocessing, and the Brillouin-zone vertex
tables of all 14 lattice types."""
import ctypes as C
import os

import numpy as np
import pytest

import fame_amd as F

REF_TREE = os.environ.get("FAME_REFERENCE_DIR", "/root/reference")


class fame_popt(C.Structure):
    _fields_ = [("grid_nums", C.c_int * 3), ("part_num", C.c_int), ("data_name", C.c_char * 50),
                ("material_type", C.c_char * 50), ("num_ele_permitt_in", C.c_int),
                ("ele_permitt_in", C.c_double * 64), ("ele_permitt_out", C.c_double),
                ("nwant", C.c_int), ("nstep", C.c_int), ("es_tol", C.c_double), ("es_maxit", C.c_int),
                ("ls_tol", C.c_double), ("ls_maxit", C.c_int)] + \
               [(n, C.c_int) for n in ("device", "printf_user_option", "printf_parameter", "create_parameter",
                                       "create_B_inout", "create_wave_vector", "save_eigen_vector",
                                       "grid_nums_max", "mem_size", "radius_adjustment")] + [("radius", C.c_double)]


dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)


class fame_material_file(C.Structure):
    _fields_ = [("data_name", C.c_char * 50), ("lattice_type", C.c_char * 50)] + \
               [(n, C.c_double) for n in ("a", "b", "c", "alpha", "beta", "gamma")] + \
               [("material_num", C.c_int), ("sphere_num", ip), ("sphere_centers", dp), ("sphere_radius", dp),
                ("cylinder_num", ip), ("cylinder_top_centers", dp), ("cylinder_bot_centers", dp),
                ("cylinder_radius", dp)]


def _lib():
    L = F.lib()
    L.fame_read_popt.argtypes = [C.c_char_p, C.POINTER(fame_popt)]
    L.fame_read_material.argtypes = [C.c_char_p, C.c_int, C.c_double, C.POINTER(fame_material_file)]
    L.fame_material_file_free.argtypes = [C.POINTER(fame_material_file)]
    L.fame_grid_nums_max.argtypes = [C.POINTER(F.fame_lattice_in), C.c_int, C.c_int, C.c_int]
    return L


POPT = """======================== Mesh information ========================
# Grid number
 24 24 24

================= Reciprocal lattice information =================
# Part number
 3

====================== Material information ======================
# Material data name
 Test_SC_Sphere
# Material type
 isotropic
# Permittivity (inner material)
 2
 13.0 11.5
# Permittivity (outer material)
 1.0

======================= Solver information =======================
# Desired eigenpair number
 6
# Dimension of Krylov subspace for Lanczos
 20
# Tolerance of Lanczos
 1e-12
# Maximum restart number of Lanczos
 300
# Tolerance of linear solver
 1e-12
# Maximum iteration number of linear solver
 1000

 ========================== Flag setting ==========================
# Device
 0
# Save Eigen Vector
 0
# Grid nums Max
 0 16
# Sphere/Cylinder radius adjustment
 0 0.6
"""

MATERIAL = """# Data name
Test_SC_Sphere
# Lattice type
simple_cubic
# Lattice constant
1 1 1 1.5707963267948966 1.5707963267948966 1.5707963267948966
# Material number
2
# Sphere number
1 0
# Sphere centers
0.5 0.5 0.5
# Sphere radius
0.25
# Cylinder number
0 1
# Cylinder top centers
0 0 1
# Cylinder bot centers
0 0 0
# Cylinder radius
0.1
"""


def test_read_popt_written_here(tmp_path):
    p = tmp_path / "Popt.txt"
    p.write_text(POPT)
    o = fame_popt()
    assert _lib().fame_read_popt(str(p).encode(), C.byref(o)) == 0
    assert list(o.grid_nums) == [24, 24, 24] and o.part_num == 3
    assert o.data_name == b"Test_SC_Sphere" and o.material_type == b"isotropic"
    assert o.num_ele_permitt_in == 2 and list(o.ele_permitt_in[:2]) == [13.0, 11.5] and o.ele_permitt_out == 1.0
    assert (o.nwant, o.nstep, o.es_maxit, o.ls_maxit) == (6, 20, 300, 1000)
    assert o.es_tol == 1e-12 and o.ls_tol == 1e-12
    assert (o.grid_nums_max, o.mem_size, o.radius_adjustment, o.radius) == (0, 16, 0, 0.6)
    assert _lib().fame_read_popt(str(tmp_path / "missing.txt").encode(), C.byref(o)) == -1
    p.write_text("# Grid number\n 24 x 24\n")
    assert _lib().fame_read_popt(str(p).encode(), C.byref(o)) == -2


@pytest.mark.skipif(not os.path.isdir(REF_TREE), reason="reference tree not present")
def test_read_reference_popt():
    o = fame_popt()
    assert _lib().fame_read_popt(os.path.join(REF_TREE, "Popt.txt").encode(), C.byref(o)) == 0
    assert list(o.grid_nums) == [48, 48, 48] and o.part_num == 10 and o.data_name == b"No221_SC_Demo"
    assert o.num_ele_permitt_in == 6 and list(o.ele_permitt_in[:6]) == [13.0] * 6
    assert (o.nwant, o.nstep, o.es_maxit, o.ls_maxit, o.device) == (10, 30, 300, 1000, 0)
    assert (o.grid_nums_max, o.mem_size, o.radius_adjustment, o.radius) == (0, 16, 0, 0.6)


def _read_material(path, adj=0, radius=0.0):
    m = fame_material_file()
    rc = _lib().fame_read_material(str(path).encode(), adj, radius, C.byref(m))
    return rc, m


def test_read_material_written_here(tmp_path):
    p = tmp_path / "m.txt"
    p.write_text(MATERIAL)
    rc, m = _read_material(p)
    assert rc == 0
    assert m.lattice_type == b"simple_cubic" and m.material_num == 2 and m.a == 1.0
    assert [m.sphere_num[i] for i in range(2)] == [1, 0] and [m.cylinder_num[i] for i in range(2)] == [0, 1]
    assert [m.sphere_centers[i] for i in range(3)] == [0.5, 0.5, 0.5] and m.sphere_radius[0] == 0.25
    assert [m.cylinder_top_centers[i] for i in range(3)] == [0, 0, 1] and m.cylinder_radius[0] == 0.1
    _lib().fame_material_file_free(C.byref(m))
    rc, m = _read_material(p, 1, 0.6)                   # radius adjustment overrides every radius
    assert rc == 0 and m.sphere_radius[0] == 0.6 and m.cylinder_radius[0] == 0.6
    _lib().fame_material_file_free(C.byref(m))
    assert _read_material(tmp_path / "none.txt")[0] == -1


MATS = sorted(f[:-4] for f in os.listdir(os.path.join(REF_TREE, "FAME_Material_Data"))) \
    if os.path.isdir(REF_TREE) else []


@pytest.mark.parametrize("name", MATS)
def test_read_every_reference_material(name):
    """Each of the reference's material files against an independent Python
    reading of the same sections."""
    path = os.path.join(REF_TREE, "FAME_Material_Data", name + ".txt")
    sec, cur = {}, None
    for ln in open(path).read().split("\n"):
        if ln.startswith("#"):
            cur = ln.strip(); sec[cur] = []
        elif cur and ln.strip():
            sec[cur] += ln.split()
    rc, m = _read_material(path)
    assert rc == 0
    assert m.lattice_type.decode() == sec["# Lattice type"][0]
    assert [m.a, m.b, m.c, m.alpha, m.beta, m.gamma] == [float(x) for x in sec["# Lattice constant"]]
    nm = int(sec["# Material number"][0])
    assert m.material_num == nm
    sn = [int(x) for x in sec.get("# Sphere number", ["0"] * nm)]
    cn = [int(x) for x in sec.get("# Cylinder number", ["0"] * nm)]
    assert [m.sphere_num[i] for i in range(nm)] == sn and [m.cylinder_num[i] for i in range(nm)] == cn
    for key, arr, n in (("# Sphere centers", m.sphere_centers, 3 * sum(sn)), ("# Sphere radius", m.sphere_radius, sum(sn)),
                        ("# Cylinder top centers", m.cylinder_top_centers, 3 * sum(cn)),
                        ("# Cylinder bot centers", m.cylinder_bot_centers, 3 * sum(cn)),
                        ("# Cylinder radius", m.cylinder_radius, sum(cn))):
        assert [arr[i] for i in range(n)] == [float(x) for x in sec.get(key, [])[:n]], (name, key)
    _lib().fame_material_file_free(C.byref(m))
