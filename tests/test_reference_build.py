"""The restated input generation and sparse curl against the reference's own
host code, compiled from the reference sources into oracle/_ref/ by
oracle/build_ref.sh.  Skipped where that build is absent (e.g. on a machine
without the reference tree).  Bit-exact bar: lattice vectors, grid, m1..m4,
t1..t4, rotation, reciprocal vectors, wave vectors and the Yee-edge material
mask must be identical; curl values agree to a few ulps (libm cexp vs std::exp).
"""
import ctypes as C
import os
import shutil
import tempfile

import numpy as np
import pytest

import fame_amd as F
import oracle as O

REF_TREE = os.environ.get("FAME_REFERENCE_DIR", "/root/reference")
pytestmark = pytest.mark.skipif(not (os.path.exists(O.REF_LIB_PATH) and os.path.isdir(REF_TREE)),
                                reason="reference host build (oracle/_ref) or reference tree not present")


class RefOut(C.Structure):
    _fields_ = [("grid_nums", C.c_int * 3), ("edge_len", C.c_double * 3), ("mesh_len", C.c_double * 3),
                ("lattice_vec_a", C.c_double * 9), ("lattice_vec_a_orig", C.c_double * 9), ("Omega", C.c_double * 9),
                ("Permutation", C.c_int * 3), ("length_a", C.c_double * 3), ("theta", C.c_double * 3),
                ("m1", C.c_int), ("m2", C.c_int), ("m3", C.c_int), ("m4", C.c_int),
                ("t1", C.c_double * 3), ("t2", C.c_double * 3), ("t3", C.c_double * 3), ("t4", C.c_double * 3),
                ("flag", C.c_char * 6), ("recip_b", C.c_double * 9), ("material_num", C.c_int),
                ("wave_vec_num", C.c_int), ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double)]


def _ref():
    return C.CDLL(O.REF_LIB_PATH)


def _parse(path):
    data, sec = {}, None
    for ln in open(path).read().split("\n"):
        if ln.startswith("#"):
            sec = ln.strip()
            data[sec] = []
        elif sec and ln.strip():
            data[sec] += ln.split()
    return data


def _run_ref(name, nx, part_num=4, grid_max=None):
    """The reference's preprocessing of material `name` at Nx = nx; grid_max =
    (mem_size GB, nwant, nstep) turns on its grid auto-sizing."""
    w = tempfile.mkdtemp()
    try:
        os.makedirs(os.path.join(w, "FAME_Material_Data"))
        shutil.copy(os.path.join(REF_TREE, "FAME_Material_Data", name + ".txt"), os.path.join(w, "FAME_Material_Data"))
        popt = open(os.path.join(REF_TREE, "Popt.txt")).read()
        popt = popt.replace(" 48 48 48", f" {nx} {nx} {nx}").replace("No221_SC_Demo", name)
        popt = popt.replace("# Part number\n 10", f"# Part number\n {part_num}")
        if grid_max is not None:
            mem, nwant, nstep = grid_max
            popt = popt.replace("# Grid nums Max\n 0 16", f"# Grid nums Max\n 1 {mem}")
            popt = popt.replace("# Desired eigenpair number\n 10", f"# Desired eigenpair number\n {nwant}")
            popt = popt.replace("# Dimension of Krylov subspace for Lanczos\n 30",
                                f"# Dimension of Krylov subspace for Lanczos\n {nstep}")
        open(os.path.join(w, "Popt.txt"), "w").write(popt)
        ro = RefOut()
        lib = _ref()
        assert lib.ref_preprocess(w.encode(), C.byref(ro)) == 0
        return lib, ro
    finally:
        shutil.rmtree(w)


def _material(d):
    nm = int(d["# Material number"][0])
    sn = [int(x) for x in d.get("# Sphere number", ["0"] * nm)]
    cn = [int(x) for x in d.get("# Cylinder number", ["0"] * nm)]
    sc = np.array([float(x) for x in d.get("# Sphere centers", [])]).reshape(-1, 3)
    sr = [float(x) for x in d.get("# Sphere radius", [])]
    ct = np.array([float(x) for x in d.get("# Cylinder top centers", [])]).reshape(-1, 3)
    cb = np.array([float(x) for x in d.get("# Cylinder bot centers", [])]).reshape(-1, 3)
    cr = [float(x) for x in d.get("# Cylinder radius", [])]
    m = F.Material()
    i = j = 0
    for q in range(nm):
        m.sphere_centers.append([tuple(v) for v in sc[i:i + sn[q]]]); m.sphere_radius.append(sr[i:i + sn[q]]); i += sn[q]
        m.cylinder_top.append([tuple(v) for v in ct[j:j + cn[q]]]); m.cylinder_bot.append([tuple(v) for v in cb[j:j + cn[q]]])
        m.cylinder_radius.append(cr[j:j + cn[q]]); j += cn[q]
    return m


ALL = sorted(f[:-4] for f in os.listdir(os.path.join(REF_TREE, "FAME_Material_Data"))) if os.path.isdir(REF_TREE) else []


@pytest.mark.parametrize("name", ALL)
def test_lattice_bit_exact(name):
    """Lattice, grid, boundary points and reciprocal vectors, and the default
    Brillouin-zone path of the lattice type (all 14 types and their variants
    occur among the shipped structures), bit for bit."""
    d = _parse(os.path.join(REF_TREE, "FAME_Material_Data", name + ".txt"))
    lc = [float(x) for x in d["# Lattice constant"]]
    for nx in (16, 23):
        lib, ro = _run_ref(name, nx)
        lo = F.build_lattice(d["# Lattice type"][0], nx, *lc)
        for fld in ("grid_nums", "edge_len", "mesh_len", "lattice_vec_a", "lattice_vec_a_orig", "Omega", "Permutation",
                    "length_a", "theta", "t1", "t2", "t3", "t4", "recip_b"):
            assert np.array_equal(np.array(getattr(ro, fld)[:]), np.array(getattr(lo, fld)[:])), (name, nx, fld)
        for fld in ("m1", "m2", "m3", "m4", "flag"):
            assert getattr(ro, fld) == getattr(lo, fld), (name, nx, fld)
        if ro.wave_vec_num == 0:
            # the reference sets no path for this type (a-base-centred orthorhombic)
            assert not F.lib().fame_bz_default_path(C.byref(lo)), name
            continue
        wv = np.zeros((ro.wave_vec_num, 3))
        lib.ref_wave_vectors(wv.ctypes.data_as(C.c_void_p))
        assert np.array_equal(F.bz_path(lo, None, 4), wv), (name, nx)


@pytest.mark.parametrize("name,mem,nwant,nstep", [("No221_SC_Demo", 1, 10, 30), ("No227_FCC_DiamondStructure", 2, 16, 48),
                                                   ("No229_BCC_W", 1, 8, 20), ("No191_Hexagonal_AlB2", 3, 12, 36),
                                                   ("No139_BCT_Si", 1, 10, 30)])
def test_grid_nums_max_bit_exact(name, mem, nwant, nstep):
    """Popt's "Grid nums Max" auto-sizing (FAME_Parameter_Generator.cpp:114-148)."""
    d = _parse(os.path.join(REF_TREE, "FAME_Material_Data", name + ".txt"))
    lc = [float(x) for x in d["# Lattice constant"]]
    for nx in (16, 64):
        _, ro = _run_ref(name, nx, grid_max=(mem, nwant, nstep))
        li = F.fame_lattice_in()
        li.lattice_type = d["# Lattice type"][0].encode()
        li.a, li.b, li.c, li.alpha, li.beta, li.gamma = lc
        li.grid_nums[0] = li.grid_nums[1] = li.grid_nums[2] = nx
        L = F.lib()
        L.fame_grid_nums_max.argtypes = [C.POINTER(F.fame_lattice_in), C.c_int, C.c_int, C.c_int]
        nx_new = L.fame_grid_nums_max(C.byref(li), nwant, nstep, mem)
        lo = F.build_lattice(d["# Lattice type"][0], nx_new, *lc)
        assert list(lo.grid_nums) == list(ro.grid_nums), (name, nx, nx_new, list(ro.grid_nums))


@pytest.mark.parametrize("name", ["No221_SC_Demo", "No227_FCC_DiamondStructure", "No229_BCC_W",
                                  "No191_Hexagonal_AlB2", "No225_FCC_NaCl", "No139_BCT_Si", "No166_rhombohedral_alphaAs"])
def test_mask_wavevectors_curl_bit_exact(name):
    d = _parse(os.path.join(REF_TREE, "FAME_Material_Data", name + ".txt"))
    lc = [float(x) for x in d["# Lattice constant"]]
    nx = 12
    lib, ro = _run_ref(name, nx, part_num=4)
    lo = F.build_lattice(d["# Lattice type"][0], nx, *lc)
    mat = _material(d)
    N = lo.grid_nums[0] * lo.grid_nums[1] * lo.grid_nums[2]
    nm = ro.material_num
    binout = np.zeros(3 * nm * N, np.int32)
    lib.ref_binout(binout.ctypes.data_as(C.c_void_p))
    idx = F.material_index(lo, mat)
    mine = np.zeros(3 * nm * N, np.int32)
    for comp in range(3):
        for i in range(N):
            m = idx[comp * N + i]
            if m >= 0:
                mine[comp * nm * N + m * N + i] = 1
    # the reference marks only the first material containing a point
    assert np.array_equal(mine, binout)
    wv = np.zeros((ro.wave_vec_num, 3))
    lib.ref_wave_vectors(wv.ctypes.data_as(C.c_void_p))
    assert np.array_equal(F.bz_path(lo, None, 4), wv)
    fl = F.to_fame_lattice(lo)
    ol = O.lattice_from(fl)
    k = np.array([0.13, -0.21, 0.05])
    r1, c1, v1 = O.curl_coo(ol, k)
    r2 = np.zeros(12 * N, np.int32); c2 = np.zeros(12 * N, np.int32); v2 = np.zeros(12 * N, np.complex128)
    lib.ref_curl(k.ctypes.data_as(C.c_void_p), r2.ctypes.data_as(C.c_void_p), c2.ctypes.data_as(C.c_void_p),
                 v2.ctypes.data_as(C.c_void_p))
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
    assert np.abs(v1 - v2).max() <= 1e-14 * np.abs(v2).max()
