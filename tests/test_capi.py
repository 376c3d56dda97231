"""The C-ABI library loads on a machine without a GPU and exports every entry
point its headers declare; host-only entry points (input generation) run here."""
import ctypes as C
import os
import re

import numpy as np

import fame_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fame_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    lib = F.lib()
    missing = [s for h in ("fame_hip.h", "fame_geometry.h", "fame_main.h", "fame_driver.h") for s in _declared(h) if not hasattr(lib, s)]
    assert not missing, missing


def test_drop_in_surface_exported():
    lib = C.CDLL(F.LIB_PATH)
    for s in ("FAME_Main_Code", "Lanczos_Isotropic", "FAME_Create_Profile_txt"):
        assert hasattr(lib, s), s


def test_profile_txt_format(tmp_path, monkeypatch):
    """Data_Profile.txt in the reference's format (FAME_Tools/FAME_Profile.cpp:27-39);
    host-only, no device needed."""
    class PROFILE(C.Structure):
        _fields_ = [("idx", C.c_int), ("ls_iter", C.POINTER(C.c_int)), ("es_iter", C.POINTER(C.c_int)),
                    ("ls_time", C.POINTER(C.c_double)), ("es_time", C.POINTER(C.c_double))]
    li = np.array([120, 7], np.int32); ei = np.array([33, 4], np.int32)
    lt = np.array([1.5, 0.25]); et = np.array([2.0, 0.125])
    pr = PROFILE(1, li.ctypes.data_as(C.POINTER(C.c_int)), ei.ctypes.data_as(C.POINTER(C.c_int)),
                 lt.ctypes.data_as(C.POINTER(C.c_double)), et.ctypes.data_as(C.POINTER(C.c_double)))
    lib = C.CDLL(F.LIB_PATH)
    lib.FAME_Create_Profile_txt.argtypes = [PROFILE]
    monkeypatch.chdir(tmp_path)
    assert lib.FAME_Create_Profile_txt(pr) == 0
    assert open(tmp_path / "Data_Profile.txt").read() == (
        "Compute_Info = [\n%idx   LS_iter     LS_time   ES_iter     ES_time\n"
        "   1,      120,       1.50,       33,       2.00;\n"
        "   2,        7,       0.25,        4,       0.12;\n];\n")


def test_no_device_needed_for_errors():
    # argument errors are reported as status codes, never aborts
    ctx = C.c_void_p()
    assert F.lib().fame_create(None, 1, 0, C.byref(ctx)) == F.FAME_ERR_ARG
    assert b"null" in F.lib().fame_last_error()


def test_grid_rule_headline():
    lo = F.build_lattice("face_centered_cubic", 96)
    assert tuple(lo.grid_nums) == (96, 83, 78)
    assert lo.general == 1
    lo = F.build_lattice("body_centered_cubic", 128)
    assert tuple(lo.grid_nums) == (128, 121, 105)
    lo = F.build_lattice("simple_cubic", 32)
    assert tuple(lo.grid_nums) == (32, 32, 32) and lo.general == 0


def test_path_by_arclength_ends_on_vertices():
    lo = F.build_lattice("face_centered_cubic", 16)
    ks = F.path_by_arclength(lo, "GXWLG", 30)
    assert ks.shape == (30, 3)
    assert np.abs(ks[0]).max() == 0 and np.abs(ks[-1]).max() < 1e-15
    X = F.rotate_wave_vector(lo, F.bz_vertex(lo, "X"))
    assert np.min(np.linalg.norm(ks - X, axis=1)) < 1e-12
