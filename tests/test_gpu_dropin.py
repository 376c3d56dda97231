"""The reference's call surface on the GPU: FAME_Main_Code(PAR, PROFILE*) with
the reference's input structs (ctypes mirrors of the reference's layout,
tests/par_structs.py), its output files (Data_Omega_array.txt, Data_EV_kk_bb.txt,
Data_Profile.txt through FAME_Create_Profile_txt), and Lanczos_Isotropic on a
context.  Frequencies are checked against the C-ABI solve of the same inputs."""
import ctypes as C
import os

import numpy as np
import pytest

import fame_amd as F

pytestmark = pytest.mark.gpu

from par_structs import ES, FLAG, LATTICE, LATTICE_CONSTANT, LS, MATERIAL, MESH, PAR, PROFILE, RECIP_LATTICE, VERTEX, dp, ip  # noqa: F401


def _parse_omega(path, nk, nwant):
    vals = []
    for line in open(path):
        line = line.strip()
        if line and not line.startswith("Freq_array") and line != "];":
            vals.append(float(line))
    return np.array(vals).reshape(nk, nwant)


def test_fame_main_code_and_outputs(gpu, tmp_path, monkeypatch):
    lo = F.build_lattice("simple_cubic", 8)
    fl = F.to_fame_lattice(lo)
    mat = F.Material(sphere_centers=[[(0.5, 0.5, 0.5)]], sphere_radius=[[0.25]])
    idx = F.material_index(lo, mat)
    N = idx.size // 3
    binout = np.ascontiguousarray((idx == 0).astype(np.int32))
    ks = np.ascontiguousarray([[0.21, -0.13, 0.07], [0.0, 0.0, 0.0], [0.25, 0.125, 0.0]], np.float64)
    nk, nwant, nstep = 3, 6, 20
    t = [np.ascontiguousarray(getattr(fl, f"t{i}"), np.float64) for i in range(1, 5)]
    eps_in = np.array([13.0])

    P = PAR()
    P.flag.device = 0
    P.flag.save_eigen_vector = 1
    for i in range(3):
        P.mesh.grid_nums[i] = fl.grid_nums[i]
        P.mesh.mesh_len[i] = fl.mesh_len[i]
    P.lattice.lattice_type = lo.lattice_type
    for i in range(9):
        P.lattice.lattice_vec_a[i] = fl.lattice_vec_a[i]
    lc = P.lattice.lattice_constant
    lc.m1, lc.m2, lc.m3, lc.m4 = fl.m1, fl.m2, fl.m3, fl.m4
    lc.t1, lc.t2, lc.t3, lc.t4 = [a.ctypes.data_as(dp) for a in t]
    P.material.material_type = b"isotropic"
    P.material.Binout = binout.ctypes.data_as(ip)
    P.material.material_num = 1
    P.material.num_ele_permitt_in = 1
    P.material.ele_permitt_in = eps_in.ctypes.data_as(dp)
    P.material.ele_permitt_out = 1.0
    P.recip_lattice.Wave_vec_num = nk
    P.recip_lattice.WaveVector = ks.ctypes.data_as(dp)
    P.ls.maxit, P.ls.tol = 1000, 1e-12
    P.es.nwant, P.es.nstep, P.es.maxit, P.es.tol = nwant, nstep, 300, 1e-12
    li, ei = np.zeros(nk, np.int32), np.zeros(nk, np.int32)
    lt, et = np.zeros(nk), np.zeros(nk)
    prof = PROFILE(0, li.ctypes.data_as(ip), ei.ctypes.data_as(ip), lt.ctypes.data_as(dp), et.ctypes.data_as(dp))

    lib = C.CDLL(F.LIB_PATH)
    lib.FAME_Main_Code.argtypes = [PAR, C.POINTER(PROFILE)]
    lib.FAME_Create_Profile_txt.argtypes = [PROFILE]
    monkeypatch.chdir(tmp_path)
    assert lib.FAME_Main_Code(P, C.byref(prof)) == 0
    assert lib.FAME_Create_Profile_txt(prof) == 0

    om = _parse_omega(tmp_path / "Data_Omega_array.txt", nk, nwant)
    S = F.Solver(fl, max_batch=nk)
    S.set_eps(F.material_eps(lo, mat, [13.0], 1.0))
    S.set_wave_vectors(ks)
    f = S.eigs(nwant, nstep)[0]
    S.close()
    # a negative omega flags a failed absolute residual gate: compare magnitudes
    m = f > 0
    assert np.max(np.abs(np.abs(om[m]) ** 2 - f[m] ** 2) / f[m] ** 2) < 1e-8
    assert np.all(om[~m] == 0)

    lines = open(tmp_path / "Data_EV_01_01.txt").read().splitlines()
    assert lines[0] == "Ele_field_mtx{1}(:, 1) = [" and lines[-1] == "];" and len(lines) == 3 * N + 2
    assert os.path.exists(tmp_path / f"Data_EV_{nk:02d}_{nwant:02d}.txt")
    prof_lines = open(tmp_path / "Data_Profile.txt").read().splitlines()
    assert prof_lines[0] == "Compute_Info = [" and len(prof_lines) == nk + 3
    assert prof.idx == nk - 1 and np.all(li > 0) and np.all(ei > 0)


def test_lanczos_isotropic_on_a_context(gpu):
    lo = F.build_lattice("face_centered_cubic", 10)
    fl = F.to_fame_lattice(lo)
    eps = F.material_eps(lo, F.Material(sphere_centers=[[(0, 0, 0)]], sphere_radius=[[0.2]]), [13.0], 1.0)
    ks = np.array([[0.21, -0.13, 0.07], [0.1, 0.2, 0.0]])
    S = F.Solver(fl, max_batch=2)
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    ref = S.eigs(6, 20)[0]
    lib = C.CDLL(F.LIB_PATH)
    lib.Lanczos_Isotropic.argtypes = [dp, C.c_void_p, C.c_void_p, ES, LS, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                      C.c_void_p]
    freq = np.zeros(2 * 6)
    nx, ny, nz = fl.grid_nums
    r = lib.Lanczos_Isotropic(freq.ctypes.data_as(dp), None, S._ctx, ES(6, 20, 300, 1e-12), LS(1000, 1e-12),
                              nx, ny, nz, b"General", None)
    assert r >= 1
    assert np.allclose(freq.reshape(2, 6), ref, rtol=1e-10, atol=0)
    # a grid that does not match the context is an argument error, not a crash
    assert lib.Lanczos_Isotropic(freq.ctypes.data_as(dp), None, S._ctx, ES(6, 20, 300, 1e-12), LS(1000, 1e-12),
                                 nx + 1, ny, nz, b"General", None) < 0
    S.close()
