"""The reference's program flow on the GPU: a directory holding Popt.txt and
FAME_Material_Data/<name>.txt, run through fame_run (Main.cpp:8-49 restated)
and through the fame_main program, against the C-ABI solve of the same
inputs (lattice, grid, mask, default Brillouin-zone path, permittivities)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import fame_amd as F

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

POPT = """# Grid number
 {nx} {nx} {nx}
# Part number
 {part}
# Material data name
 Test_Structure
# Material type
 isotropic
# Permittivity (inner material)
 1
 13.0
# Permittivity (outer material)
 1.0
# Desired eigenpair number
 {nwant}
# Dimension of Krylov subspace for Lanczos
 {nstep}
# Tolerance of Lanczos
 1e-12
# Maximum restart number of Lanczos
 300
# Tolerance of linear solver
 1e-12
# Maximum iteration number of linear solver
 1000
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
Test_Structure
# Lattice type
{lattice}
# Lattice constant
1 1 1 1.5707963267948966 1.5707963267948966 1.5707963267948966
# Material number
1
# Sphere number
1
# Sphere centers
0.5 0.5 0.5
# Sphere radius
0.25
# Cylinder number
0
"""


def _write(d, lattice, nx, part, nwant, nstep):
    (d / "FAME_Material_Data").mkdir()
    (d / "FAME_Material_Data" / "Test_Structure.txt").write_text(MATERIAL.format(lattice=lattice))
    (d / "Popt.txt").write_text(POPT.format(nx=nx, part=part, nwant=nwant, nstep=nstep))


def _omega(path):
    vals = [float(ln) for ln in open(path).read().split("\n")
            if ln.strip() and not ln.startswith("Freq_array") and ln.strip() != "];"]
    return np.array(vals)


def _reference_solve(lattice, nx, part, nwant, nstep):
    lo = F.build_lattice(lattice, nx)
    fl = F.to_fame_lattice(lo)
    eps = F.material_eps(lo, F.Material(sphere_centers=[[(0.5, 0.5, 0.5)]], sphere_radius=[[0.25]]), [13.0], 1.0)
    ks = F.bz_path(lo, None, part)
    S = F.Solver(fl, max_batch=len(ks))
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    f = S.eigs(nwant, nstep)[0]
    S.close()
    return f


@pytest.mark.parametrize("lattice,nx", [("simple_cubic", 16), ("face_centered_cubic", 16)])
def test_fame_run_matches_c_abi(gpu, tmp_path, lattice, nx):
    part, nwant, nstep = 3, 6, 20
    _write(tmp_path, lattice, nx, part, nwant, nstep)
    L = F.lib()
    L.fame_run.argtypes = [C.c_char_p]
    cwd = os.getcwd()
    assert L.fame_run(str(tmp_path).encode()) == 0
    assert os.getcwd() == cwd
    got = _omega(tmp_path / "Data_Omega_array.txt")
    ref = _reference_solve(lattice, nx, part, nwant, nstep)
    assert got.shape == (ref.size,)
    assert np.max(np.abs(np.abs(got) - ref.reshape(-1)) / np.maximum(ref.reshape(-1), 1e-300)) < 1e-9
    assert (tmp_path / "Data_Profile.txt").read_text().startswith("Compute_Info = [")


def test_fame_main_program(gpu, tmp_path):
    exe = os.path.join(ROOT, "fame-isotropic-gpu_amd", "fame_main")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (make -C fame-isotropic-gpu_amd/csrc)")
    _write(tmp_path, "simple_cubic", 16, 3, 6, 20)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    got = _omega(tmp_path / "Data_Omega_array.txt")
    ref = _reference_solve("simple_cubic", 16, 3, 6, 20)
    assert np.max(np.abs(np.abs(got) - ref.reshape(-1)) / np.maximum(ref.reshape(-1), 1e-300)) < 1e-9
    assert "Grid_nums = [  16  16  16 ]" in r.stdout
