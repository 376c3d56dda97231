"""Shared small structures for the parity tests (inputs only)."""
import numpy as np

import fame_amd as F
import oracle as O

SPHERE = F.Material(sphere_centers=[[(0.5, 0.5, 0.5)]], sphere_radius=[[0.25]])
DIAMOND = F.Material(sphere_centers=[[(0, 0, 0), (0.25, 0.25, 0.25)]], sphere_radius=[[0.15, 0.15]],
                     cylinder_top=[[(0.25, 0.25, 0.25)] * 4],
                     cylinder_bot=[[(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]],
                     cylinder_radius=[[0.10] * 4])

# (name, lattice type, Nx, lattice constants, material, eps_in)
STRUCTS = {
    "sc": ("simple_cubic", 8, {}, SPHERE, 13.0),
    "fcc": ("face_centered_cubic", 10, {}, DIAMOND, 13.0),
    "bcc": ("body_centered_cubic", 9, {}, SPHERE, 12.96),
    "hex": ("hexagonal", 8, {"c": 1.2}, SPHERE, 13.0),
}
KPTS = {"k": np.array([0.21, -0.13, 0.07]), "gamma": np.zeros(3)}


def make(name):
    lt, nx, kw, mat, ein = STRUCTS[name]
    lo = F.build_lattice(lt, nx, **kw)
    fl = F.to_fame_lattice(lo)
    eps = F.material_eps(lo, mat, [ein], 1.0)
    return lo, fl, O.lattice_from(fl), eps


def equal_mesh_general(shifts=(0, 0, 0, 0), n=8, seed=5):
    """A General-lattice grid whose three mesh lengths are equal (n^3 cells of
    h = 1/n, the given boundary shifts m1..m4, a monoclinic cell's boundary
    translations) with a seeded permittivity in [1, 13] per Yee edge.  On such
    grids Lambda is parallel to (1, 1, 1) at some nonzero modes (at Gamma with
    these shifts), where the reference's General Pi normalisers are 0/0."""
    lo = F.build_lattice("primitive_monoclinic", n, a=1.0, b=1.0, c=1 / np.sin(1.4), alpha=1.4)
    fl = F.to_fame_lattice(lo)
    assert fl.general == 1 and list(fl.grid_nums) == [n, n, n]
    for i in range(3):
        fl.mesh_len[i] = 1.0 / n
    fl.m1, fl.m2, fl.m3, fl.m4 = shifts
    eps = np.random.default_rng(seed).uniform(1.0, 13.0, 3 * n ** 3)
    return fl, O.lattice_from(fl), eps
