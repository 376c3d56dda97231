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
