"""The headline workload's General path at full size against the oracle:
FCC diamond network at 64 -> 64x55x52 (12 bands, nstep 36) and 96 -> 96x83x78
(16 bands, nstep 48), two wave vectors of the 30-point Gamma-X-W-L-Gamma path
each.  The expected frequencies were computed by the FP64 CPU oracle and are
committed in tests/golden/oracle_eigs_fcc.json (tests/golden/make_oracle_eigs.py).
The GPU's FP64 and FP32+FP64-refinement solves are held to 1e-8 relative in
lambda = omega^2, the eigenvalue tolerance of the benchmark workload; the
permittivity is checked to be
the one the fixture was made from."""
import hashlib
import json
import os

import numpy as np
import pytest

import fame_amd as F

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DIAMOND = F.Material(sphere_centers=[[(0, 0, 0), (0.25, 0.25, 0.25)]], sphere_radius=[[0.15, 0.15]],
                     cylinder_top=[[(0.25, 0.25, 0.25)] * 4],
                     cylinder_bot=[[(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]], cylinder_radius=[[0.10] * 4])


def _cases():
    with open(os.path.join(HERE, "golden", "oracle_eigs_fcc.json")) as f:
        cs = json.load(f)["cases"]
    by_nx = {}
    for c in cs:
        by_nx.setdefault(c["nx"], []).append(c)
    return sorted(by_nx.items())


@pytest.mark.parametrize("nx,cases", _cases(), ids=lambda v: str(v) if isinstance(v, int) else "")
def test_fcc_path_vs_oracle_fixture(gpu, nx, cases):
    lo = F.build_lattice("face_centered_cubic", nx)
    fl = F.to_fame_lattice(lo)
    assert list(fl.grid_nums) == cases[0]["grid"]
    eps = F.material_eps(lo, DIAMOND, [13.0], 1.0)
    assert hashlib.sha1(np.ascontiguousarray(eps).tobytes()).hexdigest() == cases[0]["eps_sha1"]
    ks = F.path_by_arclength(lo, cases[0]["path"], cases[0]["nk_path"])
    kidx = [c["k_index"] for c in cases]
    for c in cases:
        assert np.abs(ks[c["k_index"]] - np.array(c["k"])).max() < 1e-15
    nwant, nstep = cases[0]["nwant"], cases[0]["nstep"]
    ref = np.array([c["freq"] for c in cases]) ** 2
    S = F.Solver(fl, max_batch=len(kidx))
    S.set_eps(eps)
    S.set_wave_vectors(ks[kidx])
    for prec in (F.PREC_F64, F.PREC_F32_REFINE):
        print(f"FCC {nx}: {tuple(fl.grid_nums)} k {kidx} precision {prec}", flush=True)
        freq, _, E, prof = S.eigs(nwant, nstep, tol=cases[0]["tol"], precision=prec, want_fields=True)
        lam = freq ** 2
        err = np.abs(lam - ref) / ref
        print(f"  max rel eigenvalue error vs oracle {err.max():.2e}", flush=True)
        assert err.max() < 1e-8, (prec, err.max(axis=1))
        assert all(p["converged"] == 1 for p in prof)
        res = S.residuals(freq, E)
        # the oracle's own fields reach the same residual level (fixture)
        ores = np.array([c["rel_residual"] for c in cases])
        assert np.all(res < np.maximum(10 * ores, 1e-6 * lam)), (res, ores)
        del E
    # two inverse-iteration sweeps of the Ritz pairs bring the field residual
    # down to the FP64 floor, a small multiple of eps ||C^H C||
    S.set_refinement(2, 1e-16, 6000)
    bound = 100 * 2.220446049250313e-16 * float(sum((2.0 / h) ** 2 for h in fl.mesh_len))
    for prec in (F.PREC_F64, F.PREC_F32_REFINE):
        freq, _, E, prof = S.eigs(nwant, nstep, tol=cases[0]["tol"], precision=prec, want_fields=True)
        err = np.abs(freq ** 2 - ref) / ref
        res = S.residuals(freq, E)
        print(f"  refined, precision {prec}: eigenvalue error {err.max():.2e}, residual {res.max():.2e} "
              f"(bound {bound:.2e})", flush=True)
        assert err.max() < 1e-8
        assert np.all(res <= bound), (prec, res.max(), bound)
        del E
    S.close()
