"""Writes tests/golden/oracle_eigs_fcc.json: eigenvalues of the FCC diamond-network
workload (two spheres r=0.15 and four rods r=0.10, eps 13/1) computed by the FP64
CPU oracle (oracle/fame_oracle.cpp: the restatement of the reference's Lanczos /
lock-purge / CG / QBQ path, itself pinned by tests/test_oracle.py) at the
reference-adjusted grids of the 64 and 96 configurations:

  * 64 -> 64x55x52: 2 wave vectors of the 30-point Gamma-X-W-L-Gamma path, 12 bands, nstep 36
  * 96 -> 96x83x78: 2 wave vectors of the same path, 16 bands, nstep 48

Lanczos and CG tolerance 1e-12 (the workload's).  Besides the frequencies it
stores the oracle's relative residual ||C^H C E - omega^2 B E|| / ||E|| per band
and the iteration counts.  The GPU tests hold the HIP path's FP64 and
FP32-refined eigenvalues to these at 1e-8 relative.  Takes about half an hour on
8 cores.  Run from the repository root:

    python tests/golden/make_oracle_eigs.py [threads]
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "fame-isotropic-gpu_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import fame_amd as F  # noqa: E402
import oracle as O  # noqa: E402

DIAMOND = dict(sphere_centers=[[(0, 0, 0), (0.25, 0.25, 0.25)]], sphere_radius=[[0.15, 0.15]],
               cylinder_top=[[(0.25, 0.25, 0.25)] * 4],
               cylinder_bot=[[(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]], cylinder_radius=[[0.10] * 4])
NK_PATH = 30
CASES = [(64, 12, 36, (7, 22)), (96, 16, 48, (3, 17))]


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    O.set_threads(threads)
    out = []
    for nx, nwant, nstep, kidx in CASES:
        lo = F.build_lattice("face_centered_cubic", nx)
        fl = F.to_fame_lattice(lo)
        eps = F.material_eps(lo, F.Material(**DIAMOND), [13.0], 1.0)
        ks = F.path_by_arclength(lo, "GXWLG", NK_PATH)
        ol = O.lattice_from(fl)
        for ik in kidx:
            t0 = time.time()
            freq, E, prof = O.eigs(ol, eps, ks[ik], nwant, nstep, tol=1e-12, maxit=300, fields=True)
            t = O.Tables(ol, ks[ik])
            res = [float(t.rel_residual(eps, E[b], freq[b])) for b in range(nwant)]
            del E
            out.append({"nx": nx, "grid": list(map(int, fl.grid_nums)), "path": "GXWLG", "nk_path": NK_PATH,
                        "k_index": ik, "k": list(map(float, ks[ik])),
                        "eps_sha1": hashlib.sha1(np.ascontiguousarray(eps).tobytes()).hexdigest(),
                        "nwant": nwant, "nstep": nstep, "tol": 1e-12, "freq": list(map(float, freq)),
                        "rel_residual": res, "profile": prof, "seconds": time.time() - t0, "threads": threads})
            print(f"{nx} k{ik}: {time.time() - t0:.0f}s freq {freq[:4]} ... max rel res {max(res):.2e} {prof}",
                  flush=True)
            path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_eigs_fcc.json")
            with open(path, "w") as f:
                json.dump({"generator": "tests/golden/make_oracle_eigs.py", "cases": out}, f, indent=1)


if __name__ == "__main__":
    main()
