"""Writes tests/golden/dense_eigs.json: brute-force eigenvalues of small
discretised Maxwell problems C^H C E = lambda B E.

C is the sparse Yee curl of the reference's FAME_Matrix_Curl.cpp as restated in
oracle/fame_oracle.cpp (itself compared entry by entry with the reference's own
compiled curl by tests/test_reference_build.py), B = diag(eps) from the
restated material mask.  The dense generalized problem is solved with
scipy.linalg.eigh; the N-dimensional null space (gradient fields) is dropped and
the smallest nonzero eigenvalues are stored.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np
import scipy.linalg as sl
import scipy.sparse as sp

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "fame-isotropic-gpu_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from cases import KPTS, make  # noqa: E402

CASES = [("sc", 6), ("fcc", 8), ("bcc", 8), ("hex", 7)]
NEIG = 8


def dense_case(sname, nx):
    import cases
    lt, _, kw, mat, ein = cases.STRUCTS[sname]
    saved = cases.STRUCTS[sname]
    cases.STRUCTS[sname] = (lt, nx, kw, mat, ein)
    try:
        lo, fl, ol, eps = make(sname)
    finally:
        cases.STRUCTS[sname] = saved
    N = ol.N[0] * ol.N[1] * ol.N[2]
    out = []
    for kname, k in KPTS.items():
        r, c, v = O.curl_coo(ol, k)
        C = sp.coo_matrix((v, (r, c)), shape=(3 * N, 3 * N)).toarray()
        w = sl.eigh(C.conj().T @ C, np.diag(eps), eigvals_only=True)
        nz = w[w > 1e-8 * w.max()]
        out.append({"struct": sname, "nx": nx, "grid": [ol.N[0], ol.N[1], ol.N[2]], "k_name": kname,
                    "k": list(map(float, k)), "eps_sha1": hashlib.sha1(eps.tobytes()).hexdigest(),
                    "n_zero": int(len(w) - len(nz)), "lambda": list(map(float, nz[:NEIG]))})
    return out


def main():
    cases = []
    for s, nx in CASES:
        cases += dense_case(s, nx)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dense_eigs.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {path}")


if __name__ == "__main__":
    main()
