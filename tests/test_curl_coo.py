"""The product's sparse curl (fame_curl_coo, the rule that the GPU
Check_Eigendecomp stencil applies) against the oracle's FAME_Matrix_Curl
restatement, entry by entry as sparse matrices, on Simple and General lattices
with and without boundary shifts (m1..m4), at Gamma and generic wave vectors.
The oracle's curl is itself pinned against the reference's compiled
FAME_Matrix_Curl (test_reference_build.py)."""
import numpy as np
import pytest
import scipy.sparse as sp

import fame_amd as F
import oracle as O

LATTICES = [("simple_cubic", 6, {}), ("face_centered_cubic", 8, {}), ("body_centered_cubic", 7, {}),
            ("hexagonal", 6, {"c": 1.2}), ("primitive_tetragonal", 6, {"c": 1.4}),
            ("primitive_monoclinic", 8, {"a": 1.0, "b": 1.0, "c": 1 / np.sin(1.4), "alpha": 1.4}),
            ("rhombohedral", 7, {"alpha": 1.1}),
            ("triclinic", 7, {"a": 1.0, "b": 1.1, "c": 1.3, "alpha": 1.3, "beta": 1.4, "gamma": 1.2})]
KS = [np.zeros(3), np.array([0.21, -0.13, 0.07]), np.array([0.5, 0.5, 0.5])]


def _mat(r, c, v, n):
    return sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()


@pytest.mark.parametrize("lt,nx,kw", LATTICES, ids=[f"{t[0]}-{t[1]}" for t in LATTICES])
def test_curl_coo_matches_oracle(lt, nx, kw):
    lo = F.build_lattice(lt, nx, **kw)
    fl = F.to_fame_lattice(lo)
    ol = O.lattice_from(fl)
    N = int(np.prod(list(fl.grid_nums)))
    for k in KS:
        r, c, v = F.curl_coo(fl, k)
        A = _mat(r, c, v, 3 * N)
        ro, co, vo = O.curl_coo(ol, k)
        B = _mat(ro, co, vo, 3 * N)
        d = abs(A - B)
        assert d.max() <= 1e-13 * abs(B).max(), (lt, k, d.max())
        assert A.nnz == B.nnz


def test_curl_coo_shifts_are_exercised():
    """At least one of the General lattices above has nonzero boundary shifts
    in each of the C2 (m1) and C3 (m2, m3, m4) rules."""
    seen = set()
    for lt, nx, kw in LATTICES:
        fl = F.to_fame_lattice(F.build_lattice(lt, nx, **kw))
        if fl.general:
            for name in ("m1", "m2", "m3", "m4"):
                if getattr(fl, name):
                    seen.add(name)
    assert seen == {"m1", "m2", "m3", "m4"}, seen
