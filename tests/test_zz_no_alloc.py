"""fame_residuals_noalloc (the residual check without device allocation, its
fields staged in the library's static device scratch) against fame_residuals
on the same fields: two calls in a row, the second with a smaller input
(fewer wave vectors on a smaller grid), so that the second call reuses the
scratch the first one filled.  Also the size limit of the scratch."""
import numpy as np
import pytest

import fame_amd as F
from cases import KPTS, make

pytestmark = pytest.mark.gpu


def _solve(sname, ks, nwant=6, nstep=20):
    lo, fl, ol, eps = make(sname)
    S = F.Solver(fl, max_batch=len(ks))
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    freq, _, E, _ = S.eigs(nwant, nstep, want_fields=True)
    return S, freq, E


def test_noalloc_residuals_match_allocating_entry_point(gpu):
    import torch
    calls = [("fcc", np.stack([KPTS["k"], KPTS["gamma"], np.array([0.05, 0.31, -0.12])])),
             ("sc", np.stack([KPTS["k"]]))]
    sizes = []
    for sname, ks in calls:
        print(f"no-alloc residuals: {sname} with {len(ks)} wave vectors", flush=True)
        S, freq, E = _solve(sname, ks)
        sizes.append(S.N * len(ks))
        ref = S.residuals(freq, E)
        torch.cuda.synchronize()
        got = S.residuals_noalloc(freq, E)
        assert got.shape == ref.shape
        assert np.all(np.isfinite(got))
        # the same kernels and the same fixed-order sums: identical values
        assert np.array_equal(got, ref), (sname, np.abs(got - ref).max())
        # and a field that is not an eigenvector (residual of order ||C^H C||)
        Er = torch.randn_like(E.real) + 1j * torch.randn_like(E.real)
        assert np.array_equal(S.residuals_noalloc(freq, Er), S.residuals(freq, Er)), sname
        S.close()
    assert sizes[1] < sizes[0]


def test_noalloc_rejects_grids_beyond_the_scratch(gpu):
    lo = F.build_lattice("simple_cubic", 144)          # N = 2,985,984 > 2,796,202
    fl = F.to_fame_lattice(lo)
    S = F.Solver(fl, max_batch=1)
    S.set_eps(np.ones(3 * S.N))
    S.set_wave_vectors(np.array([0.1, 0.2, 0.3]))
    import torch
    E = torch.zeros((1, 1, 3 * S.N), dtype=torch.complex128, device="cuda")
    with pytest.raises(F.FameError, match="exceeds the static scratch"):
        S.residuals_noalloc(np.ones((1, 1)), E)
    S.close()
