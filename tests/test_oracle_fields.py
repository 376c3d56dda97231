"""CPU pins of the oracle's field outputs (no GPU needed).

  orc_rel_residual vs scipy.sparse    -> ||C^H C E - omega^2 B E|| / ||E|| built from the
                                         COO curl (itself pinned entry by entry against the
                                         reference's compiled FAME_Matrix_Curl); catches
                                         omega for omega^2, a dropped B, a wrong norm
  orc_eigs fields vs dense eigh       -> the restored field E = B^-1 Qr(Sigma y) of each
                                         non-degenerate band is the dense generalized
                                         eigenvector up to phase; its norm is ||E_raw||^2,
                                         the reference's multiply-by-the-norm scaling
                                         (FAME_Fast_Algorithms_Isotropic.cu:119-121)
  Gamma rule on the fields            -> E_0 = E_1 = 1/sqrt(3N) and bands moved up by two
                                         (FAME_Fast_Algorithms_Isotropic.cu:61-79)
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

import oracle as O
from cases import KPTS, STRUCTS

import cases

SIZES = {"sc": 6, "fcc": 8, "bcc": 8, "hex": 7}


def _make(sname):
    lt, _, kw, mat, ein = STRUCTS[sname]
    saved = STRUCTS[sname]
    cases.STRUCTS[sname] = (lt, SIZES[sname], kw, mat, ein)
    try:
        return cases.make(sname)
    finally:
        cases.STRUCTS[sname] = saved


def _curl(ol, k, N):
    r, c, v = O.curl_coo(ol, k)
    return sp.coo_matrix((v, (r, c)), shape=(3 * N, 3 * N)).tocsr()


def _rand(n, seed):
    g = np.random.default_rng(seed)
    return g.standard_normal(n) + 1j * g.standard_normal(n)


@pytest.mark.parametrize("kname", list(KPTS))
@pytest.mark.parametrize("sname", list(STRUCTS))
def test_rel_residual_vs_sparse(sname, kname):
    lo, fl, ol, eps = _make(sname)
    k = KPTS[kname]
    t = O.Tables(ol, k)
    N = t.N
    C = _curl(ol, k, N)
    for seed, omega in ((1, 0.7), (2, 3.1)):
        E = _rand(3 * N, seed)
        r = C.conj().T @ (C @ E) - omega ** 2 * eps * E
        ref = np.linalg.norm(r) / np.linalg.norm(E)
        got = t.rel_residual(eps, E, omega)
        assert abs(got - ref) <= 1e-12 * ref, (got, ref)
    # a planted mistake: omega instead of omega^2 changes the value
    E = _rand(3 * N, 3)
    wrong = np.linalg.norm(C.conj().T @ (C @ E) - 3.1 * eps * E) / np.linalg.norm(E)
    assert abs(t.rel_residual(eps, E, 3.1) - wrong) > 1e-3 * wrong


def _dense(ol, eps, k, N):
    C = _curl(ol, k, N).toarray()
    w, V = sl.eigh(C.conj().T @ C, np.diag(eps))
    keep = w > 1e-8 * w.max()
    return w[keep], V[:, keep]


@pytest.mark.parametrize("kname", list(KPTS))
@pytest.mark.parametrize("sname", list(STRUCTS))
def test_oracle_fields_vs_dense(sname, kname):
    lo, fl, ol, eps = _make(sname)
    k = KPTS[kname]
    t = O.Tables(ol, k)
    N = t.N
    nwant = 6
    freq, E, _ = O.eigs(ol, eps, k, nwant, 20, fields=True)
    lam_d, V = _dense(ol, eps, k, N)
    gamma = kname == "gamma"
    shift = 2 if gamma else 0
    if gamma:
        const = np.full(3 * N, 1.0 / np.sqrt(3 * N))
        assert freq[0] == 0 and freq[1] == 0
        assert np.array_equal(E[0], const) and np.array_equal(E[1], const)
    sig = np.tile(t.get(3), 2)
    C = _curl(ol, k, N)
    checked = 0
    for b in range(shift, nwant):
        j = b - shift
        lam = lam_d[j]
        gap = min(abs(lam - lam_d[j - 1]) if j > 0 else np.inf, abs(lam_d[j + 1] - lam))
        assert abs(freq[b] ** 2 - lam) <= 1e-10 * lam
        # oracle residual of its own field at the Lanczos stopping rule's level
        assert t.rel_residual(eps, E[b], freq[b]) < 1e-6 * lam
        if gap <= 1e-6 * lam:
            continue                        # degenerate pair: the field is a basis choice
        u = E[b] / np.linalg.norm(E[b])
        v = V[:, j] / np.linalg.norm(V[:, j])
        ph = np.vdot(v, u)
        assert abs(abs(ph) - 1) < 1e-9
        assert np.linalg.norm(u - ph / abs(ph) * v) < 1e-7, (sname, kname, b)
        # multiply-by-the-norm scaling: ||E|| = ||E_raw||^2, E_raw = B^-1 Qr(Sigma y), ||y|| = 1.
        # With y c = Sigma^-1 Qrs(B V_j) (Qrs = N Qr^H, Q~ = sqrt(N) Qr orthonormal) E_raw = V_j / c.
        c = np.linalg.norm(t.qrs(eps * V[:, j]) / sig)
        raw = np.linalg.norm(V[:, j]) / c
        assert abs(np.linalg.norm(E[b]) / raw ** 2 - 1) < 1e-8, (np.linalg.norm(E[b]), raw ** 2)
        # and the dense vector satisfies the residual the oracle measures
        assert t.rel_residual(eps, V[:, j], np.sqrt(lam)) < 1e-10 * lam
        r = C.conj().T @ (C @ V[:, j]) - lam * eps * V[:, j]
        assert np.linalg.norm(r) < 1e-9 * lam * np.linalg.norm(V[:, j])
        checked += 1
    assert checked >= 2
