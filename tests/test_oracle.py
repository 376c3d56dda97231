"""CPU checks that pin the FP64 oracle (no GPU needed).

What each check can catch:
  FFT vs numpy.fft                  -> wrong sign, stride, radix pass or Bluestein chirp
  FFT curl vs sparse curl           -> wrong Lambda formula, D twiddle, layout or Gamma rule
  C = P Sigma Q^H, Qrs Qr = I       -> wrong Pi_Pr / Pi_Qrs permutation or Sigma placement
  Pi orthonormal, Lambda^H Pi = 0   -> a dropped term or conjugate in the Pi formulas
  equal mesh lengths, Lambda || (1,1,1) -> the General gauge's 0/0 modes (guarded basis)
  homogeneous QBQ = I / eps         -> a missing 1/N or 1/eps
  tridiagonal eigensolver vs LAPACK -> ordering, eigenvector columns
  GVqrrq vs a dense shifted QR step -> Givens sign convention, symmetrisation
  Lanczos eigenvalues vs dense eigh -> anything in CG / Lanczos / lock-purge / Gamma rule
  CG solves QBQ x = b, invAr inverts Sigma QBQ Sigma -> a wrong CG update or Sigma placement
  a planted Lambda error            -> shows the curl check is sensitive to the tables
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

import oracle as O
from cases import KPTS, STRUCTS, equal_mesh_general, make

HERE = os.path.dirname(os.path.abspath(__file__))


def _rand(n, seed=0):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


@pytest.mark.parametrize("n", list(range(1, 70)) + [83, 97, 105, 107, 121, 128, 131, 168, 192, 243, 256])
def test_fft_vs_numpy(n):
    x = np.stack([_rand(n, s) for s in range(3)])
    f = O.fft(x, -1)
    b = O.fft(x, +1)
    assert np.abs(f - np.fft.fft(x, axis=1)).max() <= 1e-13 * max(1.0, np.abs(x).sum())
    assert np.abs(b - np.fft.ifft(x, axis=1) * n).max() <= 1e-13 * max(1.0, np.abs(x).sum())


@pytest.mark.parametrize("kname", list(KPTS))
@pytest.mark.parametrize("sname", list(STRUCTS))
def test_factorisation(sname, kname):
    lo, fl, ol, eps = make(sname)
    k = KPTS[kname]
    t = O.Tables(ol, k)
    N = t.N
    r, c, v = O.curl_coo(ol, k)
    C = sp.coo_matrix((v, (r, c)), shape=(3 * N, 3 * N)).tocsr()
    E = _rand(3 * N, 1)
    CE = C @ E
    assert np.abs(t.curl(E) - CE).max() <= 1e-12 * np.abs(CE).max()
    # C^H through the factorisation
    assert np.abs(t.curlh(E) - C.conj().T @ E).max() <= 1e-12 * np.abs(C.conj().T @ E).max()
    # C = P Sigma Q^H  (Check_Eigendecomp, FAME_Main_Code.cu:320-362, with Qrs = N Qr^H)
    sig = t.get(3)
    q = t.qrs(E)
    q[: t.Nd] *= sig
    q[t.Nd:] *= sig
    assert np.abs(t.pr(q) - CE).max() <= 1e-12 * np.abs(CE).max()
    x = _rand(2 * t.Nd, 2)
    assert np.abs(t.qrs(t.qr(x)) - x).max() <= 1e-12 * np.abs(x).max()
    assert np.abs(t.qbq(np.full(3 * N, 1 / 13.0), x) - x / 13.0).max() <= 1e-13 * np.abs(x).max()
    # Pi bases
    Pi = t.get(4).reshape(6, t.Nd)
    Lx, Ly, Lz = t.get(0), t.get(1), t.get(2)
    P1, P2 = Pi[0:3], Pi[3:6]
    one = np.ones(t.Nd)
    assert np.abs(np.sum(np.abs(P1) ** 2, 0) - one).max() < 1e-12
    assert np.abs(np.sum(np.abs(P2) ** 2, 0) - one).max() < 1e-12
    assert np.abs(np.sum(P1.conj() * P2, 0)).max() < 1e-12
    L = np.stack([Lx, Ly, Lz])
    scale = np.sqrt(np.sum(np.abs(L) ** 2, 0))
    assert np.abs(np.sum(L.conj() * P1, 0) / scale).max() < 1e-12
    assert np.abs(np.sum(L.conj() * P2, 0) / scale).max() < 1e-12
    assert np.abs(sig - scale).max() < 1e-12 * scale.max()


def test_planted_lambda_error_is_caught():
    """Perturbing Lambda_x by one part in 1e9 must break the curl identity."""
    lo, fl, ol, eps = make("fcc")
    k = KPTS["k"]
    good = O.Tables(ol, k)
    h = list(ol.h)
    ol.h[0] = h[0] * (1 + 1e-9)
    bad = O.Tables(ol, k)
    ol.h[0] = h[0]
    r, c, v = O.curl_coo(ol, k)
    N = good.N
    C = sp.coo_matrix((v, (r, c)), shape=(3 * N, 3 * N)).tocsr()
    E = _rand(3 * N, 3)
    CE = C @ E
    assert np.abs(good.curl(E) - CE).max() <= 1e-12 * np.abs(CE).max()
    assert np.abs(bad.curl(E) - CE).max() > 1e-11 * np.abs(CE).max()


@pytest.mark.parametrize("n", [2, 5, 12, 30, 48])
def test_tridiagonal_vs_lapack(n):
    rng = np.random.default_rng(n)
    d = rng.uniform(1.0, 2.0, n)
    e = rng.uniform(-0.3, 0.3, n - 1)
    w, Z, st = O.tridiag_eig(d, e)
    assert st == 0
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ref = np.sort(np.linalg.eigvalsh(T))[::-1]
    assert np.abs(w - ref).max() < 1e-13
    assert np.all(np.diff(w) <= 0)
    assert np.abs(T @ Z - Z * w).max() < 1e-13
    # zpteqr (the reference's routine, Lanczos_Isotropic.cu:53) gives the same order
    from scipy.linalg import lapack
    dd, ee, zz, info = lapack.zpteqr(d.copy(), e.copy(), np.eye(n, dtype=complex), compute_z=2)
    assert info == 0
    assert np.abs(dd - w).max() < 1e-12
    assert np.abs(np.abs(np.real(zz)) - np.abs(Z)).max() < 1e-8


def test_gvqrrq_is_a_shifted_qr_step():
    import ctypes as C
    n = 9
    rng = np.random.default_rng(4)
    d = rng.uniform(1, 3, n)
    e = rng.uniform(0.1, 0.5, n - 1)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    mu = float(np.linalg.eigvalsh(T)[3])
    T0 = d.copy(); T1 = np.zeros(n); T1[: n - 1] = e
    T2 = np.zeros(n); T3 = np.zeros(n); c = np.zeros(n); s = np.zeros(n)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    O.lib().orc_gvqrrq(p(T0), p(T1), p(T2), p(T3), p(c), p(s), mu, n)
    Q, R = np.linalg.qr(T - mu * np.eye(n))
    Tn = R @ Q + mu * np.eye(n)
    assert np.abs(np.abs(np.diag(Tn)) - np.abs(T0)).max() < 1e-12 or np.abs(np.diag(Tn) - T0).max() < 1e-12
    assert np.abs(np.abs(np.diag(Tn, 1)) - np.abs(T1[: n - 1])).max() < 1e-12
    assert np.abs(np.sort(np.linalg.eigvalsh(np.diag(T0) + np.diag(T1[:n-1], 1) + np.diag(T1[:n-1], -1)))
                  - np.sort(np.linalg.eigvalsh(T))).max() < 1e-12


def _golden():
    with open(os.path.join(HERE, "golden", "dense_eigs.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _golden(), ids=lambda c: f"{c['struct']}-{c['k_name']}")
def test_oracle_eigs_vs_dense(case):
    import cases
    import hashlib
    lt, _, kw, mat, ein = cases.STRUCTS[case["struct"]]
    saved = cases.STRUCTS[case["struct"]]
    cases.STRUCTS[case["struct"]] = (lt, case["nx"], kw, mat, ein)
    try:
        lo, fl, ol, eps = make(case["struct"])
    finally:
        cases.STRUCTS[case["struct"]] = saved
    assert hashlib.sha1(eps.tobytes()).hexdigest() == case["eps_sha1"]
    k = np.array(case["k"])
    nwant = 6
    freq, _, prof = O.eigs(ol, eps, k, nwant, 20)
    lam = freq ** 2
    gamma = case["k_name"] == "gamma"
    ref = np.concatenate([[0.0, 0.0], case["lambda"][: nwant - 2]]) if gamma else np.array(case["lambda"][:nwant])
    nz = ref > 0
    assert np.all(lam[~nz] == 0)
    assert np.max(np.abs(lam[nz] - ref[nz]) / ref[nz]) < 1e-10


@pytest.mark.parametrize("sname", ["sc", "fcc"])
def test_cg_and_invar_solve(sname):
    """orc_cg returns x with ||QBQ x - b|| at the requested absolute tolerance
    (CG.cu stops on <r,r> <= tol^2), and orc_invar returns y with
    Sigma QBQ Sigma y = x (A_r^{-1}, ..._Isotropic_invAr.cu:8-48)."""
    lo, fl, ol, eps = make(sname)
    t = O.Tables(ol, KPTS["k"])
    invB = 1.0 / eps
    b = _rand(2 * t.Nd, 4) * 1e-2
    x, it, res = t.cg(invB, b, 2000, 1e-12)
    assert 1 < it < 2000
    assert np.linalg.norm(t.qbq(invB, x) - b) <= 2e-12
    sig = np.tile(t.get(3), 2)
    y, _ = t.invar(invB, b, 2000, 1e-13)
    assert np.linalg.norm(sig * t.qbq(invB, sig * y) - b) <= 1e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("shifts", [(0, 0, 0, 0), (0, 2, 0, 2)])
@pytest.mark.parametrize("k", [(0.0, 0.0, 0.0), (0.3, 0.3, 0.3), (0.21, -0.13, 0.07)])
def test_pi_gauge_where_lambda_is_parallel_to_ones(shifts, k):
    """General grid with three equal mesh lengths: at Gamma some nonzero modes
    have Lambda exactly parallel to (1, 1, 1), where the reference's General
    normalisers (FAME_Matrix_Lambdas.cu:441-461) are 0/0; at (0.3, 0.3, 0.3)
    some are nearly parallel.  The oracle's basis must stay orthonormal and
    orthogonal to Lambda there, and the factorisation C = P Sigma Q^H hold."""
    fl, ol, eps = equal_mesh_general(shifts)
    k = np.array(k)
    t = O.Tables(ol, k)
    Lx, Ly, Lz = t.get(0), t.get(1), t.get(2)
    L = np.stack([Lx, Ly, Lz])
    Lq = np.sum(np.abs(L) ** 2, 0)
    d = np.abs(Lz - Ly) ** 2 + np.abs(Lx - Lz) ** 2 + np.abs(Ly - Lx) ** 2
    if not k.any():
        assert np.sum(d <= 1e-30 * Lq) >= 1          # the degenerate modes are present
    Pi = t.get(4).reshape(6, t.Nd)
    assert np.isfinite(Pi).all()
    P1, P2 = Pi[0:3], Pi[3:6]
    # Away from the exact 0/0 the oracle keeps the reference's General formulas,
    # whose normalisers cancel to |d|^2 / Lq: their rounding grows as eps Lq / |d|^2.
    # The exactly parallel modes take the Simple gauge, accurate to rounding.
    guarded = d <= 1e-30 * Lq
    tol = 1e-13 + np.where(guarded, 0.0, 50 * 2.2e-16 * Lq / np.maximum(d, 1e-300))
    assert np.all(np.abs(np.sum(np.abs(P1) ** 2, 0) - 1) < tol)
    assert np.all(np.abs(np.sum(np.abs(P2) ** 2, 0) - 1) < tol)
    assert np.all(np.abs(np.sum(P1.conj() * P2, 0)) < tol)
    assert np.all(np.abs(np.sum(L.conj() * P1, 0) / np.sqrt(Lq)) < tol)
    assert np.all(np.abs(np.sum(L.conj() * P2, 0) / np.sqrt(Lq)) < tol)
    N = t.N
    r, c, v = O.curl_coo(ol, k)
    C = sp.coo_matrix((v, (r, c)), shape=(3 * N, 3 * N)).tocsr()
    E = _rand(3 * N, 1)
    CE = C @ E
    q = t.qrs(E)
    sig = t.get(3)
    q[: t.Nd] *= sig
    q[t.Nd:] *= sig
    assert np.abs(t.pr(q) - CE).max() <= 1e-12 * np.abs(CE).max()
    x = _rand(2 * t.Nd, 2)
    # the nearly parallel modes' basis is orthonormal only to their tolerance above
    assert np.abs(t.qrs(t.qr(x)) - x).max() <= 2e-11 * np.abs(x).max()


def test_eigs_where_lambda_is_parallel_to_ones():
    """The oracle's eigenvalues on the equal-mesh General grid at Gamma (with
    exactly degenerate Pi modes) against a dense generalized eigensolve of
    C^H C E = lambda B E with the sparse curl."""
    fl, ol, eps = equal_mesh_general((0, 2, 0, 2), n=6)
    k = np.zeros(3)
    nwant = 8
    freq, _, prof = O.eigs(ol, eps, k, nwant, 24)
    N = ol.N[0] * ol.N[1] * ol.N[2]
    r, c, v = O.curl_coo(ol, k)
    C = sp.coo_matrix((v, (r, c)), shape=(3 * N, 3 * N)).toarray()
    w = sl.eigh(C.conj().T @ C, np.diag(eps), eigvals_only=True)
    w = np.sort(w[w > 1e-8 * w.max()])
    lam = freq ** 2
    assert np.all(lam[:2] == 0)
    assert np.max(np.abs(lam[2:] - w[: nwant - 2]) / w[: nwant - 2]) < 1e-10
