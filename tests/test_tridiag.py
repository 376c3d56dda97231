"""The product library's host routines of the Lanczos restart, on the CPU:
fame_tridiag_eig (replaces LAPACKE_zpteqr, FAME_Main/Lanczos_Isotropic.cu:48-53)
and fame_tridiag_qr_step (replaces GVqrrq_g, FAME_Main/Lanczos_LockPurge_gpu.cu:95-190),
against LAPACK zpteqr / numpy eigh and a dense shifted QR step."""
import ctypes as C

import numpy as np
import pytest
from scipy.linalg import lapack

import fame_amd as F


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _lib():
    L = F.lib()
    L.fame_tridiag_eig.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.fame_tridiag_eig.restype = C.c_int
    L.fame_tridiag_qr_step.argtypes = [C.c_void_p] * 6 + [C.c_double, C.c_int]
    L.fame_tridiag_qr_step.restype = C.c_int
    return L


def tridiag_eig(d, e):
    n = len(d)
    w = np.zeros(n); z = np.zeros((n, n))
    ee = np.zeros(max(n - 1, 1)); ee[: n - 1] = e
    st = _lib().fame_tridiag_eig(n, _p(np.ascontiguousarray(d, float)), _p(ee), _p(w), _p(z))
    return w, z.T.copy(), st      # z is column-major: row j of the buffer is eigenvector j


def qr_step(d, e, mu):
    n = len(d)
    dd = np.array(d, float); ee = np.zeros(n); ee[: n - 1] = e
    ru = np.zeros(n); rw = np.zeros(n); c = np.zeros(n); s = np.zeros(n)
    st = _lib().fame_tridiag_qr_step(_p(dd), _p(ee), _p(ru), _p(rw), _p(c), _p(s), mu, n)
    return dd, ee[: n - 1], c[: n - 1], s[: n - 1], st


def _tri(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


@pytest.mark.parametrize("n", [1, 2, 5, 12, 30, 48, 73, 127])
def test_tridiag_eig_vs_lapack(n):
    rng = np.random.default_rng(100 + n)
    d = rng.uniform(1.0, 2.0, n)
    e = rng.uniform(-0.3, 0.3, n - 1)
    w, Z, st = tridiag_eig(d, e)
    assert st == 0
    T = _tri(d, e)
    ref = np.sort(np.linalg.eigvalsh(T))[::-1]
    assert np.abs(w - ref).max() < 1e-13
    assert np.all(np.diff(w) <= 0)                                   # descending, as zpteqr
    assert np.abs(T @ Z - Z * w).max() < 1e-13
    assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-13
    if n > 1:
        dd, ee, zz, info = lapack.zpteqr(d.copy(), e.copy(), np.eye(n, dtype=complex), compute_z=2)
        assert info == 0
        assert np.abs(dd - w).max() < 1e-12
        assert np.abs(np.abs(np.real(zz)) - np.abs(Z)).max() < 1e-8


def test_tridiag_eig_clustered_and_graded():
    # Lanczos T near convergence: tiny off-diagonals and a clustered pair
    d = np.array([3.0, 3.0 + 1e-13, 1.0, 0.5, 0.25, 1e-3])
    e = np.array([1e-14, 1e-9, 0.2, 1e-16, 0.1])
    w, Z, st = tridiag_eig(d, e)
    T = _tri(d, e)
    assert st == 0
    assert np.abs(w - np.sort(np.linalg.eigvalsh(T))[::-1]).max() < 1e-14
    assert np.abs(T @ Z - Z * w).max() < 1e-14


def _dense_qr_step(T, mu):
    """T - mu I = QR with R's diagonal made non-negative (Givens with r = hypot >= 0)."""
    n = T.shape[0]
    Q, R = np.linalg.qr(T - mu * np.eye(n))
    sg = np.where(np.diag(R) < 0, -1.0, 1.0)
    Q, R = Q * sg, sg[:, None] * R
    return Q, R, R @ Q + mu * np.eye(n)


@pytest.mark.parametrize("n", [2, 3, 9, 30, 127])
def test_qr_step_matches_dense_shifted_qr(n):
    rng = np.random.default_rng(n)
    d = rng.uniform(1, 3, n)
    e = rng.uniform(0.1, 0.5, n - 1) * rng.choice([-1, 1], n - 1)
    T = _tri(d, e)
    mu = 0.37 * d.min()                       # generic shift: R nonsingular, Q unique
    Q, R, Tn = _dense_qr_step(T, mu)
    d1, e1, c, s, st = qr_step(d, e, mu)
    assert st == 0
    assert np.abs(d1 - np.diag(Tn)).max() < 1e-13
    assert np.abs(e1 - 0.5 * (np.diag(Tn, 1) + np.diag(Tn, -1))).max() < 1e-13
    # rotations: Q = G_0^T ... G_{n-2}^T with G_i = [[c, -s], [s, c]] on (i, i+1)
    Qg = np.eye(n)
    for i in range(n - 1):
        G = np.eye(n)
        G[i, i], G[i, i + 1], G[i + 1, i], G[i + 1, i + 1] = c[i], -s[i], s[i], c[i]
        Qg = Qg @ G.T
    assert np.abs(Qg - Q).max() < 1e-12
    assert np.abs(c * c + s * s - 1).max() < 1e-15
    # spectrum preserved
    assert np.abs(np.sort(np.linalg.eigvalsh(_tri(d1, e1))) - np.sort(np.linalg.eigvalsh(T))).max() < 1e-12


def test_qr_step_exact_shift_deflates():
    """The restart's use: with an exact eigenvalue as shift, the last row decouples
    and carries the shift (Lanczos_LockPurge_gpu.cu:23-45 then purges it)."""
    n = 12
    rng = np.random.default_rng(7)
    d = rng.uniform(1, 3, n)
    e = rng.uniform(0.1, 0.5, n - 1)
    T = _tri(d, e)
    lam = np.linalg.eigvalsh(T)
    mu = float(lam[4])
    d1, e1, c, s, st = qr_step(d, e, mu)
    assert st == 0
    assert abs(e1[-1]) < 1e-12
    assert abs(d1[-1] - mu) < 1e-12
    rest = np.sort(np.linalg.eigvalsh(_tri(d1[:-1], e1[:-1])))
    assert np.abs(rest - np.delete(lam, 4)).max() < 1e-12


def test_qr_step_zero_column_and_args():
    # a zero pivot and zero sub-diagonal (r = 0) is an identity rotation, not a NaN
    d = np.array([0.0, 1.0, 2.0]); e = np.array([0.0, 0.5])
    d1, e1, c, s, st = qr_step(d, e, 0.0)
    assert st == 0 and np.all(np.isfinite(d1)) and np.all(np.isfinite(e1))
    assert c[0] == 1.0 and s[0] == 0.0
    assert np.abs(np.sort(np.linalg.eigvalsh(_tri(d1, e1))) - np.sort(np.linalg.eigvalsh(_tri(d, e)))).max() < 1e-14
    assert _lib().fame_tridiag_qr_step(None, None, None, None, None, None, 0.0, 0) == F.FAME_ERR_ARG
    assert _lib().fame_tridiag_eig(0, None, None, None, None) == F.FAME_ERR_ARG


def test_qr_step_agrees_with_oracle_step():
    """The oracle's step (test infrastructure) and the product's step give the same T'."""
    import oracle as O
    n = 10
    rng = np.random.default_rng(11)
    d = rng.uniform(1, 3, n); e = rng.uniform(0.05, 0.5, n - 1)
    mu = float(np.linalg.eigvalsh(_tri(d, e))[2])
    T0 = d.copy(); T1 = np.zeros(n); T1[: n - 1] = e
    T2 = np.zeros(n); T3 = np.zeros(n); c = np.zeros(n); s = np.zeros(n)
    O.lib().orc_gvqrrq(_p(T0), _p(T1), _p(T2), _p(T3), _p(c), _p(s), mu, n)
    d1, e1, c1, s1, _ = qr_step(d, e, mu)
    assert np.abs(d1 - T0).max() < 1e-13
    assert np.abs(e1 - T1[: n - 1]).max() < 1e-13
    assert np.abs(c1 - c[: n - 1]).max() < 1e-13 and np.abs(s1 - s[: n - 1]).max() < 1e-13
