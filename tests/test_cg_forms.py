"""The GPU CG's two fused iteration forms, restated in numpy on small Hermitian
positive-definite systems, against the reference CG's order of operations
(CG.cu:53-78: alpha = <r,r>/<p,Ap>; x += alpha p; r -= alpha Ap; beta from the
new <r,r>; p = r + beta p), in FP64 and FP32 vector arithmetic with FP64 dot
products.  The fused forms must stop after the same number of iterations and
reach the same residual level:

* "ap stored": the first pass applies x += alpha p and r -= alpha Ap of the
  previous iteration, sums <r,r>; the last pass sums <p,Ap> and <Ap,Ap>; the
  next <r,r> for beta and the stopping rule comes from one step of
  |r - alpha Ap|^2 = |alpha|^2 <Ap,Ap> - <r,r> (conjugacy: <r,Ap> = <p,Ap>);
* "residual on the last pass": alpha from <p,Ap> before the last pass, which
  stores r - alpha Ap and sums its exact norm; the first pass applies only
  x += alpha p and p = r + beta p.
A dropped term or a wrong sign in either form changes the iteration count or
stalls the residual."""
import numpy as np
import pytest


def _dot(u, v):
    return np.vdot(u.astype(np.complex128), v.astype(np.complex128))


def _system(n, kappa, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    a = (q * np.linspace(1.0 / kappa, 1.0, n)) @ q.conj().T
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return a, b


def cg_reference(a, b, tol, maxit, dt):
    a = a.astype(dt)
    x = np.zeros(len(b), dt)
    r = b.astype(dt)
    p = r.copy()
    r1 = _dot(r, r).real
    k = 1
    while r1 > tol * tol and k <= maxit:
        ap = (a @ p).astype(dt)
        alpha = r1 / _dot(p, ap)
        x = (x + alpha * p).astype(dt)
        r = (r - alpha * ap).astype(dt)
        r0, r1 = r1, _dot(r, r).real
        p = (r + (r1 / r0) * p).astype(dt)
        k += 1
    return x, k


def cg_ap_stored(a, b, tol, maxit, dt):
    a = a.astype(dt)
    x = np.zeros(len(b), dt)
    r = b.astype(dt)
    p = r.copy()
    ap = np.zeros(len(b), dt)
    alpha, beta, pend, k = 0.0, 0.0, False, 1
    active = _dot(r, r).real > tol * tol
    while active:
        if pend:
            x = (x + alpha * p).astype(dt)
            r = (r - alpha * ap).astype(dt)
        rr = _dot(r, r).real
        p = (r + beta * p).astype(dt)
        ap = (a @ p).astype(dt)
        alpha = rr / _dot(p, ap)
        pend = True
        s = max(0.0, abs(alpha) ** 2 * _dot(ap, ap).real - rr)
        beta = s / rr
        k += 1
        active = s > tol * tol and k <= maxit
    if pend:
        x = (x + alpha * p).astype(dt)
    return x, k


def cg_residual_last(a, b, tol, maxit, dt):
    a = a.astype(dt)
    x = np.zeros(len(b), dt)
    r = b.astype(dt)
    p = r.copy()
    rr = _dot(r, r).real
    alpha, beta, pend, k = 0.0, 0.0, False, 1
    active = rr > tol * tol
    while active:
        if pend:
            x = (x + alpha * p).astype(dt)
        p = (r + beta * p).astype(dt)
        ap = (a @ p).astype(dt)          # never stored on the GPU
        alpha = rr / _dot(p, ap).real    # <p,Ap> as the real quadratic form
        pend = True
        r = (r - alpha * ap).astype(dt)
        s = _dot(r, r).real
        beta, rr = s / rr, s
        k += 1
        active = s > tol * tol and k <= maxit
    if pend:
        x = (x + alpha * p).astype(dt)
    return x, k


@pytest.mark.parametrize("kappa", [13.0, 400.0])
@pytest.mark.parametrize("dt,reltol", [(np.complex128, 1e-11), (np.complex64, 1e-6)])
@pytest.mark.parametrize("form", [cg_ap_stored, cg_residual_last])
def test_fused_forms_match_reference_cg(kappa, dt, reltol, form):
    a, b = _system(200, kappa, 3)
    tol = reltol * np.linalg.norm(b)
    x0, k0 = cg_reference(a, b, tol, 2000, dt)
    x1, k1 = form(a, b, tol, 2000, dt)
    assert abs(k1 - k0) <= 1, (k0, k1)
    r0 = np.linalg.norm(b - a @ x0)
    r1 = np.linalg.norm(b - a @ x1)
    assert r1 <= 2.0 * max(r0, tol), (r0, r1)


def test_a_wrong_recurrence_sign_is_caught():
    """The one-step recurrence with +<r,r> instead of -<r,r> no longer tracks
    the residual: the stopping rule then fires at the wrong iteration."""
    a, b = _system(200, 13.0, 3)
    tol = 1e-11 * np.linalg.norm(b)
    _, k0 = cg_reference(a, b, tol, 2000, np.complex128)

    def wrong(a, b, tol, maxit, dt):
        a = a.astype(dt)
        x = np.zeros(len(b), dt); r = b.astype(dt); p = r.copy(); ap = np.zeros(len(b), dt)
        alpha, beta, pend, k = 0.0, 0.0, False, 1
        active = True
        while active:
            if pend:
                x = x + alpha * p; r = r - alpha * ap
            rr = _dot(r, r).real
            p = r + beta * p; ap = a @ p
            alpha = rr / _dot(p, ap); pend = True
            s = max(0.0, abs(alpha) ** 2 * _dot(ap, ap).real + rr)
            beta = s / rr; k += 1
            active = s > tol * tol and k <= maxit
        return x, k

    _, kw = wrong(a, b, tol, 2000, np.complex128)
    assert abs(kw - k0) > 1
