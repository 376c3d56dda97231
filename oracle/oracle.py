"""ctypes view of oracle/liboracle.so -- the FP64 CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libfame_ref_host.so")


class orc_lattice(C.Structure):
    _fields_ = [("N", C.c_int * 3), ("general", C.c_int), ("h", C.c_double * 3), ("a", C.c_double * 9),
                ("m1", C.c_int), ("m2", C.c_int), ("m3", C.c_int), ("m4", C.c_int),
                ("t1", C.c_double * 3), ("t2", C.c_double * 3), ("t3", C.c_double * 3), ("t4", C.c_double * 3)]


class orc_es(C.Structure):
    _fields_ = [("nwant", C.c_int), ("nstep", C.c_int), ("maxit", C.c_int), ("tol", C.c_double)]


class orc_ls(C.Structure):
    _fields_ = [("maxit", C.c_int), ("tol", C.c_double)]


class orc_profile(C.Structure):
    _fields_ = [("ls_iter", C.c_int), ("es_iter", C.c_int), ("restarts", C.c_int), ("nd", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        L.orc_fft.argtypes = [i, i, vp, i]
        L.orc_set_threads.argtypes = [i]
        L.orc_tables_create.argtypes = [C.POINTER(orc_lattice), vp]
        L.orc_tables_create.restype = vp
        L.orc_tables_destroy.argtypes = [vp]
        L.orc_tables_nd.argtypes = [vp]
        L.orc_tables_get.argtypes = [vp, i, vp]
        for n in ("orc_T", "orc_Tstar", "orc_qr", "orc_qrs", "orc_pr", "orc_prs", "orc_curl", "orc_curlh"):
            getattr(L, n).argtypes = [vp, vp, vp]
        L.orc_qbq.argtypes = [vp, vp, vp, vp]
        L.orc_cg.argtypes = [vp, vp, vp, vp, i, d, C.POINTER(d)]
        L.orc_invar.argtypes = [vp, vp, vp, vp, i, d]
        L.orc_curl_coo.argtypes = [C.POINTER(orc_lattice), vp, vp, vp, vp]
        L.orc_tridiag_eig.argtypes = [i, vp, vp, vp, vp]
        L.orc_gvqrrq.argtypes = [vp, vp, vp, vp, vp, vp, d, i]
        L.orc_eigs.argtypes = [C.POINTER(orc_lattice), vp, vp, C.POINTER(orc_es), C.POINTER(orc_ls), vp, vp,
                               C.POINTER(orc_profile)]
        L.orc_rel_residual.argtypes = [vp, vp, vp, d]
        L.orc_rel_residual.restype = d
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def lattice_from(fl) -> orc_lattice:
    """orc_lattice from a fame_lattice (same fields, other names)."""
    o = orc_lattice()
    for i in range(3):
        o.N[i] = fl.grid_nums[i]
        o.h[i] = fl.mesh_len[i]
        o.t1[i], o.t2[i], o.t3[i], o.t4[i] = fl.t1[i], fl.t2[i], fl.t3[i], fl.t4[i]
    for i in range(9):
        o.a[i] = fl.lattice_vec_a[i]
    o.general = fl.general
    o.m1, o.m2, o.m3, o.m4 = fl.m1, fl.m2, fl.m3, fl.m4
    return o


def fft(x: np.ndarray, sign: int) -> np.ndarray:
    y = np.ascontiguousarray(x, np.complex128).copy()
    n = y.shape[-1]
    lib().orc_fft(n, sign, _p(y), y.size // n)
    return y


class Tables:
    def __init__(self, lat: orc_lattice, k):
        self.lat = lat
        self.k = np.ascontiguousarray(k, np.float64)
        self.h = lib().orc_tables_create(C.byref(lat), _p(self.k))
        self.N = lat.N[0] * lat.N[1] * lat.N[2]
        self.Nd = lib().orc_tables_nd(self.h)

    def __del__(self):
        try:
            lib().orc_tables_destroy(self.h)
        except Exception:
            pass

    def get(self, which: int) -> np.ndarray:
        n = self.Nd if which < 4 else 6 * self.Nd
        out = np.zeros(n, np.float64 if which == 3 else np.complex128)
        lib().orc_tables_get(self.h, which, _p(out))
        return out

    def _op(self, name, x, out_len):
        xx = np.ascontiguousarray(x, np.complex128)
        y = np.zeros(out_len, np.complex128)
        getattr(lib(), name)(self.h, _p(xx), _p(y))
        return y

    def qr(self, x): return self._op("orc_qr", x, 3 * self.N)
    def pr(self, x): return self._op("orc_pr", x, 3 * self.N)
    def qrs(self, v): return self._op("orc_qrs", v, 2 * self.Nd)
    def prs(self, v): return self._op("orc_prs", v, 2 * self.Nd)
    def T(self, u): return self._op("orc_T", u, self.N)
    def Tstar(self, v): return self._op("orc_Tstar", v, self.N)
    def curl(self, E): return self._op("orc_curl", E, 3 * self.N)
    def curlh(self, F): return self._op("orc_curlh", F, 3 * self.N)

    def qbq(self, invB, x):
        xx = np.ascontiguousarray(x, np.complex128)
        ib = np.ascontiguousarray(invB, np.float64)
        y = np.zeros(2 * self.Nd, np.complex128)
        lib().orc_qbq(self.h, _p(ib), _p(xx), _p(y))
        return y

    def cg(self, invB, rhs, maxit=1000, tol=1e-12):
        b = np.ascontiguousarray(rhs, np.complex128)
        ib = np.ascontiguousarray(invB, np.float64)
        x = np.zeros(2 * self.Nd, np.complex128)
        res = C.c_double()
        it = lib().orc_cg(self.h, _p(ib), _p(b), _p(x), maxit, tol, C.byref(res))
        return x, it, res.value

    def invar(self, invB, x, maxit=1000, tol=1e-12):
        xx = np.ascontiguousarray(x, np.complex128)
        ib = np.ascontiguousarray(invB, np.float64)
        y = np.zeros(2 * self.Nd, np.complex128)
        it = lib().orc_invar(self.h, _p(ib), _p(xx), _p(y), maxit, tol)
        return y, it

    def rel_residual(self, eps, E, omega):
        return lib().orc_rel_residual(self.h, _p(np.ascontiguousarray(eps, np.float64)),
                                      _p(np.ascontiguousarray(E, np.complex128)), float(omega))


def curl_coo(lat: orc_lattice, k):
    N = lat.N[0] * lat.N[1] * lat.N[2]
    r = np.zeros(12 * N, np.int32)
    c = np.zeros(12 * N, np.int32)
    v = np.zeros(12 * N, np.complex128)
    kk = np.ascontiguousarray(k, np.float64)
    lib().orc_curl_coo(C.byref(lat), _p(kk), _p(r), _p(c), _p(v))
    return r, c, v


def tridiag_eig(d, e):
    n = len(d)
    w = np.zeros(n)
    z = np.zeros((n, n))
    dd = np.ascontiguousarray(d, np.float64)
    ee = np.ascontiguousarray(e, np.float64)
    st = lib().orc_tridiag_eig(n, _p(dd), _p(ee), _p(w), _p(z))
    return w, z.T.copy(), st   # columns are eigenvectors


def eigs(lat: orc_lattice, eps, k, nwant, nstep, tol=1e-12, maxit=300, ls_tol=1e-12, ls_maxit=1000, fields=False):
    N = lat.N[0] * lat.N[1] * lat.N[2]
    e = np.ascontiguousarray(eps, np.float64)
    kk = np.ascontiguousarray(k, np.float64)
    freq = np.zeros(nwant)
    ev = np.zeros(3 * N * nwant, np.complex128) if fields else None
    prof = orc_profile()
    es = orc_es(nwant, nstep, maxit, tol)
    ls = orc_ls(ls_maxit, ls_tol)
    lib().orc_eigs(C.byref(lat), _p(e), _p(kk), C.byref(es), C.byref(ls), _p(freq),
                   _p(ev) if ev is not None else None, C.byref(prof))
    out = dict(ls_iter=prof.ls_iter, es_iter=prof.es_iter, restarts=prof.restarts, nd=prof.nd)
    return freq, (ev.reshape(nwant, 3 * N) if ev is not None else None), out
