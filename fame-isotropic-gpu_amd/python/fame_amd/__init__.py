"""fame_amd -- Python bindings of the MI355X band solver (ctypes over the C-ABI).

The compute path is libfame_hip.so (hand-written gfx950 kernels behind
include/fame_hip.h).  PyTorch is used only to own device memory and to pass
device pointers; nothing here computes on the CPU on behalf of the solver, and
every entry point raises if the native library is missing.

Reference interfaces mirrored (file:line under the reference tree):
  Solver.eigs        FAME_Main/FAME_Fast_Algorithms_Isotropic.cu:20 + Lanczos_Isotropic.cu:7
  Solver.apply       FAME_Main/FAME_Matrix_Vector_Production_{Qr,Qrs,Pr,Prs,Isotropic_QBQ,Isotropic_invAr}.cu
  Solver.cg          FAME_Main/CG.cu:5
  build_lattice      FAME_Preprocessing/FAME_Parameter_Generator.cpp:21
  material_eps       FAME_Preprocessing/FAME_Material_Locate_Index.cpp:101 + FAME_Main/FAME_Matrix_B_Isotropic.cu:4
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(os.path.dirname(_HERE))          # fame-isotropic-gpu_amd/
LIB_PATH = os.path.join(_PKG, "libfame_hip.so")

FAME_OK, FAME_ERR_ARG, FAME_ERR_HIP, FAME_ERR_NOMEM, FAME_ERR_STATE, FAME_ERR_UNSUPPORTED = range(6)
PREC_F64, PREC_F32_REFINE = 0, 1
OP_QR, OP_QRS, OP_PR, OP_PRS, OP_QBQ, OP_INVAR = range(6)

SIMPLE_TYPES = ("simple_cubic", "primitive_orthorhombic", "primitive_tetragonal")


class FameError(RuntimeError):
    pass


class fame_lattice(C.Structure):
    _fields_ = [("grid_nums", C.c_int * 3), ("general", C.c_int), ("mesh_len", C.c_double * 3),
                ("lattice_vec_a", C.c_double * 9), ("m1", C.c_int), ("m2", C.c_int), ("m3", C.c_int), ("m4", C.c_int),
                ("t1", C.c_double * 3), ("t2", C.c_double * 3), ("t3", C.c_double * 3), ("t4", C.c_double * 3)]


class fame_es(C.Structure):
    _fields_ = [("nwant", C.c_int), ("nstep", C.c_int), ("maxit", C.c_int), ("tol", C.c_double)]


class fame_ls(C.Structure):
    _fields_ = [("maxit", C.c_int), ("tol", C.c_double)]


class fame_profile(C.Structure):
    _fields_ = [("ls_iter", C.c_int), ("es_iter", C.c_int), ("restarts", C.c_int), ("nd", C.c_int),
                ("converged", C.c_int), ("qbq64", C.c_int), ("qbq32", C.c_int), ("pad", C.c_int),
                ("es_time", C.c_double), ("ls_time", C.c_double)]


class fame_lattice_in(C.Structure):
    _fields_ = [("lattice_type", C.c_char * 50), ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double), ("grid_nums", C.c_int * 3)]


class fame_lattice_out(C.Structure):
    _fields_ = [("lattice_type", C.c_char * 50), ("general", C.c_int), ("grid_nums", C.c_int * 3),
                ("edge_len", C.c_double * 3), ("mesh_len", C.c_double * 3),
                ("a", C.c_double), ("b", C.c_double), ("c", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
                ("lattice_vec_a", C.c_double * 9), ("lattice_vec_a_orig", C.c_double * 9), ("Omega", C.c_double * 9),
                ("Permutation", C.c_int * 3), ("length_a", C.c_double * 3), ("theta", C.c_double * 3),
                ("m1", C.c_int), ("m2", C.c_int), ("m3", C.c_int), ("m4", C.c_int),
                ("t1", C.c_double * 3), ("t2", C.c_double * 3), ("t3", C.c_double * 3), ("t4", C.c_double * 3),
                ("flag", C.c_char * 6), ("recip_b", C.c_double * 9)]


class fame_material_in(C.Structure):
    _fields_ = [("material_num", C.c_int), ("sphere_num", C.POINTER(C.c_int)),
                ("sphere_centers", C.POINTER(C.c_double)), ("sphere_radius", C.POINTER(C.c_double)),
                ("cylinder_num", C.POINTER(C.c_int)), ("cylinder_top_centers", C.POINTER(C.c_double)),
                ("cylinder_bot_centers", C.POINTER(C.c_double)), ("cylinder_radius", C.POINTER(C.c_double))]


EXPORTS = ["fame_create", "fame_destroy", "fame_set_eps", "fame_set_wave_vectors", "fame_apply", "fame_cg",
           "fame_eigs", "fame_residuals", "fame_eig_grad", "fame_set_stream", "fame_query", "fame_last_error",
           "fame_lattice_build", "fame_material_locate", "fame_material_eps", "fame_bz_vertex",
           "fame_wave_vec_rotate", "fame_bz_path_count", "fame_bz_path", "fame_bz_default_path"]

_lib = None


def lib():
    """Load libfame_hip.so once.  Raises if it is not built: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FameError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
    # PyTorch ships its own HIP runtime; load it first so that this library and
    # torch share one runtime (the library's libamdhip64 dependency then
    # resolves to the copy already in the process)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.fame_create.argtypes = [C.POINTER(fame_lattice), i, i, C.POINTER(vp)]
    L.fame_destroy.argtypes = [vp]
    L.fame_set_eps.argtypes = [vp, vp]
    L.fame_set_wave_vectors.argtypes = [vp, i, vp, vp]
    L.fame_apply.argtypes = [vp, i, i, vp, vp]
    L.fame_cg.argtypes = [vp, i, vp, vp, i, d, C.POINTER(i)]
    L.fame_eigs.argtypes = [vp, i, C.POINTER(fame_es), C.POINTER(fame_ls), vp, vp, vp, vp]
    L.fame_residuals.argtypes = [vp, i, vp, vp, vp]
    L.fame_residuals_noalloc.argtypes = [vp, i, vp, vp, vp]
    L.fame_check_eigendecomp.argtypes = [vp, i, C.c_ulonglong, vp]
    L.fame_curl_coo.argtypes = [C.POINTER(fame_lattice), vp, vp, vp, vp]
    L.fame_eig_grad.argtypes = [vp, i, vp, vp, vp, vp]
    L.fame_time_qbq_passes.argtypes = [vp, i, i, vp]
    L.fame_set_refinement.argtypes = [vp, i, d, i]
    L.fame_set_groups.argtypes = [vp, i]
    L.fame_time_cg_passes.argtypes = [vp, i, i, vp]
    L.fame_time_qbq_passes.restype = C.c_int
    L.fame_set_stream.argtypes = [vp, vp]
    L.fame_query.argtypes = [vp, vp, vp, vp]
    L.fame_last_error.restype = C.c_char_p
    L.fame_lattice_build.argtypes = [C.POINTER(fame_lattice_in), C.POINTER(fame_lattice_out)]
    L.fame_material_locate.argtypes = [C.POINTER(fame_lattice_out), C.POINTER(fame_material_in), vp]
    L.fame_material_eps.argtypes = [i, i, vp, vp, i, d, vp]
    L.fame_bz_vertex.argtypes = [C.POINTER(fame_lattice_out), C.c_char, vp]
    L.fame_wave_vec_rotate.argtypes = [C.POINTER(fame_lattice_out), vp, vp]
    L.fame_bz_path_count.argtypes = [C.c_char_p, i]
    L.fame_bz_path.argtypes = [C.POINTER(fame_lattice_out), C.c_char_p, i, vp]
    L.fame_bz_default_path.argtypes = [C.POINTER(fame_lattice_out)]
    L.fame_bz_default_path.restype = C.c_char_p
    for name in ("fame_create", "fame_destroy", "fame_set_eps", "fame_set_wave_vectors", "fame_apply", "fame_cg",
                 "fame_eigs", "fame_residuals", "fame_eig_grad", "fame_set_stream", "fame_query"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(st: int, what: str):
    if st != FAME_OK:
        msg = lib().fame_last_error()
        raise FameError(f"{what} failed with status {st}: {msg.decode() if msg else ''}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------- geometry
@dataclass
class Material:
    """Inclusions in fractional coordinates of the original lattice vectors
    (the reference's FAME_Material_Data format, FAME_Set_User_Define.cpp:4-120)."""
    sphere_centers: list = field(default_factory=list)      # per material: list of (x,y,z)
    sphere_radius: list = field(default_factory=list)       # per material: list of r
    cylinder_top: list = field(default_factory=list)
    cylinder_bot: list = field(default_factory=list)
    cylinder_radius: list = field(default_factory=list)

    @property
    def material_num(self):
        return max(len(self.sphere_centers), len(self.cylinder_top))


def build_lattice(lattice_type: str, nx: int, a=1.0, b=0.0, c=0.0, alpha=0.0, beta=0.0, gamma=0.0):
    li = fame_lattice_in()
    li.lattice_type = lattice_type.encode()
    li.a, li.b, li.c, li.alpha, li.beta, li.gamma = a, b, c, alpha, beta, gamma
    li.grid_nums[0] = nx
    lo = fame_lattice_out()
    if lib().fame_lattice_build(C.byref(li), C.byref(lo)) != 0:
        raise FameError(f"unknown lattice type {lattice_type}")
    return lo


def to_fame_lattice(lo: fame_lattice_out) -> fame_lattice:
    L = fame_lattice()
    for i in range(3):
        L.grid_nums[i] = lo.grid_nums[i]
        L.mesh_len[i] = lo.mesh_len[i]
        L.t1[i], L.t2[i], L.t3[i], L.t4[i] = lo.t1[i], lo.t2[i], lo.t3[i], lo.t4[i]
    for i in range(9):
        L.lattice_vec_a[i] = lo.lattice_vec_a[i]
    L.general = lo.general
    L.m1, L.m2, L.m3, L.m4 = lo.m1, lo.m2, lo.m3, lo.m4
    return L


def material_index(lo: fame_lattice_out, mat: Material) -> np.ndarray:
    nm = mat.material_num
    sn = np.array([len(mat.sphere_centers[m]) if m < len(mat.sphere_centers) else 0 for m in range(nm)], np.int32)
    cn = np.array([len(mat.cylinder_top[m]) if m < len(mat.cylinder_top) else 0 for m in range(nm)], np.int32)
    sc = np.array([p for m in range(nm) if m < len(mat.sphere_centers) for p in mat.sphere_centers[m]], np.float64).reshape(-1)
    sr = np.array([r for m in range(nm) if m < len(mat.sphere_radius) for r in mat.sphere_radius[m]], np.float64)
    ct = np.array([p for m in range(nm) if m < len(mat.cylinder_top) for p in mat.cylinder_top[m]], np.float64).reshape(-1)
    cb = np.array([p for m in range(nm) if m < len(mat.cylinder_bot) for p in mat.cylinder_bot[m]], np.float64).reshape(-1)
    cr = np.array([r for m in range(nm) if m < len(mat.cylinder_radius) for r in mat.cylinder_radius[m]], np.float64)
    keep = [sn, cn, sc, sr, ct, cb, cr]
    for k, arr in enumerate(keep):
        if arr.size == 0:
            keep[k] = np.zeros(1, arr.dtype)
    sn, cn, sc, sr, ct, cb, cr = keep
    mi = fame_material_in()
    mi.material_num = nm
    mi.sphere_num = sn.ctypes.data_as(C.POINTER(C.c_int))
    mi.sphere_centers = sc.ctypes.data_as(C.POINTER(C.c_double))
    mi.sphere_radius = sr.ctypes.data_as(C.POINTER(C.c_double))
    mi.cylinder_num = cn.ctypes.data_as(C.POINTER(C.c_int))
    mi.cylinder_top_centers = ct.ctypes.data_as(C.POINTER(C.c_double))
    mi.cylinder_bot_centers = cb.ctypes.data_as(C.POINTER(C.c_double))
    mi.cylinder_radius = cr.ctypes.data_as(C.POINTER(C.c_double))
    N = lo.grid_nums[0] * lo.grid_nums[1] * lo.grid_nums[2]
    idx = np.empty(3 * N, np.int32)
    lib().fame_material_locate(C.byref(lo), C.byref(mi), _ptr(idx))
    return idx


def material_eps(lo: fame_lattice_out, mat: Material, eps_in: Sequence[float], eps_out: float) -> np.ndarray:
    idx = material_index(lo, mat)
    N = idx.size // 3
    e_in = np.asarray(eps_in, np.float64)
    eps = np.empty(3 * N, np.float64)
    st = lib().fame_material_eps(N, mat.material_num, _ptr(idx), _ptr(e_in), e_in.size, float(eps_out), _ptr(eps))
    if st != 0:
        raise FameError("fewer permittivity values than materials")
    return eps


def bz_vertex(lo: fame_lattice_out, name: str) -> np.ndarray:
    v = np.zeros(3)
    if lib().fame_bz_vertex(C.byref(lo), name.encode(), _ptr(v)) != 0:
        raise FameError(f"unknown Brillouin-zone vertex {name!r} for {lo.lattice_type.decode()}")
    return v


def rotate_wave_vector(lo: fame_lattice_out, v) -> np.ndarray:
    vin = np.ascontiguousarray(v, np.float64)
    out = np.zeros(3)
    lib().fame_wave_vec_rotate(C.byref(lo), _ptr(vin), _ptr(out))
    return out


def bz_path(lo: fame_lattice_out, path: Optional[str] = None, part_num: int = 10) -> np.ndarray:
    if path is None:
        path = lib().fame_bz_default_path(C.byref(lo)).decode()
    n = lib().fame_bz_path_count(path.encode(), part_num)
    out = np.zeros((n, 3))
    got = lib().fame_bz_path(C.byref(lo), path.encode(), part_num, _ptr(out))
    if got < 0:
        raise FameError(f"bad path {path!r}")
    return out[:got]


def path_by_arclength(lo: fame_lattice_out, vertices: str, npts: int) -> np.ndarray:
    """npts wave vectors on the piecewise-linear vertex path, every vertex included
    and the remaining points shared among the segments in proportion to their
    length (largest remainder), evenly spaced inside each segment; computational
    frame.  Used for explicit k lists whose length the reference's part_num rule
    (4(p-1)+1 points for 5 vertices) cannot give."""
    V = np.array([bz_vertex(lo, ch) for ch in vertices])
    nseg = len(V) - 1
    if npts < nseg + 1:
        raise FameError("need at least one point per vertex")
    seg = np.linalg.norm(np.diff(V, axis=0), axis=1)
    inner = npts - (nseg + 1)
    share = inner * seg / seg.sum()
    cnt = np.floor(share).astype(int)
    for j in np.argsort(-(share - cnt))[: inner - cnt.sum()]:
        cnt[j] += 1
    pts = [V[0]]
    for j in range(nseg):
        for q in range(1, cnt[j] + 2):
            pts.append(V[j] + (q / (cnt[j] + 1)) * (V[j + 1] - V[j]))
    return np.array([rotate_wave_vector(lo, p) for p in pts])


# --------------------------------------------------------------------------- solver
def _torch():
    import torch
    return torch


class Solver:
    """One structure (lattice + permittivity), up to `max_batch` wave vectors."""

    def __init__(self, lattice: fame_lattice, max_batch: int = 1, device: int = 0):
        self.L = lattice
        self.max_batch = max_batch
        self.device = device
        self._ctx = C.c_void_p()
        _check(lib().fame_create(C.byref(lattice), max_batch, device, C.byref(self._ctx)), "fame_create")
        self.N = lattice.grid_nums[0] * lattice.grid_nums[1] * lattice.grid_nums[2]
        self.nd = []

    def close(self):
        if self._ctx:
            lib().fame_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_eps(self, eps3N: np.ndarray):
        e = np.ascontiguousarray(eps3N, np.float64)
        if e.size != 3 * self.N:
            raise FameError("eps must have 3N entries")
        _check(lib().fame_set_eps(self._ctx, _ptr(e)), "fame_set_eps")

    def set_wave_vectors(self, k):
        kk = np.ascontiguousarray(np.atleast_2d(k), np.float64)
        nd = np.zeros(kk.shape[0], np.int32)
        _check(lib().fame_set_wave_vectors(self._ctx, kk.shape[0], _ptr(kk), _ptr(nd)), "fame_set_wave_vectors")
        self.nd = [int(v) for v in nd]
        self.k = kk.copy()
        return self.nd

    def fft_lengths(self):
        n = C.c_int(); nk = C.c_int(); l3 = (C.c_int * 3)()
        _check(lib().fame_query(self._ctx, C.byref(n), C.byref(nk), l3), "fame_query")
        return list(l3)

    def apply(self, op: int, ik: int, x):
        """x, result: torch complex128 tensors on cuda (reference layouts)."""
        torch = _torch()
        nd = self.nd[ik]
        out_len = 3 * self.N if op in (OP_QR, OP_PR) else 2 * nd
        x = x.contiguous()
        y = torch.empty(out_len, dtype=torch.complex128, device=x.device)
        _check(lib().fame_apply(self._ctx, op, ik, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())), "fame_apply")
        return y

    def cg(self, ik: int, rhs, maxit=1000, tol=1e-12):
        torch = _torch()
        x = torch.empty_like(rhs)
        it = C.c_int()
        _check(lib().fame_cg(self._ctx, ik, C.c_void_p(rhs.contiguous().data_ptr()), C.c_void_p(x.data_ptr()),
                             maxit, tol, C.byref(it)), "fame_cg")
        return x, it.value

    def eigs(self, nwant, nstep, tol=1e-12, maxit=300, ls_tol=1e-12, ls_maxit=1000, precision=PREC_F64,
             want_reduced=False, want_fields=False):
        torch = _torch()
        nk = len(self.nd)
        es = fame_es(nwant, nstep, maxit, tol)
        ls = fame_ls(ls_maxit, ls_tol)
        freq = np.zeros((nk, nwant))
        prof = (fame_profile * nk)()
        dev = torch.device("cuda", self.device)
        evr = torch.empty((nk, nwant, 2 * self.N), dtype=torch.complex128, device=dev) if want_reduced else None
        E = torch.empty((nk, nwant, 3 * self.N), dtype=torch.complex128, device=dev) if want_fields else None
        _check(lib().fame_eigs(self._ctx, precision, C.byref(es), C.byref(ls), _ptr(freq),
                               C.c_void_p(evr.data_ptr()) if evr is not None else None,
                               C.c_void_p(E.data_ptr()) if E is not None else None, prof), "fame_eigs")
        profs = [dict(ls_iter=p.ls_iter, es_iter=p.es_iter, restarts=p.restarts, nd=p.nd, converged=p.converged,
                      qbq64=p.qbq64, qbq32=p.qbq32, es_time=p.es_time, ls_time=p.ls_time) for p in prof]
        return freq, evr, E, profs

    def set_refinement(self, sweeps: int, tol: float = 1e-14, maxit: int = 2000):
        """Inverse-iteration sweeps that eigs applies to the converged Ritz pairs
        (0: the reference's algorithm); see fame_set_refinement."""
        _check(lib().fame_set_refinement(self._ctx, int(sweeps), float(tol), int(maxit)), "fame_set_refinement")

    def set_groups(self, groups: int):
        """How eigs runs the batch: 1 one batch, g >= 2 that many concurrent groups,
        0 (default) two groups from two wave vectors on; see fame_set_groups."""
        _check(lib().fame_set_groups(self._ctx, int(groups)), "fame_set_groups")

    def restore(self, evr):
        """E fields (nk, nwant, 3N) recomputed from reduced Ritz vectors (nk, nwant, 2N)."""
        torch = _torch()
        nk, nwant, _ = evr.shape
        E = torch.empty((nk, nwant, 3 * self.N), dtype=torch.complex128, device=evr.device)
        _check(lib().fame_restore(self._ctx, nwant, C.c_void_p(evr.contiguous().data_ptr()), C.c_void_p(E.data_ptr())),
               "fame_restore")
        return E

    def residuals(self, freq, E):
        nk, nwant = freq.shape
        res = np.zeros((nk, nwant))
        f = np.ascontiguousarray(freq, np.float64)
        _check(lib().fame_residuals(self._ctx, nwant, _ptr(f), C.c_void_p(E.data_ptr()), _ptr(res)), "fame_residuals")
        return res

    def time_cg_passes(self, prec_bytes: int, reps: int = 10):
        """Average microseconds of the kernels of one fused CG iteration (all wave
        vectors active): inner expand, mid inverse, outer eps, alpha, mid forward,
        inner contract, beta, whole iteration."""
        us = np.zeros(8)
        _check(lib().fame_time_cg_passes(self._ctx, prec_bytes, reps, _ptr(us)), "fame_time_cg_passes")
        return us

    def check_eigendecomp(self, ik: int, seed: int = 1):
        """[||C_i x_i - T Lambda_i T* x_i|| for i = x, y, z, ||C x - P Sigma Q^H x||]
        for a seeded random field x (the reference's Check_Eigendecomp)."""
        err = np.zeros(4)
        _check(lib().fame_check_eigendecomp(self._ctx, ik, seed, _ptr(err)), "fame_check_eigendecomp")
        return err

    def residuals_noalloc(self, freq, E):
        """As residuals, through fame_residuals_noalloc (no device allocation;
        the fields are staged in the library's static device scratch)."""
        nk, nwant = freq.shape
        res = np.zeros((nk, nwant))
        f = np.ascontiguousarray(freq, np.float64)
        _check(lib().fame_residuals_noalloc(self._ctx, nwant, _ptr(f), C.c_void_p(E.data_ptr()), _ptr(res)),
               "fame_residuals_noalloc")
        return res

    def autotune(self, prec_bytes: int = 8):
        """Tune the pass kernels on the current batch; returns {pass: (engine, lines)}."""
        ch = np.zeros(6, np.int32)
        _check(lib().fame_autotune(self._ctx, prec_bytes, _ptr(ch)), "fame_autotune")
        return {name: (int(ch[2 * i]), int(ch[2 * i + 1])) for i, name in enumerate(("inner", "mid", "outer"))}

    def set_tiles(self, prec_bytes: int, tiles):
        """Fix the pass tiles: {pass: (engine, lines)} as returned by autotune."""
        ch = np.array([v for name in ("inner", "mid", "outer") for v in tiles[name]], np.int32)
        _check(lib().fame_set_tiles(self._ctx, prec_bytes, _ptr(ch)), "fame_set_tiles")

    def time_qbq_passes(self, prec_bytes: int = 8, reps: int = 20):
        """hipEvent-timed average microseconds of the 5 QBQ passes and one QBQ (all wave vectors)."""
        us = np.zeros(6)
        _check(lib().fame_time_qbq_passes(self._ctx, prec_bytes, reps, _ptr(us)), "fame_time_qbq_passes")
        return us

    def eig_vjp(self, freq, E, weights):
        """sum_{k,j} w_kj d lambda_kj / d eps  (3N, reference layout) by Hellmann-Feynman."""
        torch = _torch()
        nk, nwant = freq.shape
        f = np.ascontiguousarray(freq, np.float64)
        w = np.ascontiguousarray(weights, np.float64)
        g = torch.zeros(3 * self.N, dtype=torch.float64, device=E.device)
        _check(lib().fame_eig_grad(self._ctx, nwant, _ptr(f), C.c_void_p(E.data_ptr()), _ptr(w),
                                   C.c_void_p(g.data_ptr())), "fame_eig_grad")
        return g

OP_INVAR_MIXED, OP_QBQ32 = 6, 7


def curl_coo(lat, k):
    """The sparse curl of FAME_Matrix_Curl for wave vector k: (rows, cols, complex vals), 12N entries."""
    N = int(lat.grid_nums[0]) * int(lat.grid_nums[1]) * int(lat.grid_nums[2])
    r = np.zeros(12 * N, np.int32)
    c = np.zeros(12 * N, np.int32)
    v = np.zeros(12 * N, np.complex128)
    kk = np.ascontiguousarray(k, np.float64)
    _check(lib().fame_curl_coo(C.byref(lat), _ptr(kk), _ptr(r), _ptr(c), _ptr(v)), "fame_curl_coo")
    return r, c, v


def eigs_many(solvers, nwant, nstep, tol=1e-12, maxit=300, ls_tol=1e-12, ls_maxit=1000, precision=PREC_F64,
              want_fields=False):
    """Solve several structures (Solvers of any grid and batch) in one call;
    nwant / nstep / precision may be per-solver lists.  Returns [(freq, profs)],
    or [(freq, profs, E)] with the restored fields (nk, nwant, 3N) on the GPU
    when want_fields."""
    n = len(solvers)
    per = lambda v: list(v) if isinstance(v, (list, tuple)) else [v] * n
    nw, ns, pr = per(nwant), per(nstep), per(precision)
    ctxs = (C.c_void_p * n)(*[s._ctx.value for s in solvers])
    esa = (fame_es * n)(*[fame_es(nw[i], ns[i], maxit, tol) for i in range(n)])
    lsa = (fame_ls * n)(*[fame_ls(ls_maxit, ls_tol) for _ in range(n)])
    prec = (C.c_int * n)(*pr)
    freqs = [np.zeros((len(s.nd), nw[i])) for i, s in enumerate(solvers)]
    profs = [(fame_profile * len(s.nd))() for s in solvers]
    fp = (C.c_void_p * n)(*[f.ctypes.data for f in freqs])
    pp = (C.c_void_p * n)(*[C.addressof(p) for p in profs])
    Es, ep = None, None
    if want_fields:
        torch = _torch()
        Es = [torch.empty((len(s.nd), nw[i], 3 * s.N), dtype=torch.complex128, device=torch.device("cuda", s.device))
              for i, s in enumerate(solvers)]
        ep = (C.c_void_p * n)(*[E.data_ptr() for E in Es])
    _check(lib().fame_eigs_many(n, ctxs, prec, esa, lsa, fp, None, ep, pp), "fame_eigs_many")
    out = []
    for i, (f, ps) in enumerate(zip(freqs, profs)):
        pd = [dict(ls_iter=p.ls_iter, es_iter=p.es_iter, restarts=p.restarts, nd=p.nd, converged=p.converged,
                   qbq64=p.qbq64, qbq32=p.qbq32, es_time=p.es_time, ls_time=p.ls_time) for p in ps]
        out.append((f, pd, Es[i]) if want_fields else (f, pd))
    return out


# --------------------------------------------------------------------------- adjoint
def _make_eig_function():
    torch = _torch()

    class BandEigenvalues(torch.autograd.Function):
        """lambda = omega^2 of the `nwant` lowest bands at every wave vector of
        `solver`, as a differentiable function of the permittivity on the Yee
        points (3N, reference layout).  The backward pass is the Hellmann-Feynman
        adjoint d lambda / d eps_e = -lambda |E_e|^2 / (E^H B E) evaluated on the
        GPU through fame_eig_grad; inside a degenerate cluster it returns the
        derivative of the cluster mean (the only basis-independent quantity)."""

        @staticmethod
        def forward(ctx, eps, solver, nwant, nstep, tol, precision):
            e = eps.detach().to("cpu", torch.float64).numpy().copy()
            solver.set_eps(e)
            freq, evr, _, prof = solver.eigs(nwant, nstep, tol=tol, precision=precision, want_reduced=True)
            if not all(p["converged"] for p in prof):
                raise FameError("eigensolve did not converge")
            ctx.solver, ctx.freq, ctx.evr = solver, freq, evr
            ctx.eps_np, ctx.k = e, solver.k.copy()
            ctx.eps_device = eps.device
            return torch.as_tensor(freq ** 2, dtype=torch.float64, device=eps.device)

        @staticmethod
        def backward(ctx, grad_lam):
            S = ctx.solver
            # the fields are recomputed from the reduced eigenvectors (not stored)
            S.set_eps(ctx.eps_np)
            if not np.array_equal(S.k, ctx.k):
                S.set_wave_vectors(ctx.k)
            E = S.restore(ctx.evr)
            w = grad_lam.detach().to("cpu", torch.float64).numpy()
            g = S.eig_vjp(ctx.freq, E, w)
            del E
            return g.to(ctx.eps_device), None, None, None, None, None

    return BandEigenvalues


_EIG_FN = None


def band_eigenvalues(solver: "Solver", eps, nwant: int, nstep: int, tol: float = 1e-12, precision: int = PREC_F64):
    """Differentiable lambda[k, band] (torch float64) of eps (torch float64, 3N)."""
    global _EIG_FN
    if _EIG_FN is None:
        _EIG_FN = _make_eig_function()
    return _EIG_FN.apply(eps, solver, nwant, nstep, tol, precision)


def cell_to_yee(rho_eps):
    """Permittivity on the Yee edges from a periodic per-cell field (torch,
    shape (Nz, Ny, Nx), x fastest like the reference's real-space layout): the
    edge along axis d between cells r and r + e_d takes the mean of the two.
    Returns 3N entries [x-edges; y-edges; z-edges]; differentiable."""
    torch = _torch()
    ex = 0.5 * (rho_eps + torch.roll(rho_eps, -1, dims=2))
    ey = 0.5 * (rho_eps + torch.roll(rho_eps, -1, dims=1))
    ez = 0.5 * (rho_eps + torch.roll(rho_eps, -1, dims=0))
    return torch.cat([ex.reshape(-1), ey.reshape(-1), ez.reshape(-1)])
