"""GPU parity: the HIP path through the C-ABI against the FP64 oracle.

Operators (Qr, Qrs, Pr, Prs, QBQ, invAr) on random vectors, CG, and the full
restarted Lanczos eigensolve (FP64 and FP32+FP64-refinement), for Simple (SC)
and General (FCC, BCC, hexagonal) lattices, at a generic k and at Gamma.
Bars: operators 1e-11 relative (FP64 FFT rounding), eigenvalues 1e-8 relative
(the eigenvalue tolerance of the benchmark workload), residual
||A E - lambda B E||/||E|| small.
Both line-transform engines are covered: Stockham/Bluestein passes and the
direct DFT on the matrix cores (FAME_FFT selects the engine at context
creation; the default picks the matrix cores for axis lengths 16..127).
"""
import os

import numpy as np
import scipy.sparse as sp
import pytest

import fame_amd as F
import oracle as O
from cases import KPTS, STRUCTS, equal_mesh_general, make

pytestmark = pytest.mark.gpu

O.set_threads(8)


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(params=["stockham", "mfma"])
def engine(request):
    old = os.environ.get("FAME_FFT")
    os.environ["FAME_FFT"] = request.param
    yield request.param
    if old is None:
        del os.environ["FAME_FFT"]
    else:
        os.environ["FAME_FFT"] = old


@pytest.mark.parametrize("kname", list(KPTS))
@pytest.mark.parametrize("sname", list(STRUCTS))
def test_operators(gpu, engine, sname, kname):
    import torch
    lo, fl, ol, eps = make(sname)
    k = KPTS[kname]
    S = F.Solver(fl, max_batch=2)
    S.set_eps(eps)
    nd = S.set_wave_vectors(np.stack([np.array([0.4, 0.1, 0.2]), k]))
    t = O.Tables(ol, k)
    assert nd[1] == t.Nd
    N = t.N
    ik = 1
    x = _rand(2 * t.Nd, 1)
    v = _rand(3 * N, 2)
    xd = torch.tensor(x, device=gpu)
    vd = torch.tensor(v, device=gpu)
    assert _rel(S.apply(F.OP_QR, ik, xd).cpu().numpy(), t.qr(x)) < 1e-11
    assert _rel(S.apply(F.OP_PR, ik, xd).cpu().numpy(), t.pr(x)) < 1e-11
    assert _rel(S.apply(F.OP_QRS, ik, vd).cpu().numpy(), t.qrs(v)) < 1e-11
    assert _rel(S.apply(F.OP_PRS, ik, vd).cpu().numpy(), t.prs(v)) < 1e-11
    invB = 1.0 / eps
    assert _rel(S.apply(F.OP_QBQ, ik, xd).cpu().numpy(), t.qbq(invB, x)) < 1e-11
    y_gpu = S.apply(F.OP_INVAR, ik, xd).cpu().numpy()
    y_orc, _ = t.invar(invB, x)
    assert _rel(y_gpu, y_orc) < 1e-9
    S.close()


@pytest.mark.parametrize("sname", ["sc", "fcc"])
def test_cg(gpu, engine, sname):
    import torch
    lo, fl, ol, eps = make(sname)
    k = KPTS["k"]
    S = F.Solver(fl, max_batch=1)
    S.set_eps(eps)
    S.set_wave_vectors(k)
    t = O.Tables(ol, k)
    b = _rand(2 * t.Nd, 3) * 1e-2
    xg, itg = S.cg(0, torch.tensor(b, device=gpu), maxit=1000, tol=1e-12)
    xo, ito, _ = t.cg(1.0 / eps, b, 1000, 1e-12)
    assert abs(itg - ito) <= 1
    assert _rel(xg.cpu().numpy(), xo) < 1e-9
    S.close()


@pytest.mark.parametrize("prec", [F.PREC_F64, F.PREC_F32_REFINE])
@pytest.mark.parametrize("sname", list(STRUCTS))
def test_eigs_batched(gpu, engine, sname, prec):
    lo, fl, ol, eps = make(sname)
    ks = np.stack([KPTS["k"], KPTS["gamma"], np.array([0.05, 0.31, -0.12])])
    nwant, nstep = 6, 20
    S = F.Solver(fl, max_batch=len(ks))
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    freq, _, E, prof = S.eigs(nwant, nstep, precision=prec, want_fields=True)
    res = S.residuals(freq, E)
    for i, k in enumerate(ks):
        fo, _, po = O.eigs(ol, eps, k, nwant, nstep)
        lam_g, lam_o = freq[i] ** 2, fo ** 2
        nz = lam_o > 0
        assert np.array_equal(lam_g > 0, nz)
        assert np.max(np.abs(lam_g[nz] - lam_o[nz]) / lam_o[nz]) < 1e-8, (sname, i, lam_g, lam_o)
        assert prof[i]["converged"] == 1
        # The reference's stopping rule bounds the Ritz residual of A_r^{-1} (theta = 1/lambda)
        # by es.tol = 1e-12; in lambda space that bound is amplified by lambda*||A_r||
        # (||A_r|| ~ max|Lambda|^2 / eps_min), so the field residual is checked at 1e-6 lambda.
        assert np.all(res[i][nz] < 1e-6 * lam_g[nz]), res[i]
    S.close()


@pytest.mark.parametrize("sname,nx", [("fcc", 22), ("sc", 20)])
def test_operators_default_engine_medium(gpu, sname, nx):
    """Grids whose axes the default engine puts on the matrix cores (16..127),
    one with a prime axis (FCC 22 -> 22x19x18: 19 is prime)."""
    import torch
    lt, _, kw, mat, ein = __import__("cases").STRUCTS[sname]
    lo = F.build_lattice(lt, nx, **kw)
    fl = F.to_fame_lattice(lo)
    eps = F.material_eps(lo, mat, [ein], 1.0)
    ol = O.lattice_from(fl)
    k = KPTS["k"]
    old = os.environ.pop("FAME_FFT", None)
    try:
        S = F.Solver(fl, max_batch=1)
    finally:
        if old is not None:
            os.environ["FAME_FFT"] = old
    S.set_eps(eps)
    S.set_wave_vectors(k)
    t = O.Tables(ol, k)
    x = _rand(2 * t.Nd, 5)
    xd = torch.tensor(x, device=gpu)
    invB = 1.0 / eps
    assert _rel(S.apply(F.OP_QR, 0, xd).cpu().numpy(), t.qr(x)) < 1e-11
    assert _rel(S.apply(F.OP_QBQ, 0, xd).cpu().numpy(), t.qbq(invB, x)) < 1e-11
    # the FP32 operator against the FP64 oracle: single-precision rounding only
    assert _rel(S.apply(F.OP_QBQ32, 0, xd).cpu().numpy(), t.qbq(invB, x)) < 2e-6
    S.close()


@pytest.mark.parametrize("fft", ["stockham", "auto"])
def test_cg_and_eigs_prime_axis(gpu, fft):
    """FCC 20 -> 20x17x16.  With the Stockham engine the prime 17 takes a
    Bluestein plan, so the CG iteration keeps Ap and reads p in its last pass;
    with the default engine (17 on the matrix cores) the last pass updates r
    and the outer pass sums <p,Ap>.  Both against the oracle."""
    import torch
    lt, _, kw, mat, ein = __import__("cases").STRUCTS["fcc"]
    lo = F.build_lattice(lt, 20, **kw)
    fl = F.to_fame_lattice(lo)
    assert list(fl.grid_nums) == [20, 17, 16]
    eps = F.material_eps(lo, mat, [ein], 1.0)
    ol = O.lattice_from(fl)
    k = KPTS["k"]
    old = os.environ.get("FAME_FFT")
    os.environ["FAME_FFT"] = fft
    try:
        S = F.Solver(fl, max_batch=1)
    finally:
        if old is None:
            del os.environ["FAME_FFT"]
        else:
            os.environ["FAME_FFT"] = old
    S.set_eps(eps)
    S.set_wave_vectors(k)
    t = O.Tables(ol, k)
    b = _rand(2 * t.Nd, 7) * 1e-2
    xg, itg = S.cg(0, torch.tensor(b, device=gpu), maxit=1000, tol=1e-12)
    xo, ito, _ = t.cg(1.0 / eps, b, 1000, 1e-12)
    assert abs(itg - ito) <= 1
    assert _rel(xg.cpu().numpy(), xo) < 1e-9
    fo, _, _ = O.eigs(ol, eps, k, 6, 20)
    lam_o = fo ** 2
    for prec in (F.PREC_F64, F.PREC_F32_REFINE):
        freq, _, _, prof = S.eigs(6, 20, precision=prec)
        assert prof[0]["converged"] == 1
        assert np.max(np.abs(freq[0] ** 2 - lam_o) / lam_o) < 1e-8, (fft, prec, freq[0] ** 2, lam_o)
    S.close()


def test_graph_replay_matches_eager(gpu):
    """The CG iteration replayed from a captured hipGraph gives bit-identical
    results to eager launches: several solves in a row (each a fresh replay
    sequence with host synchronisation and eager work in between), and a full
    batched eigensolve in both precisions."""
    import torch
    lo, fl, ol, eps = make("fcc")
    ks = np.stack([KPTS["k"], KPTS["gamma"]])
    out = {}
    for graphs in ("1", "0"):
        # the autotuner is off so that both contexts use the same tiles
        saved = {k: os.environ.get(k) for k in ("FAME_GRAPH", "FAME_AUTOTUNE")}
        os.environ["FAME_GRAPH"] = graphs
        os.environ["FAME_AUTOTUNE"] = "0"
        try:
            S = F.Solver(fl, max_batch=2)
        finally:
            for k, v in saved.items():
                if v is None:
                    del os.environ[k]
                else:
                    os.environ[k] = v
        S.set_eps(eps)
        S.set_wave_vectors(ks)
        res = []
        for seed in range(3):
            b = torch.tensor(_rand(2 * S.nd[0], 10 + seed) * 1e-2, device=gpu)
            x, it = S.cg(0, b, maxit=500, tol=1e-12)
            res.append((x.cpu().numpy(), it))
            torch.cuda.synchronize()
            _ = (b * 2).sum().item()                      # eager work between replays
        f64 = S.eigs(6, 20, precision=F.PREC_F64)[0]
        f32 = S.eigs(6, 20, precision=F.PREC_F32_REFINE)[0]
        out[graphs] = (res, f64, f32)
        S.close()
    for (xg, ig), (xe, ie) in zip(out["1"][0], out["0"][0]):
        assert ig == ie and np.array_equal(xg, xe)
    assert np.array_equal(out["1"][1], out["0"][1])
    assert np.array_equal(out["1"][2], out["0"][2])


def test_batch_with_uneven_convergence(gpu):
    """Problems of one batch that converge at different restarts keep their own
    Lanczos bases: every band's field must satisfy the eigen-equation, and the
    batched frequencies must equal one-at-a-time solves."""
    lo, fl, ol, eps = make("fcc")
    ks = np.array([[0.21, -0.13, 0.07], [0.0, 0.0, 0.0], [0.4, 0.1, 0.2], [0.02, 0.01, -0.01],
                   [0.33, 0.33, 0.0], [-0.15, 0.27, 0.4]])
    nwant, nstep = 8, 14
    S = F.Solver(fl, max_batch=len(ks))
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    freq, _, E, prof = S.eigs(nwant, nstep, want_fields=True)
    restarts = [p["restarts"] for p in prof]
    assert len(set(restarts)) > 1, restarts          # the scenario this test is about
    res = S.residuals(freq, E)
    for i in range(len(ks)):
        nz = freq[i] > 0
        assert np.all(res[i][nz] < 1e-6 * freq[i][nz] ** 2), (i, restarts, res[i])
    one = F.Solver(fl, max_batch=1)
    one.set_eps(eps)
    for i in range(len(ks)):
        one.set_wave_vectors(ks[i:i + 1])
        f1 = one.eigs(nwant, nstep)[0][0]
        nz = f1 > 0
        assert np.max(np.abs(f1[nz] ** 2 - freq[i][nz] ** 2) / f1[nz] ** 2) < 1e-9
    one.close()
    S.close()


def test_qbq_counts_are_applications(gpu):
    """The profile's QBQ counts are operator applications, not launch slots:
    in a batch whose problems finish their CG solves at different iterations,
    the FP64 count equals the CG iterations (ls_iter minus one per solve) and
    the FP32 count of the refined solve equals its FP32 iterations, averaged
    over the batch's wave vectors as the profile reports them (over each
    group's when the batch is split)."""
    lo, fl, ol, eps = make("fcc")
    ks = np.array([[0.21, -0.13, 0.07], [0.0, 0.0, 0.0], [0.4, 0.1, 0.2], [0.33, 0.33, 0.0]])
    S = F.Solver(fl, max_batch=len(ks))
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    n = len(ks)
    for prec in (F.PREC_F64, F.PREC_F32_REFINE):
        S.set_groups(1)
        _, _, _, prof = S.eigs(8, 14, precision=prec, want_fields=False)
        cg_its = sum(p["ls_iter"] - p["es_iter"] for p in prof)
        solves = sum(p["es_iter"] for p in prof)
        if prec == F.PREC_F64:
            assert prof[0]["qbq64"] == cg_its // n and prof[0]["qbq32"] == 0, (prof[0], cg_its)
        else:
            assert prof[0]["qbq32"] == cg_its // n, (prof[0], cg_its)
            # one FP64 residual pass per sweep, at least one sweep per solve
            assert solves // n <= prof[0]["qbq64"] < prof[0]["qbq32"], (prof[0], solves)
        # split in two groups, each entry holds its group's mean: same total
        S.set_groups(2)
        _, _, _, prof2 = S.eigs(8, 14, precision=prec, want_fields=False)
        key = "qbq64" if prec == F.PREC_F64 else "qbq32"
        assert abs(sum(p[key] for p in prof2) - cg_its) < n, (prof2, cg_its)
    S.close()


def test_split_batch_is_bit_identical(gpu):
    """fame_eigs splits a batch of 16 or more wave vectors over concurrent child
    contexts (fame_set_groups): one batch, the automatic two groups and three
    groups give bit-identical eigenvalues, fields and iteration counts, in both
    precisions, and the children follow the parent's new wave vectors (another
    partition) and new permittivity."""
    import torch
    lo, fl, ol, eps = make("fcc")
    rng = np.random.default_rng(11)
    ks = np.vstack([np.zeros(3), rng.uniform(-0.4, 0.4, (17, 3))])
    S = F.Solver(fl, max_batch=len(ks))
    S.set_eps(eps)

    def run(groups, prec):
        S.set_groups(groups)
        f, _, E, prof = S.eigs(8, 14, precision=prec, want_fields=True)
        return f, E.cpu(), [p["ls_iter"] for p in prof]

    for kk, e in ((ks, eps), (ks[:16], eps), (ks[:16], 1.0 + 0.5 * eps)):
        S.set_eps(e)
        S.set_wave_vectors(kk)
        for prec in (F.PREC_F64, F.PREC_F32_REFINE):
            f1, E1, it1 = run(1, prec)
            for g in (0, 3):
                f2, E2, it2 = run(g, prec)
                assert np.array_equal(f1, f2), (len(kk), prec, g, np.max(np.abs(f1 - f2)))
                assert torch.equal(E1, E2), (len(kk), prec, g)
                assert it1 == it2, (len(kk), prec, g)
    S.close()


def _align_err(u, v):
    """|| u/|u| - phase * v/|v| || for the best unit phase."""
    u = u / np.linalg.norm(u)
    v = v / np.linalg.norm(v)
    ph = np.vdot(v, u)
    return float(np.linalg.norm(u - ph / abs(ph) * v))


@pytest.mark.parametrize("prec", [F.PREC_F64, F.PREC_F32_REFINE])
@pytest.mark.parametrize("sname", list(STRUCTS))
def test_fields_and_residuals_vs_oracle(gpu, sname, prec):
    """The restored E fields of fame_eigs against the oracle's (each
    non-degenerate band up to a global phase, and the multiply-by-the-norm
    scale of FAME_Fast_Algorithms_Isotropic.cu:119-121), the Gamma fields, and
    fame_residuals against the oracle's residual of the same field."""
    lo, fl, ol, eps = make(sname)
    ks = np.stack([KPTS["k"], KPTS["gamma"]])
    nwant, nstep = 6, 20
    S = F.Solver(fl, max_batch=len(ks))
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    freq, _, E, prof = S.eigs(nwant, nstep, precision=prec, want_fields=True)
    res = S.residuals(freq, E)
    Eh = E.cpu().numpy()
    N3 = 3 * S.N
    checked = 0
    for i, k in enumerate(ks):
        t = O.Tables(ol, k)
        fo, Eo, _ = O.eigs(ol, eps, k, nwant, nstep, fields=True)
        lam = fo ** 2
        for b in range(nwant):
            if lam[b] == 0:
                const = np.full(N3, 1.0 / np.sqrt(N3))
                assert np.abs(Eh[i, b] - const).max() == 0 and freq[i, b] == 0
                continue
            # the top band's partner above the wanted set is unknown, so its
            # field may be any vector of a degenerate eigenspace: not aligned
            gaps = [abs(lam[b] - lam[c]) for c in range(nwant) if c != b and lam[c] > 0]
            if b < nwant - 1 and gaps and min(gaps) > 1e-6 * lam[b]:
                assert _align_err(Eh[i, b], Eo[b]) < 1e-6, (sname, i, b)
                nr = np.linalg.norm(Eh[i, b]) / np.linalg.norm(Eo[b])
                assert abs(nr - 1) < 1e-6, (sname, i, b, nr)
                checked += 1
            # the GPU residual kernel against the oracle's residual of the same field
            ro = t.rel_residual(eps, Eh[i, b], freq[i, b])
            assert abs(res[i, b] - ro) <= 1e-3 * ro + 1e-11 * lam[b], (res[i, b], ro)
    # at Gamma the cubic structures' bands come in degenerate sets; the generic
    # wave vector has non-degenerate ones
    assert checked >= 2
    S.close()


@pytest.mark.parametrize("sname", list(STRUCTS))
def test_residual_kernel_on_arbitrary_fields(gpu, sname):
    """fame_residuals on fields that are not eigenvectors (residual of order
    ||C^H C||): the FFT curl-curl of the GPU against the oracle's sparse curl."""
    import torch
    lo, fl, ol, eps = make(sname)
    ks = np.stack([KPTS["k"], KPTS["gamma"]])
    S = F.Solver(fl, max_batch=2)
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    nw = 3
    E = np.stack([np.stack([_rand(3 * S.N, 50 + 7 * i + b) for b in range(nw)]) for i in range(2)])
    freq = np.array([[0.5, 1.7, 3.3], [0.9, 2.2, 4.1]])
    res = S.residuals(freq, torch.tensor(E, device=gpu))
    for i, k in enumerate(ks):
        t = O.Tables(ol, k)
        for b in range(nw):
            ro = t.rel_residual(eps, E[i, b], freq[i, b])
            assert abs(res[i, b] - ro) <= 1e-10 * ro, (sname, i, b, res[i, b], ro)
    S.close()


@pytest.mark.parametrize("prec", [F.PREC_F64, F.PREC_F32_REFINE])
def test_krylov_dimension_above_64(gpu, prec):
    """nstep = 3 nwant = 72 (the uneven-batch configuration's rule): the Gram
    kernel takes its 128-column form; eigenvalues against the oracle."""
    lo, fl, ol, eps = make("fcc")
    k = KPTS["k"]
    nwant, nstep = 24, 72
    S = F.Solver(fl, max_batch=1)
    S.set_eps(eps)
    S.set_wave_vectors(k)
    freq, _, E, prof = S.eigs(nwant, nstep, precision=prec, want_fields=True)
    fo, _, _ = O.eigs(ol, eps, k, nwant, nstep)
    assert prof[0]["converged"] == 1
    assert np.max(np.abs(freq[0] ** 2 - fo ** 2) / fo ** 2) < 1e-8
    res = S.residuals(freq, E)
    assert np.all(res[0] < 1e-6 * freq[0] ** 2)
    S.close()


def test_general_gauge_on_modes_parallel_to_111(gpu):
    """A cubic lattice solved through the General path (z-fastest Fourier
    layout, skew twiddles with m = 0) at k along (1,1,1): the modes with
    m_x = m_y = m_z have Lambda_x = Lambda_y = Lambda_z, where the reference's
    General normalisers are 0/0.  The spectrum must be finite and equal the
    Simple path's and the oracle's."""
    lo, fl, ol, eps = make("sc")
    k = np.array([0.11, 0.11, 0.11])
    out = {}
    for general in (0, 1):
        fl.general = general
        S = F.Solver(fl, max_batch=1)
        S.set_eps(eps)
        S.set_wave_vectors(k)
        out[general] = S.eigs(8, 24)[0][0]
        S.close()
    fl.general = 0
    assert np.all(np.isfinite(out[1]))
    assert np.max(np.abs(out[1] ** 2 - out[0] ** 2) / out[0] ** 2) < 1e-9
    og = O.lattice_from(fl)
    og.general = 1
    fo, _, _ = O.eigs(og, eps, k, 8, 24)
    assert np.max(np.abs(out[1] ** 2 - fo ** 2) / fo ** 2) < 1e-9


@pytest.mark.parametrize("shifts", [(0, 0, 0, 0), (0, 2, 0, 2)])
def test_operators_on_equal_mesh_general_grid(gpu, engine, shifts):
    """A General grid whose three mesh lengths are equal takes the guarded Pi
    basis in the inner kernels; at Gamma some of its modes have Lambda exactly
    parallel to (1, 1, 1).  Operators in FP64 (1e-11) and the FP32 QBQ (2e-6)
    against the oracle, at Gamma and at a generic wave vector."""
    import torch
    fl, ol, eps = equal_mesh_general(shifts, n=16)
    ks = np.stack([np.zeros(3), np.array([0.21, -0.13, 0.07])])
    S = F.Solver(fl, max_batch=2)
    S.set_eps(eps)
    nd = S.set_wave_vectors(ks)
    invB = 1.0 / eps
    for ik, k in enumerate(ks):
        t = O.Tables(ol, k)
        assert nd[ik] == t.Nd
        x = _rand(2 * t.Nd, 10 + ik)
        v = _rand(3 * t.N, 20 + ik)
        xd = torch.tensor(x, device=gpu)
        vd = torch.tensor(v, device=gpu)
        assert _rel(S.apply(F.OP_QR, ik, xd).cpu().numpy(), t.qr(x)) < 1e-11
        assert _rel(S.apply(F.OP_PR, ik, xd).cpu().numpy(), t.pr(x)) < 1e-11
        assert _rel(S.apply(F.OP_QRS, ik, vd).cpu().numpy(), t.qrs(v)) < 1e-11
        assert _rel(S.apply(F.OP_PRS, ik, vd).cpu().numpy(), t.prs(v)) < 1e-11
        assert _rel(S.apply(F.OP_QBQ, ik, xd).cpu().numpy(), t.qbq(invB, x)) < 1e-11
        assert _rel(S.apply(F.OP_QBQ32, ik, xd).cpu().numpy(), t.qbq(invB, x)) < 2e-6
    S.close()


@pytest.mark.parametrize("sname", list(STRUCTS))
def test_check_eigendecomp(gpu, engine, sname):
    """fame_check_eigendecomp (the reference's Check_Eigendecomp on the GPU):
    the sparse forward differences and curl against the factorised operators,
    at a generic wave vector and at Gamma, below the reference's 1e-6 and at
    FP64 rounding relative to ||C x||; and its sparse side equals the product's
    COO curl applied on the host."""
    import torch
    lo, fl, ol, eps = make(sname)
    ks = np.stack([KPTS["k"], KPTS["gamma"]])
    S = F.Solver(fl, max_batch=2)
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    N = S.N
    for ik, k in enumerate(ks):
        err = S.check_eigendecomp(ik, seed=7 + ik)
        r, c, v = F.curl_coo(fl, k)
        Cm = sp.coo_matrix((v, (r, c)), shape=(3 * N, 3 * N)).tocsr()
        x = np.random.default_rng(3).random(3 * N) + 1j * np.random.default_rng(4).random(3 * N)
        scale = np.linalg.norm(Cm @ x)
        assert np.all(err < 1e-6), (sname, ik, err)
        assert np.all(err < 1e-13 * scale), (sname, ik, err, scale)
        # the factorised side through the public operators: Pr(Sigma Qrs x) = C x
        q = S.apply(F.OP_QRS, ik, torch.tensor(x, device=gpu)).cpu().numpy()
        t = O.Tables(ol, k)
        sig = t.get(3)
        q[: t.Nd] *= sig
        q[t.Nd:] *= sig
        y = S.apply(F.OP_PR, ik, torch.tensor(q, device=gpu)).cpu().numpy()
        assert np.linalg.norm(y - Cm @ x) <= 1e-13 * scale
    S.close()
