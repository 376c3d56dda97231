"""The pass kernels under every code path that a grid and a tile choice
select, against the FP64 oracle: two modes per thread with 16-byte accesses
(even line lengths, planes and lines per block) or one mode per thread (odd
ones), the fused matrix-core middle kernel (one direct DFT, 16 or 32 lines per
block) or the general middle kernel (Stockham, or fewer lines), and the General
gauge applied without forming the basis.  Tiles are fixed with set_tiles, so
each case runs the configuration it names; the FP32 QBQ is held to 2e-6 of the
oracle's FP64 QBQ, the FP64 QBQ to 1e-11 (its fused matrix-core middle kernel
among the paths), and an FP32-refined eigensolve to 1e-8 in lambda."""
import itertools

import numpy as np
import pytest

import fame_amd as F
import oracle as O
from cases import SPHERE

pytestmark = pytest.mark.gpu

O.set_threads(8)

# (lattice type, nx, keyword arguments): odd and even line lengths, prime
# middle axes (direct DFT), Simple and General gauges
GRIDS = [("simple_cubic", 9, {}), ("simple_cubic", 18, {}), ("face_centered_cubic", 22, {}),
         ("body_centered_cubic", 17, {}), ("hexagonal", 16, {"c": 1.2})]


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _setup(lt, nx, kw):
    lo = F.build_lattice(lt, nx, **kw)
    fl = F.to_fame_lattice(lo)
    eps = F.material_eps(lo, SPHERE, [13.0], 1.0)
    return fl, O.lattice_from(fl), eps


def _choices(S, prec=4):
    """Feasible tile sets of one precision (4: FP32, 8: FP64): engine 0 / 1 and
    several lines per block for each pass, combined so that every (pass,
    engine, lines) value occurs."""
    opts = {}
    for name in ("inner", "mid", "outer"):
        opts[name] = []
        for eng, L in itertools.product((0, 1), (5, 8, 12, 16, 32)):
            base = {"inner": (0, 1), "mid": (0, 1), "outer": (0, 1)}
            base[name] = (eng, L)
            try:
                S.set_tiles(prec, base)
            except RuntimeError:
                continue
            opts[name].append((eng, L))
    n = max(len(v) for v in opts.values())
    return [{k: v[i % len(v)] for k, v in opts.items()} for i in range(n)]


@pytest.mark.parametrize("lt,nx,kw", GRIDS, ids=[f"{g[0]}-{g[1]}" for g in GRIDS])
def test_fp32_qbq_every_tile_path(gpu, lt, nx, kw):
    import torch
    fl, ol, eps = _setup(lt, nx, kw)
    ks = np.stack([np.array([0.21, -0.13, 0.07]), np.zeros(3)])
    S = F.Solver(fl, max_batch=2)
    S.set_eps(eps)
    nd = S.set_wave_vectors(ks)
    tabs = [O.Tables(ol, k) for k in ks]
    xs = [_rand(2 * nd[i], 40 + i) for i in range(2)]
    refs = [tabs[i].qbq(1.0 / eps, xs[i]) for i in range(2)]
    sets = _choices(S)
    assert sets, "no feasible FP32 tiles"
    print(f"{lt} {nx}: grid {tuple(fl.grid_nums)}, {len(sets)} tile sets", flush=True)
    for tiles in sets:
        S.set_tiles(4, tiles)
        for i in range(2):
            y = S.apply(F.OP_QBQ32, i, torch.tensor(xs[i], device=gpu)).cpu().numpy()
            assert _rel(y, refs[i]) < 2e-6, (tiles, i, _rel(y, refs[i]))
    S.close()


@pytest.mark.parametrize("lt,nx,kw", GRIDS, ids=[f"{g[0]}-{g[1]}" for g in GRIDS])
def test_fp32_refined_eigs_every_tile_path(gpu, lt, nx, kw):
    fl, ol, eps = _setup(lt, nx, kw)
    k = np.array([0.21, -0.13, 0.07])
    fo, _, _ = O.eigs(ol, eps, k, 6, 20)
    S = F.Solver(fl, max_batch=1)
    S.set_eps(eps)
    S.set_wave_vectors(k)
    for tiles in _choices(S):
        S.set_tiles(4, tiles)
        f, _, _, prof = S.eigs(6, 20, precision=F.PREC_F32_REFINE, want_fields=False)
        assert prof[0]["converged"] == 1, tiles
        assert np.max(np.abs(f[0] ** 2 - fo ** 2) / fo ** 2) < 1e-8, (tiles, f[0], fo)
    S.close()


@pytest.mark.parametrize("lt,nx,kw", GRIDS, ids=[f"{g[0]}-{g[1]}" for g in GRIDS])
def test_fp64_qbq_every_tile_path(gpu, lt, nx, kw):
    import torch
    fl, ol, eps = _setup(lt, nx, kw)
    ks = np.stack([np.array([0.21, -0.13, 0.07]), np.zeros(3)])
    S = F.Solver(fl, max_batch=2)
    S.set_eps(eps)
    nd = S.set_wave_vectors(ks)
    tabs = [O.Tables(ol, k) for k in ks]
    xs = [_rand(2 * nd[i], 60 + i) for i in range(2)]
    refs = [tabs[i].qbq(1.0 / eps, xs[i]) for i in range(2)]
    sets = _choices(S, 8)
    assert sets, "no feasible FP64 tiles"
    for tiles in sets:
        S.set_tiles(8, tiles)
        for i in range(2):
            y = S.apply(F.OP_QBQ, i, torch.tensor(xs[i], device=gpu)).cpu().numpy()
            assert _rel(y, refs[i]) < 1e-11, (tiles, i, _rel(y, refs[i]))
    S.close()
