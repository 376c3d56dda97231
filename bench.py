#!/usr/bin/env python3
"""Headline benchmark: k-points/s of the band solver on one MI355X.

Workload (the metric's configuration): FCC lattice a=1, diamond network -- two
spheres r=0.15a at fractional (0,0,0) and (1/4,1/4,1/4) joined by four rods
r=0.10a (the bonds of the reference's No227_FCC_DiamondStructure file), eps 13/1,
nominal grid 96 -> 96x83x78 by the reference's grid rule, 30 wave vectors spaced
evenly in arc length on Gamma-X-W-L-Gamma, 16 bands, Krylov dimension 48,
Lanczos and CG tolerance 1e-12.  Synthetic input: the structure is generated,
nothing is read from disk.

One step = one batched FP32+FP64-refinement eigensolve of all 30 wave vectors
(one fame_eigs call).  The plain FP64 solver (one wave vector at a time, FP64
throughout: the reference algorithm) is timed beside it on `--plain-nk` of the
wave vectors, and the per-band eigenvalue error of the batched solve against it
is reported.  The roofline entry prices the dominant QBQ pass kernel with its
hipEvent-timed average duration; cpu_baseline times the FP64 CPU restatement
(oracle) on a bounded sample of QBQ applies at the same grid.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fame-isotropic-gpu_amd", "python"))

HBM_PEAK_GBS = 8000.0


def build_workload(nx, nk):
    import fame_amd as F
    lo = F.build_lattice("face_centered_cubic", nx)
    fl = F.to_fame_lattice(lo)
    mat = F.Material(sphere_centers=[[(0, 0, 0), (0.25, 0.25, 0.25)]], sphere_radius=[[0.15, 0.15]],
                     cylinder_top=[[(0.25, 0.25, 0.25)] * 4],
                     cylinder_bot=[[(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]],
                     cylinder_radius=[[0.10] * 4])
    eps = F.material_eps(lo, mat, [13.0], 1.0)
    ks = F.path_by_arclength(lo, "GXWLG", nk)
    return F, lo, fl, eps, ks


def cpu_baseline(fl, eps, k, qbq_per_k, seconds):
    """FP64 oracle QBQ applies on all host cores for ~`seconds`; k-pts/s from the
    measured QBQ count per k-point of the same algorithm."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    O.set_threads(cores)
    t = O.Tables(O.lattice_from(fl), k)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(2 * t.Nd) + 1j * rng.standard_normal(2 * t.Nd)
    invB = 1.0 / eps
    t.qbq(invB, x)                       # warm (plans, pages)
    n, t0 = 0, time.perf_counter()
    while True:
        t.qbq(invB, x)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    per_qbq = el / n
    return {"value": 1.0 / (per_qbq * qbq_per_k), "unit": "k-pts/s", "cores": cores, "kind": "port",
            "sample": f"{n} FP64 QBQ applies of the CPU restatement at {tuple(fl.grid_nums)} on {cores} threads "
                      f"({per_qbq:.3f} s each) x {qbq_per_k} QBQ applies per k-point measured by the plain FP64 "
                      f"GPU solve; CPU CG/Lanczos vector work not included"}


COPY_BYTES = 1 << 30
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_fcc96_30k.json")


def measured_traffic(kernel_key, tiles, grid, nk):
    """HBM bytes per launch of a pass kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (scaled by the 1 GiB streaming-kernel
    calibration), used when this run's tuned tiles are the profiled ones."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None, "no traffic profile"
    want = ";".join(f"{p}:" + ",".join(str(v) for name in ("inner", "mid", "outer") for v in tiles[p][name])
                    for p in ("fp32", "fp64"))
    if prof.get("tiles") != want or list(prof.get("grid", [])) != list(grid) or prof.get("nk") != nk:
        return None, f"profiled tiles {prof.get('tiles')} differ from this run's {want}"
    k = prof["kernels"].get(kernel_key)
    return (k["traffic_bytes"], "rocprofv3 FETCH_SIZE+WRITE_SIZE, scaled") if k else (None, "kernel not profiled")


def passes_only(F, fl, eps, ks, args, torch, dev):
    """Per-pass QBQ timings of the batch and a 1 GiB elementwise negation
    (a plain streaming kernel that reads and writes COPY_BYTES once each: the
    reference for scaling the HBM byte counters), printed as one JSON line."""
    S = F.Solver(fl, max_batch=args.nk)
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    if args.tiles:
        tune = {}
        for part in args.tiles.split(";"):
            key, vals = part.split(":")
            v = [int(t) for t in vals.split(",")]
            tune[key] = {"inner": v[0:2], "mid": v[2:4], "outer": v[4:6]}
            S.set_tiles(4 if key == "fp32" else 8, tune[key])
    else:
        tune = {"fp32": S.autotune(4), "fp64": S.autotune(8)}
    us32 = S.time_qbq_passes(4, args.reps)
    us64 = S.time_qbq_passes(8, args.reps)
    S.close()
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    torch.neg(a, out=b)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.reps):
        torch.neg(a, out=b)
    e1.record()
    e1.synchronize()
    cp_us = e0.elapsed_time(e1) / args.reps * 1e3
    names = ["inner_expand", "mid_inverse", "outer_eps", "mid_forward", "inner_contract", "qbq"]
    print(json.dumps({"grid": list(fl.grid_nums), "nk": args.nk,
                      "fp32_us": dict(zip(names, map(float, us32))), "fp64_us": dict(zip(names, map(float, us64))),
                      "copy_bytes": COPY_BYTES, "copy_us": cp_us, "copy_gbs": 2 * COPY_BYTES / cp_us / 1e3,
                      "tuned_tiles": tune}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=96)
    ap.add_argument("--nk", type=int, default=30)
    ap.add_argument("--nwant", type=int, default=16)
    ap.add_argument("--nstep", type=int, default=48)
    ap.add_argument("--plain-nk", type=int, default=30, help="wave vectors timed with the plain FP64 solver")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--passes-only", action="store_true",
                    help="only time the five QBQ passes on the batch (FP32 and FP64) and one plain copy of "
                         "known size (the reference kernel for HBM counter scaling); for profiler runs")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--tiles", type=str, default=None,
                    help="with --passes-only: fixed tiles instead of tuning, 'fp32:e,L,e,L,e,L;fp64:e,L,e,L,e,L' "
                         "(engine, lines per block of the inner, middle and outer passes, as bench.py reports them)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solver shards over wave vectors with no communication, "
              "and multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import torch
    F, lo, fl, eps, ks = build_workload(args.grid, args.nk)
    dev = torch.device("cuda", 0)
    tol, maxit = 1e-12, 300
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    log(f"grid {tuple(fl.grid_nums)} N={fl.grid_nums[0]*fl.grid_nums[1]*fl.grid_nums[2]} nk={args.nk}")

    if args.passes_only:
        passes_only(F, fl, eps, ks, args, torch, dev)
        return

    # ---- batched FP32 + FP64 refinement (the headline step) ----
    S = F.Solver(fl, max_batch=args.nk)
    S.set_eps(eps)
    S.set_wave_vectors(ks)
    for _ in range(args.warmup):
        S.eigs(args.nwant, args.nstep, tol=tol, maxit=maxit, precision=F.PREC_F32_REFINE)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        S.set_wave_vectors(ks)
        freq_b, _, _, prof_b = S.eigs(args.nwant, args.nstep, tol=tol, maxit=maxit, precision=F.PREC_F32_REFINE)
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    kps_b = args.nk * args.steps / t_b
    log(f"batched FP32+refine: {t_b:.2f}s for {args.steps} step(s) -> {kps_b:.3f} k-pts/s")

    # ---- roofline of the dominant QBQ pass (FP32 passes dominate the batched solve) ----
    tiles = {"fp32": S.autotune(4), "fp64": S.autotune(8)}    # the choices the solve used (same batch shape)
    us32 = S.time_qbq_passes(4, 10)
    us64 = S.time_qbq_passes(8, 10)
    N = S.N
    s32, s64, r32 = 8, 16, 4
    bytes_pass = lambda s: [5 * N * s, 6 * N * s, 6 * N * s + 3 * N * (s // 2), 6 * N * s, 5 * N * s]
    names = ["inner_expand", "mid_inverse", "outer_eps", "mid_forward", "inner_contract"]
    tot32 = [us32[i] for i in range(5)]
    dom = int(np.argmax(tot32))
    alg_bytes = bytes_pass(s32)[dom] * args.nk
    achieved = alg_bytes / (us32[dom] * 1e-6) / 1e9
    traffic, traffic_note = measured_traffic(f"k_{names[dom]}_fp32", tiles, list(fl.grid_nums), args.nk)
    qbq_alg = (10 * N * s32 + 3 * N * r32) * args.nk
    qbq_gbs = qbq_alg / (us32[5] * 1e-6) / 1e9
    S.close()

    # ---- plain FP64, one wave vector at a time ----
    plain_nk = min(args.plain_nk, args.nk)
    freq_p = None
    kps_p = None
    qbq_per_k = None
    if plain_nk > 0:
        P = F.Solver(fl, max_batch=1)
        P.set_eps(eps)
        P.set_wave_vectors(ks[:1])
        P.eigs(args.nwant, args.nstep, tol=tol, maxit=maxit, precision=F.PREC_F64)   # warm
        freq_p = np.zeros((plain_nk, args.nwant))
        qbqs = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(plain_nk):
            P.set_wave_vectors(ks[i:i + 1])
            f, _, _, pr = P.eigs(args.nwant, args.nstep, tol=tol, maxit=maxit, precision=F.PREC_F64)
            freq_p[i] = f[0]
            qbqs.append(pr[0]["qbq64"])
        torch.cuda.synchronize()
        t_p = time.perf_counter() - t0
        kps_p = plain_nk / t_p
        qbq_per_k = int(np.mean(qbqs))
        P.close()
        log(f"plain FP64: {t_p:.2f}s for {plain_nk} wave vectors -> {kps_p:.3f} k-pts/s")

    err = None
    if freq_p is not None:
        lb, lp = freq_b[:plain_nk] ** 2, freq_p ** 2
        m = lp > 0
        err = float(np.max(np.abs(lb[m] - lp[m]) / lp[m]))

    cpu = None
    if not args.no_cpu and qbq_per_k:
        cpu = cpu_baseline(fl, eps, ks[min(3, args.nk - 1)], qbq_per_k, args.cpu_seconds)

    out = {
        "metric": "k-pts/s @96^3,16 bands (batched FP32+refine vs plain FP64); HBM GB/s",
        "value": kps_b,
        "unit": "k-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * t_b / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+c128 (complex FP32 FFT/CG inside complex FP64 refinement and Lanczos)",
        "data": "synthetic (generated structure, no files)",
        "config": {"workload": "FCC a=1 diamond network (2 spheres r=0.15 + 4 rods r=0.10), eps 13/1, "
                               f"grid {args.grid}->{fl.grid_nums[0]}x{fl.grid_nums[1]}x{fl.grid_nums[2]}, "
                               f"{args.nk} k on Gamma-X-W-L-Gamma, {args.nwant} bands, nstep {args.nstep}, tol 1e-12",
                   "global_batch": args.nk, "parallelism": "single GPU, all wave vectors in one batch"},
        "plain_fp64": {"value": kps_p, "unit": "k-pts/s", "nk_timed": plain_nk,
                       "speedup_batched_vs_plain": (kps_b / kps_p) if kps_p else None},
        "max_rel_eig_err_vs_fp64": err,
        "profile_batched": {"ls_iter_mean": float(np.mean([p["ls_iter"] for p in prof_b])),
                            "es_iter_mean": float(np.mean([p["es_iter"] for p in prof_b])),
                            "restarts_max": int(max(p["restarts"] for p in prof_b)),
                            "qbq32_per_k": prof_b[0]["qbq32"], "qbq64_per_k": prof_b[0]["qbq64"],
                            "all_converged": bool(all(p["converged"] for p in prof_b))},
        "qbq_fp32_us_per_batch": us32[5], "qbq_fp64_us_per_batch": us64[5],
        "qbq_path_gbs": qbq_gbs,
        "tuned_tiles": tiles,
        "roofline": {"bound": "hbm", "kernel": f"k_{names[dom]} (FP32, {args.nk} problems per launch)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "avg_launch_us": us32[dom], "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
