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
(one fame_eigs call) with the restored E fields and the GPU relative-residual
check of every band (fame_residuals), as the metric counts a k-point's
eigensolve.  The plain FP64 solver (one wave vector at a time, FP64 throughout:
the reference algorithm, also with restoration and residuals) is timed beside
it on `--plain-nk` of the wave vectors, and the per-band eigenvalue error of the
batched solve against it is reported.  The roofline entry prices the solve's
dominant kernel -- one of the FP32 CG iteration kernels with their fused vector
updates -- with its hipEvent-timed average duration and algorithmic bytes;
cpu_baseline times the FP64 CPU restatement (oracle/) on complete invAr solves
and Gram-Schmidt steps at the same grid, priced with the plain solve's invAr
count per k-point.
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


def _host_cores():
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    return cores


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(fl, eps, k, es_iter_per_k, nwant, nstep, reps=5):
    """The FP64 CPU restatement of the reference algorithm (oracle/) on the
    host cores, priced per wave vector of the same workload:

      seconds per k = (invAr solves per k) x (t_invAr + t_reorth)

    t_invAr: median of `reps` complete invAr solves (Sigma^-1 CG(QBQ, Sigma^-1 y)
    to the workload's absolute tolerance, CG vector work included) from unit
    random right-hand sides, on all host cores.  t_reorth: median of `reps`
    classical Gram-Schmidt steps (h = U^H w, w -= U h, numpy/BLAS) against the
    average number of basis columns a step sees, (nwant + nstep) / 2.  The invAr
    count per k is the one the plain FP64 GPU solve measured (same algorithm).
    The single-thread figure times `reps` QBQ applies on one thread and scales
    t_invAr by the all-core QBQ-to-invAr ratio; its reorthogonalisation is timed
    on one BLAS thread.  Lanczos restarts and the tridiagonal solves (O(nstep^2)
    per restart) are left out."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from threadpoolctl import threadpool_limits
    cores = _host_cores()
    O.set_threads(cores)
    t = O.Tables(O.lattice_from(fl), k)
    rng = np.random.default_rng(0)
    invB = 1.0 / eps
    nd2 = 2 * t.Nd

    def unit():
        y = rng.standard_normal(nd2) + 1j * rng.standard_normal(nd2)
        return y / np.linalg.norm(y)

    x = unit()
    t.qbq(invB, x)                                        # warm (plans, pages)
    t_qbq, _ = _median_time(lambda: t.qbq(invB, x), reps)
    its = []
    t_inv, _ = _median_time(lambda: its.append(t.invar(invB, unit(), maxit=1000, tol=1e-12)[1]), reps)
    m = (nwant + nstep) // 2
    U = rng.standard_normal((nd2, m)) + 1j * rng.standard_normal((nd2, m))
    w = unit()

    def cgs():
        h = U.conj().T @ w
        w[:] = w - U @ h

    t_ro, _ = _median_time(cgs, reps)
    O.set_threads(1)
    t_qbq1, _ = _median_time(lambda: t.qbq(invB, x), reps)
    with threadpool_limits(limits=1):
        t_ro1, _ = _median_time(cgs, reps)
    O.set_threads(cores)
    t_inv1 = t_qbq1 * (t_inv / t_qbq)
    per_k = es_iter_per_k * (t_inv + t_ro)
    per_k1 = es_iter_per_k * (t_inv1 + t_ro1)
    del U
    return {"value": 1.0 / per_k, "unit": "k-pts/s", "cores": cores, "kind": "port",
            "value_1thread": 1.0 / per_k1,
            "qbq_applies_per_s": 1.0 / t_qbq, "qbq_applies_per_s_1thread": 1.0 / t_qbq1,
            "sample": f"FP64 CPU restatement (oracle/) at {tuple(fl.grid_nums)}: median of {reps} complete invAr "
                      f"solves on {cores} threads ({t_inv:.3f} s, {int(np.median(its))} CG iterations) and of {reps} "
                      f"Gram-Schmidt steps against {m} columns ({t_ro:.3f} s), x {es_iter_per_k} invAr per k-point "
                      f"(the plain FP64 GPU solve's count); 1 thread: {reps} QBQ applies ({t_qbq1:.3f} s vs "
                      f"{t_qbq:.3f} s on {cores}) scaling the invAr time, Gram-Schmidt on one BLAS thread "
                      f"({t_ro1:.3f} s); restarts and tridiagonal solves not included"}


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
    # each timing runs twice and the second is kept: the first brings the
    # clocks and caches to the state a running solve has
    if args.cg_only:
        S.time_cg_passes(4, args.reps)
        cg32 = S.time_cg_passes(4, args.reps)
        us32 = us64 = None
    else:
        S.time_qbq_passes(4, args.reps)
        us32 = S.time_qbq_passes(4, args.reps)
        S.time_qbq_passes(8, args.reps)
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
    out = {"grid": list(fl.grid_nums), "nk": args.nk, "copy_bytes": COPY_BYTES, "copy_us": cp_us,
           "copy_gbs": 2 * COPY_BYTES / cp_us / 1e3, "tuned_tiles": tune}
    if args.cg_only:
        names = ["cg_inner_expand", "cg_mid_inverse", "cg_outer_eps", "cg_alpha", "cg_mid_forward",
                 "cg_inner_contract", "cg_beta", "cg_iteration"]
        out["fp32_cg_us"] = dict(zip(names, map(float, cg32)))
    else:
        names = ["inner_expand", "mid_inverse", "outer_eps", "mid_forward", "inner_contract", "qbq"]
        out["fp32_us"] = dict(zip(names, map(float, us32)))
        out["fp64_us"] = dict(zip(names, map(float, us64)))
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=96)
    ap.add_argument("--nk", type=int, default=30)
    ap.add_argument("--nwant", type=int, default=16)
    ap.add_argument("--nstep", type=int, default=48)
    ap.add_argument("--plain-nk", type=int, default=30, help="wave vectors timed with the plain FP64 solver")
    ap.add_argument("--cpu-reps", type=int, default=5, help="repetitions (median) of each CPU baseline timing")
    ap.add_argument("--refine-sweeps", type=int, default=2,
                    help="inverse-iteration sweeps of one extra, untimed batched solve reported under 'refined' (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--passes-only", action="store_true",
                    help="only time the five QBQ passes on the batch (FP32 and FP64) and one plain copy of "
                         "known size (the reference kernel for HBM counter scaling); for profiler runs")
    ap.add_argument("--cg-only", action="store_true",
                    help="with --passes-only: time only the kernels of one fused FP32 CG iteration (for counter runs "
                         "of those kernels alone)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the warm-up and timed batched solves, one short JSON line (for kernel traces of "
                         "the timed step)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--tiles", type=str, default=None,
                    help="with --passes-only: fixed tiles instead of tuning, 'fp32:e,L,e,L,e,L;fp64:e,L,e,L,e,L' "
                         "(engine, lines per block of the inner, middle and outer passes, as bench.py reports them)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solver shards over wave vectors with no communication, "
              "and runs on several GPUs are not implemented", file=sys.stderr)
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
    def solve(solver, prec):
        # one k-point eigensolve as the metric counts it: Lanczos, restoration
        # of the fields, and the relative residual check on the GPU
        freq, _, E, prof = solver.eigs(args.nwant, args.nstep, tol=tol, maxit=maxit, precision=prec, want_fields=True)
        res = solver.residuals(freq, E)
        del E
        return freq, res, prof

    for _ in range(args.warmup):
        solve(S, F.PREC_F32_REFINE)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_s = []
    for _ in range(args.steps):
        ts = time.perf_counter()
        S.set_wave_vectors(ks)
        freq_b, res_b, prof_b = solve(S, F.PREC_F32_REFINE)
        torch.cuda.synchronize()
        step_s.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    kps_b = args.nk * args.steps / t_b
    log(f"batched FP32+refine: {t_b:.2f}s for {args.steps} step(s) -> {kps_b:.3f} k-pts/s")
    if args.batched_only:
        S.close()
        print(json.dumps({"metric": "k-pts/s (batched FP32+refine only)", "value": kps_b, "unit": "k-pts/s",
                          "steps": args.steps, "warmup": args.warmup, "step_ms": [1000.0 * v for v in step_s],
                          "qbq32_per_k": float(np.mean([p["qbq32"] for p in prof_b])),
                          "qbq64_per_k": float(np.mean([p["qbq64"] for p in prof_b])),
                          "es_iter_mean": float(np.mean([p["es_iter"] for p in prof_b]))}))
        return

    # ---- the same batched solve with two inverse-iteration sweeps of the Ritz
    # pairs (fame_set_refinement), outside the timed steps: its cost and the field
    # residual it reaches, which the reference's stopping rule alone does not ----
    refined = None
    if args.refine_sweeps > 0:
        S.set_refinement(args.refine_sweeps, 1e-16, 6000)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S.set_wave_vectors(ks)
        freq_r, res_r, prof_r = solve(S, F.PREC_F32_REFINE)
        torch.cuda.synchronize()
        t_r = time.perf_counter() - t0
        S.set_refinement(0, 1e-16, 6000)
        m = freq_r > 0
        refined = {"sweeps": args.refine_sweeps, "cg_tol": 1e-16, "k_pts_per_s": args.nk / t_r,
                   "max_rel_residual": float(np.max(res_r[m])),
                   "max_rel_eig_change": float(np.max(np.abs(freq_r[m] ** 2 - freq_b[m] ** 2) / freq_b[m] ** 2)),
                   "curlcurl_norm_bound": float(sum((2.0 / h) ** 2 for h in fl.mesh_len)),
                   "max_rel_residual_over_eps_norm": float(np.max(res_r[m])) /
                   (2.220446049250313e-16 * float(sum((2.0 / h) ** 2 for h in fl.mesh_len)))}
        log(f"refined ({args.refine_sweeps} sweeps): {t_r:.2f}s, max rel residual {refined['max_rel_residual']:.2e}")

    # ---- roofline of the solve's dominant kernel ----
    # FP32 CG iterations are most of the batched solve (qbq32 vs qbq64 per k in
    # the profile); its kernels are timed as the solver launches them, fused
    # vector updates included, with all wave vectors active.
    tiles = {"fp32": S.autotune(4), "fp64": S.autotune(8)}    # the choices the solve used (same batch shape)
    us32 = S.time_qbq_passes(4, 20)
    us64 = S.time_qbq_passes(8, 20)
    S.time_cg_passes(4, 20)            # the first pass brings clocks and caches to the solve's state
    cg32 = S.time_cg_passes(4, 20)
    N = S.N
    s32, r32 = 8, 4
    cg_names = ["cg_inner_expand", "cg_mid_inverse", "cg_outer_eps", "cg_alpha", "cg_mid_forward",
                "cg_inner_contract", "cg_beta"]
    # algorithmic bytes per wave vector: expand reads p, r, x (3 x 2N) and writes
    # p, x (2 x 2N) and the field (3N); the outer pass adds 1/eps (3N reals); the
    # contracting pass reads the field (3N) and reads and writes r (2 x 2N)
    cg_bytes = [13 * N * s32, 6 * N * s32, 6 * N * s32 + 3 * N * r32, 0, 6 * N * s32, 7 * N * s32, 0]
    dom = int(np.argmax(cg32[:7]))
    alg_bytes = cg_bytes[dom] * args.nk
    achieved = alg_bytes / (cg32[dom] * 1e-6) / 1e9
    traffic, traffic_note = measured_traffic(f"k_{cg_names[dom]}_fp32", tiles, list(fl.grid_nums), args.nk)
    qbq_alg = (10 * N * s32 + 3 * N * r32) * args.nk
    qbq_gbs = qbq_alg / (us32[5] * 1e-6) / 1e9
    cg_iter_gbs = sum(cg_bytes) * args.nk / (cg32[7] * 1e-6) / 1e9
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
        solve(P, F.PREC_F64)                                                        # warm
        freq_p = np.zeros((plain_nk, args.nwant))
        res_p = np.zeros((plain_nk, args.nwant))
        invar_per_k, cg_per_k = [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(plain_nk):
            P.set_wave_vectors(ks[i:i + 1])
            f, r, pr = solve(P, F.PREC_F64)
            freq_p[i], res_p[i] = f[0], r[0]
            invar_per_k.append(pr[0]["es_iter"])
            cg_per_k.append(pr[0]["ls_iter"])
        torch.cuda.synchronize()
        t_p = time.perf_counter() - t0
        kps_p = plain_nk / t_p
        qbq_per_k = int(np.mean(invar_per_k))
        P.close()
        log(f"plain FP64: {t_p:.2f}s for {plain_nk} wave vectors -> {kps_p:.3f} k-pts/s")

    err = None
    if freq_p is not None:
        lb, lp = freq_b[:plain_nk] ** 2, freq_p ** 2
        m = lp > 0
        err = float(np.max(np.abs(lb[m] - lp[m]) / lp[m]))

    cpu = None
    if not args.no_cpu and qbq_per_k:
        cpu = cpu_baseline(fl, eps, ks[min(3, args.nk - 1)], qbq_per_k, args.nwant, args.nstep, args.cpu_reps)

    # each profile entry holds the mean over the wave vectors solved with it (a
    # batch, or one of the concurrent groups a batch is split into)
    q32_k = float(np.mean([p["qbq32"] for p in prof_b]))
    q64_k = float(np.mean([p["qbq64"] for p in prof_b]))

    def max_rel_res(freq, res):
        m = freq > 0
        return float(np.max(res[m] / freq[m] ** 2)) if m.any() else None, float(np.max(res[m])) if m.any() else None

    out = {
        "metric": "k-pts/s @96^3,16 bands (batched FP32+refine vs plain FP64); HBM GB/s",
        "value": kps_b,
        "unit": "k-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * t_b / args.steps,
        "step_ms": [1000.0 * v for v in step_s],
        "median_step_ms": 1000.0 * float(np.median(step_s)),
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
                       "speedup_batched_vs_plain": (kps_b / kps_p) if kps_p else None,
                       "invar_per_k": qbq_per_k,
                       "qbq_applies_per_s": (float(np.sum(cg_per_k)) / t_p) if kps_p else None},
        "qbq_applies_per_s": {"batched": (q32_k + q64_k) * args.nk / (t_b / args.steps),
                              "definition": "QBQ matvecs (one structure at one k each; FP32 and FP64) per second"},
        "max_rel_eig_err_vs_fp64": err,
        "max_rel_residual": {"batched": max_rel_res(freq_b, res_b)[1],
                             "plain_fp64": max_rel_res(freq_p, res_p)[1] if freq_p is not None else None,
                             "definition": "max over bands of ||C^H C E - omega^2 B E|| / ||E|| (GPU fame_residuals)"},
        "refined": refined,
        "profile_batched": {"ls_iter_mean": float(np.mean([p["ls_iter"] for p in prof_b])),
                            "es_iter_mean": float(np.mean([p["es_iter"] for p in prof_b])),
                            "restarts_max": int(max(p["restarts"] for p in prof_b)),
                            "qbq32_per_k": q32_k, "qbq64_per_k": q64_k,
                            "all_converged": bool(all(p["converged"] for p in prof_b))},
        "qbq_fp32_us_per_batch": us32[5], "qbq_fp64_us_per_batch": us64[5],
        "qbq_path_gbs": qbq_gbs,
        "cg_iteration_us": float(cg32[7]), "cg_iteration_gbs": cg_iter_gbs,
        "cg_kernels_us": dict(zip(cg_names, map(float, cg32[:7]))),
        "tuned_tiles": tiles,
        "roofline": {"bound": "hbm", "kernel": f"{cg_names[dom]} (FP32 k_inner/k_mid/k_outer with the fused CG "
                                               f"updates, {args.nk} wave vectors active per launch)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "avg_launch_us": float(cg32[dom]), "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
