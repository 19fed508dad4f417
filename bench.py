#!/usr/bin/env python3
"""Benchmark of the FP64 finite-volume Euler hot path on one MI355X.

Headline workload (--workload batch, the default): 64 independent
four-quadrant Riemann problems on [0,1]^2 (seeds 1000..1063, states from
numpy.random.default_rng(seed): rho log-uniform [0.1,2], u,v uniform [-1,1],
p uniform [0.1,2]), each with its own finest level drawn from {5,6,7,8}
(numpy.random.default_rng(2024)), refined by the density-gradient indicator
every 10 steps with 2:1 balance, transmissive boundaries, SSP-RK3 with input
CFL 0.4, FP64.

Setup (untimed): the 64 problems are advanced together to t = --batch-t-end
in one device context (amr_evolve_batch: one launch per RK stage over every
instance that has not reached its stop, instances leaving the device mesh at
the regrid after their stop), once with 8x8 and once with 16x16 patches
(16x16 one level coarser at both ends: the same finest sub-cell size), the
stage-kernel tile autotuned on the run's own mesh.  The patch size with the
shorter time to t_end is kept ("autotuner" in the output, both runs beside
it); the 64 meshes and states it reached are the benchmark input, resident
in HBM (about 6-10 M cells, far larger than the 256 MB Infinity Cache).

A step = one SSP-RK3 step (3 fused residual + stage-update launches, CFL
folded into the last) over all 64 instances, each with its own dt.
value = sub-grid cells x 3 stages x steps / wall time of the timed region
(HIP events on the launch stream; direct launches or graph replay, as the
autotuner measured faster).

roofline: the fused stage kernel; algorithmic bytes = 85.33 B per cell-stage
(stage 1 reads U^n and writes U^1; stages 2-3 read U^s and U^n and write
U^s+1; 32 B per state), achieved = bytes per launch / average launch
duration = the timed region / launches.  traffic = FETCH_SIZE (scaled by the
copy kernel's reading of a known size) + WRITE_SIZE per launch from the
committed rocprofv3 summary of the same mesh
(profiles/stage_kernel_<workload>_traffic.txt), when its cell count matches.

cpu_baseline: the C oracle (restatement of the reference CPU path: AoS
patches, face loop + element loop, k-form SSP-RK3) on the host, one worker per
core on contiguous quadrant ranges (as many as the GPU's share of the host
allows: at most 16), on a bounded number of steps of the same 64-instance mesh
(one dt for all: the cost per cell-stage is the same); its -O2 and
-O3 -march=native builds both reported, the faster as `value`.

Secondary lines: the single adaptive Riemann problem (levels 3-7, 8x8, its
207k-cell mesh at t = 0.25: cache-resident, so a throughput figure only),
the learned-tagger comparison, the batched-mode comparison (one by one vs
batched, plain vs learned, with L1 errors) and the double-Mach reflection.
--workload riemann4 / vortex put those single meshes in the headline instead.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "amr-demo_amd"))

METRIC = "Mcell-updates/s (RK stage) on 1 GPU, %HBM peak; plain vs learned+batched at equal L1"
HBM_PEAK_GBS = 8000.0
B_ALG_STAGE = (64.0 + 96.0 + 96.0) / 3.0     # bytes per cell-stage, SSP-RK3 Shu-Osher
CFL_EFF = 0.25 * 0.4 * 3                       # input CFL 0.4, SSP-RK3 (physics_initialize.c:193)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# ---------------------------------------------------------------------------
# the 64-instance batch (headline)
# ---------------------------------------------------------------------------
def batch_problem(k):
    rng = np.random.default_rng(2024)
    lmaxs = [int(x) for x in rng.choice([5, 6, 7, 8], size=k)]
    return [1000 + i for i in range(k)], lmaxs


def batch_forests(L, seeds, lmaxs, lmin, tol, n):
    """16x16 patches one level coarser at both ends (the same finest
    sub-cell size as 8x8 at lmax)"""
    d = {8: 0, 16: 1}[n]
    return [L.riemann_forest(s, lmin - d, lm - d, tol, n=n) for s, lm in zip(seeds, lmaxs)]


def run_batch(amrhip, forests, insts, tol, interval, n, tagger=None, lag=0, groups=1):
    t0 = time.perf_counter()
    per, tot = amrhip.evolve_batch(forests, insts, tol, interval, tile="auto", tagger=tagger, regrid_lag=lag,
                                   groups=groups)
    wall = time.perf_counter() - t0
    useful = sum(p.cell_stages for p in per)
    r = {"patch": n, "wall_s": round(wall, 3), "steps": tot.steps, "regrids": tot.regrids,
         "peak_cells": tot.peak_quads * n * n, "final_cells": sum(f.nquad for f in forests) * n * n,
         "gpu_s": round(tot.wall_gpu, 3), "regrid_s": round(tot.wall_regrid, 3),
         "regrid_split_s": {"tags": round(tot.wall_tags, 3), "host_forest": round(tot.wall_forest, 3),
                            "device_remesh": round(tot.wall_remesh, 3)},
         # cells of instances that have not reached their stop (an instance
         # leaves the launches at the regrid after its stop)
         "useful_cell_stages": useful, "launched_cell_stages": tot.cell_stages,
         "end_to_end_Mcell_upd_s": useful / wall / 1e6,
         "gpu_Mcell_upd_s": useful / tot.wall_gpu / 1e6,
         "gpu_hbm_frac": useful / tot.wall_gpu * B_ALG_STAGE / (HBM_PEAK_GBS * 1e9)}
    if groups > 1:
        # the groups' GPU phases overlap one another: no single GPU time
        r.update({"groups": groups, "gpu_s": None, "gpu_Mcell_upd_s": None, "gpu_hbm_frac": None})
    return r, per


def batch_headline(amrhip, args):
    """runs the 64 problems to t_end with each candidate patch size, keeps the
    faster; returns (chosen n, the final forests of every run by patch size,
    the runs, (seeds, finest levels))"""
    from amrhip import learned as L
    seeds, lmaxs = batch_problem(args.batch_instances)
    runs, forests = {}, {}
    # one small run first: the runtime's one-time set-up stays out of the timings
    L.riemann_params(0.01, regrid=10)
    L.riemann_forest(0, 2, 4, args.tol).evolve_device(args.tol, 10)
    for n in args.batch_patches:
        L.riemann_params(args.batch_t_end, regrid=10, n=n)
        p = amrhip.get_params()
        insts = [amrhip.make_inst(1.0, (0, 0, 0, 0), p.cfl, p.dt, args.batch_t_end) for _ in seeds]
        fb = batch_forests(L, seeds, lmaxs, args.lmin, args.tol, n)
        r, _ = run_batch(amrhip, fb, insts, args.tol, 10, n)
        runs[n], forests[n] = r, fb
        log(f"batch of {len(seeds)}, {n}x{n} patches: {r['wall_s']:.2f} s to t={args.batch_t_end}, "
            f"{r['steps']} steps, {r['regrids']} regrids, peak {r['peak_cells']} cells, final {r['final_cells']}, "
            f"{r['end_to_end_Mcell_upd_s']:.0f} Mcell-upd/s end to end, {r['gpu_Mcell_upd_s']:.0f} over GPU time")
    best = min(runs, key=lambda k: runs[k]["wall_s"])
    return best, forests, runs, (seeds, lmaxs)


def save_forests(forests, d):
    os.makedirs(d, exist_ok=True)
    for k, f in enumerate(forests):
        f.save(os.path.join(d, f"inst_{k:03d}.bin"))


def load_forests(amrhip, d, k):
    return [amrhip.Forest.load(os.path.join(d, f"inst_{i:03d}.bin"))[0] for i in range(k)]


# ---------------------------------------------------------------------------
# single adaptive Riemann problem (the 207k-cell secondary line)
# ---------------------------------------------------------------------------
def build_riemann_mesh(amrhip, seed, lmin, lmax, n, tol, t_end, regrid, device=True):
    """run the adaptive Riemann problem on the GPU path to t_end; return the
    forest (mesh + state) and run statistics.  device=True keeps the state in
    HBM across regrids (amr_evolve_device); False is the interval-by-interval
    path through hpc_amr_evolve_solver with host regrids (amr_evolve)"""
    ic = amrhip.riemann4_ic(seed)
    f = amrhip.Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, n, lmin, lmax)
    f.apply_ic(ic)
    f.regrid_initial(ic, tol)
    amrhip.set_params(nelem_subgrid=n, time_scheme=3, time_steps=10 ** 9, dt=1.0,
                      cfl=CFL_EFF, coarse_grid_h=1.0, regrid_interval=regrid,
                      checkpoint_interval=2 ** 31 - 1, visualization_interval=2 ** 31 - 1,
                      bc=(0, 0, 0, 0, 1, 1), t_end=t_end, verbose=0)
    t0 = time.perf_counter()
    st = f.evolve_device(tol, regrid) if device else f.evolve(tol)
    wall = time.perf_counter() - t0
    return f, st, wall


def time_steps(amrhip, n, insts, arrays, steps, warmup, tile=None, autotune=True):
    """Gpu context on the given mesh: autotuned tile, warm-up, timed steps"""
    ci, fi, geom, soln, cinst = arrays
    g = amrhip.Gpu()
    g.setup(n, 3, insts, ci, fi, geom, soln, cinst)
    tune = None
    if tile:
        g.set_tile(*tile)
    elif autotune:
        tune = g.autotune(10)
        log("autotune (ppb, minw, design, ms/step): " +
            ", ".join(f"({a},{b},{c},{d:.4f})" for a, b, c, d in tune))
    if warmup:
        g.run(warmup)
    ms, stage_ms = g.bench(steps)
    return g, tune, ms, stage_ms


def roofline_of(cells, ms, steps, n, tile, launch, stage_ms):
    avg_stage = ms / (3 * steps)
    bytes_launch = cells * B_ALG_STAGE
    achieved = bytes_launch / (avg_stage * 1e-3) / 1e9
    return {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": f"stage_kernel<{n},{tile[0]},{tile[1]},{tile[2]}>",
        "bytes_per_launch": int(bytes_launch),
        "avg_launch_us": round(avg_stage * 1e3, 3),
        "avg_launch_from": f"timed region / stage launches, 3 per step, the step's dt prep launch counted in "
                           f"the time (HIP events on the launch stream, {launch})",
        "stage_launch_us_bracketed": [round(s * 1e3, 3) for s in stage_ms],
    }


def measured_traffic(workload, cells):
    """HBM bytes per stage-kernel launch from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE per dispatch, scaled
    by the copy kernel's reading of a known size), when it was taken on the
    same mesh size; None otherwise (the summary names the tile it used)"""
    path = os.path.join(ROOT, "profiles", f"stage_kernel_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    try:
        if int(kv["cells"]) != cells:
            return None
        return int(float(kv["hbm_bytes_per_launch"])), kv.get("tile", "")
    except (KeyError, ValueError):
        return None


def cpu_baseline(ci, fi, geom, soln, n, seconds, threads, h_base):
    """the C oracle on the host: each build timed on a bounded number of steps"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    lines = {}
    for path, label in ((pyoracle.ORACLE_SO, "-O2"), (pyoracle.ORACLE_NATIVE_SO, "-O3 -march=native")):
        if not os.path.exists(path):
            continue
        orc = pyoracle.Oracle(path)
        p = orc.params(n, bc=(0, 0, 0, 0), cfl=CFL_EFF, dt=1.0, h_base=h_base, scheme=3)
        t0 = time.perf_counter()
        orc.run(p, ci, fi, geom, soln, 1, threads=threads)
        one = max(time.perf_counter() - t0, 1e-4)
        k = max(1, int(seconds / 2 / one))
        t0 = time.perf_counter()
        orc.run(p, ci, fi, geom, soln, k, threads=threads)
        dt = time.perf_counter() - t0
        cells = soln.size // 4
        v = cells * 3 * k / dt / 1e6
        log(f"cpu baseline {label}: {k} steps, {dt:.2f} s, {v:.2f} Mcell-updates/s on {threads} threads")
        lines[label] = {"value": round(v, 3), "steps": k, "seconds": round(dt, 2)}
    return lines


# ---------------------------------------------------------------------------
# secondary experiments
# ---------------------------------------------------------------------------
def riemann_line(amrhip, args):
    """the single adaptive Riemann problem (levels 3-7, 8x8) to t = 0.25 on
    the state-resident loop, then timed steps on the ~207k-cell mesh reached
    (cache-resident: a throughput figure, not an HBM one)"""
    f_cold, st_cold, wall_cold = build_riemann_mesh(amrhip, args.seed, args.lmin, args.lmax, 8, args.tol,
                                                    args.t_end, 10)
    f, st, wall = build_riemann_mesh(amrhip, args.seed, args.lmin, args.lmax, 8, args.tol, args.t_end, 10)
    if (st.steps, st.regrids, f.nquad) != (st_cold.steps, st_cold.regrids, f_cold.nquad):
        raise RuntimeError("two identical adaptive runs differ")
    ci, fi, geom, soln = f.external()
    cells = f.nquad * 64
    g, tune, ms, stage_ms = time_steps(amrhip, 8, [amrhip.make_inst(1.0, (0, 0, 0, 0), CFL_EFF, 1.0, 0.0)],
                                       (ci, fi, geom, soln, None), args.steps, args.warmup)
    tile = g.tile()
    launch = "graph replay" if g.graphs() else "direct stream launches"
    g.close()
    value = cells * 3 * args.steps / (ms * 1e-3) / 1e6
    rl = roofline_of(cells, ms, args.steps, 8, tile, launch, stage_ms)
    rl["note"] = "the mesh (~6.6 MB per state vector) sits in the 256 MB Infinity Cache: not an HBM figure"
    log(f"riemann4 207k mesh: {value:.0f} Mcell-updates/s, {ms / args.steps * 1e3:.1f} us/step")
    return {
        "workload": f"riemann4 seed {args.seed}, adaptive levels {args.lmin}-{args.lmax}, 8x8 FP64 patches, "
                    f"mesh at t={st.t:.3f}",
        "cells": cells, "value_Mcell_upd_s": round(value, 3), "ms_per_step": round(ms / args.steps, 6),
        "roofline": rl,
        "adaptive_run": {
            "loop": "state resident in HBM, device indicator + remesh, host topology",
            "steps": st.steps, "regrids": st.regrids, "t": round(st.t, 6),
            "cell_stages": st.cell_stages, "wall_s": round(wall, 3),
            "gpu_s": round(st.wall_gpu, 3), "regrid_s": round(st.wall_regrid, 3),
            "end_to_end_Mcell_updates_per_s": round(st.cell_stages / wall / 1e6, 3),
            "first_run_in_process_wall_s": round(wall_cold, 3),
        },
    }


def equal_l1_plain(L, args, lmax, target, ref, interval=10, lag=0, log_fn=log):
    """plain runs with the indicator tolerance bisected (in log tol) until
    their L1 error is within 1 % of `target`; returns the matching run (or
    the closest one found) and the search trail"""
    lo, hi = 0.005, 0.5         # small tol refines more (lower L1)
    trail, best = [], None
    for _ in range(9):
        tol = float(np.sqrt(lo * hi))
        r, _ = L.deploy(args.seed, args.lmin, lmax, args.t_end, interval, None, tol, ref, lmax, lag)
        trail.append({"tol": round(tol, 5), "l1": r["l1_density"], "wall_s": round(r["wall_s"], 3)})
        log_fn(f"  equal-L1 search: plain tol {tol:.4f}: L1 {r['l1_density']:.5e} (target {target:.5e}), "
               f"{r['wall_s']:.2f} s")
        if best is None or abs(r["l1_density"] - target) < abs(best["l1_density"] - target):
            best = r
        if abs(r["l1_density"] - target) <= 0.01 * target:
            break
        if r["l1_density"] > target:
            hi = tol
        else:
            lo = tol
    return best, trail


def interpolated_wall(trail, target):
    """wall time at the target L1, linear between the closest runs on either
    side of it (L1 is not monotone in the tolerance at the 1 % level, so the
    bisection may not land within 1 %); None without a run on each side"""
    below = [t for t in trail if t["l1"] <= target]
    above = [t for t in trail if t["l1"] > target]
    if not below or not above:
        return None
    a = max(below, key=lambda t: t["l1"])
    b = min(above, key=lambda t: t["l1"])
    w = (target - a["l1"]) / (b["l1"] - a["l1"])
    return {"wall_s": a["wall_s"] + w * (b["wall_s"] - a["wall_s"]), "between_tol": [a["tol"], b["tol"]],
            "between_l1": [a["l1"], b["l1"]]}


def learned_experiment(amrhip, args):
    """plain indicator every 10 steps vs the learned tagger (with the plain
    indicator as fallback) every --learned-interval steps, on a seeded
    Riemann run with levels lmin..--learned-lmax: end-to-end cell-updates/s,
    wall time and L1 density error against the plain-method FP64 solution on
    the uniform finest level.  The network is trained on the GPU from
    snapshots of other seeds (1..k); recall is measured on two held-out
    seeds.  Plain every 40 runs with the learned run's regrid lag, so the
    network's own contribution is isolated; a plain run with its tolerance
    tuned to the learned run's L1 (within 1 %) gives the wall-time-to-solution
    comparison at equal accuracy."""
    from amrhip import learned as L
    lmax, iv, lag = args.learned_lmax, args.learned_interval, args.learned_lag
    t0 = time.perf_counter()
    ts = L.collect(range(1, 1 + args.learned_seeds), lmin=args.lmin, lmax=args.lmax, t_end=args.t_end,
                   tol=args.tol, window=iv, log=log)
    held = L.collect([100, 101], lmin=args.lmin, lmax=args.lmax, t_end=args.t_end, tol=args.tol, window=iv)
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    m, losses = L.train(ts, epochs=args.learned_epochs)
    t_train = time.perf_counter() - t0
    rec = L.recall_report(m, held, args.tol)
    rec["at_equal_refine_fraction"] = L.recall_at_equal_fraction(m, held, args.tol)
    log(f"tagger: {ts.size} training rows ({t_collect:.1f} s to collect), trained in {t_train:.2f} s, "
        f"loss {losses[0]:.3f} -> {losses[-1]:.3f}; held-out recall mlp {rec['recall_mlp']:.3f} "
        f"vs indicator-now {rec['recall_indicator_now']:.3f}")
    ref, _ = L.uniform_reference(args.seed, lmax, args.t_end, log=log)
    runs = {}
    for name, interval, tagger, lg in (("plain_every_10", 10, None, 0),
                                       (f"plain_every_{iv}_lag_{lag}", iv, None, lag),
                                       (f"learned_every_{iv}_lag_{lag}", iv,
                                        amrhip.MlpTagger(m, args.tol, True), lag)):
        r, _ = L.deploy(args.seed, args.lmin, lmax, args.t_end, interval, tagger, args.tol, ref, lmax, lg)
        runs[name] = r
        log(f"{name}: {r['steps']} steps, {r['regrids']} regrids, peak {r['peak_cells']} cells, "
            f"{r['wall_s']:.2f} s, {r['end_to_end_Mcell_upd_s']:.0f} Mcell-upd/s end to end, "
            f"L1 {r['l1_density']:.5e}")
    base = runs["plain_every_10"]
    lr = runs[f"learned_every_{iv}_lag_{lag}"]
    out = {
        "deployed_workload": f"riemann4 seed {args.seed}, levels {args.lmin}-{lmax}, 8x8 FP64, t={args.t_end}",
        "reference": f"plain FP64 on the uniform level-{lmax} forest ({ref.shape[0]}^2 cells)",
        "training": {"rows": ts.size, "label_counts": L.class_counts(ts.get(features=False)[1]),
                     "seeds": args.learned_seeds, "window": iv, "label_stride": 5,
                     "epochs": args.learned_epochs, "loss_first_last": [float(losses[0]), float(losses[-1])],
                     "collect_s": round(t_collect, 2), "train_s": round(t_train, 3)},
        "held_out_recall": rec,
        "runs": runs,
        "speedup_end_to_end": lr["end_to_end_Mcell_upd_s"] / base["end_to_end_Mcell_upd_s"],
        "l1_ratio": lr["l1_density"] / base["l1_density"],
    }
    if not args.no_equal_l1:
        eq, trail = equal_l1_plain(L, args, lmax, lr["l1_density"], ref)
        rec = {"run": eq, "search": trail,
               "l1_within_1pct": abs(eq["l1_density"] / lr["l1_density"] - 1) <= 0.01,
               "wall_time_ratio_plain_over_learned": eq["wall_s"] / lr["wall_s"]}
        iw = interpolated_wall(trail, lr["l1_density"])
        if iw is not None:
            iw["wall_time_ratio_plain_over_learned"] = iw["wall_s"] / lr["wall_s"]
            rec["interpolated_at_learned_l1"] = iw
        out["plain_every_10_at_learned_l1"] = rec
        log(f"equal L1: plain every 10 at tol {eq['tol']:.4f}: L1 {eq['l1_density']:.5e}, {eq['wall_s']:.2f} s "
            f"vs learned {lr['wall_s']:.2f} s ({eq['wall_s'] / lr['wall_s']:.2f}x)")
    return out, m


def dmr_experiment(amrhip, args, model):
    """config 5 at the deepest depth the plain scheme carries to t = 0.2: the
    Mach-10 double-Mach reflection on [0,4] x [0,1] with 8x8 patches (levels
    2..--dmr-lmax) and with 16x16 patches (one level coarser at both ends: the
    same finest cells), plain indicator every 10 steps vs the learned tagger
    (fine-tuned on double-Mach snapshots of a run one level coarser, 8x8 then
    16x16) every --learned-interval steps; L1 density error against the plain
    solution on the uniform finest level (the same grid for both patch
    sizes)"""
    from amrhip import learned as L
    lmax, iv, t_end = args.dmr_lmax, args.learned_interval, args.dmr_t_end
    out = {"workload": f"double-Mach reflection, Mach 10 at 60 deg, [0,4]x[0,1], finest level {lmax} of 8x8 "
                       f"patches (16x16: level {lmax - 1}), FP64, t={t_end}"}
    ref, _ = L.dmr_uniform_reference(lmax, t_end, log=log)
    for n in args.dmr_patches:
        d = {8: 0, 16: 1}[n]
        res = {}
        if model is not None:
            t0 = time.perf_counter()
            ts = L.dmr_collect(2 - d, lmax - 1 - d, t_end, tol=args.tol, window=iv, log=log, n=n)
            losses = model.train(ts, epochs=10, batch=4096, lr=1e-3, cw=(1.0, 1.0, 2.0), seed=7)
            res["fine_tune"] = {"rows": ts.size, "epochs": 10,
                                "loss_first_last": [float(losses[0]), float(losses[-1])],
                                "wall_s": round(time.perf_counter() - t0, 2)}
            ts.close()
        runs = [("plain_every_10", 10, None)]
        if model is not None:
            runs.append((f"learned_every_{iv}", iv, amrhip.MlpTagger(model, args.tol, True)))
        for name, interval, tagger in runs:
            r, _ = L.dmr_deploy(2 - d, lmax - d, t_end, interval, tagger, args.tol, ref, lmax - d,
                                args.learned_lag if tagger is not None else 0,
                                verbose=1 if lmax >= 9 else 0, n=n, tile="auto")
            res[name] = r
            log(f"double-Mach {n}x{n} {name}: {r['steps']} steps, {r['regrids']} regrids, peak {r['peak_cells']} "
                f"cells, {r['wall_s']:.2f} s, {r['end_to_end_Mcell_upd_s']:.0f} Mcell-upd/s, L1 {r['l1_density']:.4e}")
        if model is not None:
            a, b = res["plain_every_10"], res[f"learned_every_{iv}"]
            res["speedup_end_to_end"] = b["end_to_end_Mcell_upd_s"] / a["end_to_end_Mcell_upd_s"]
            res["wall_time_ratio_plain_over_learned"] = a["wall_s"] / b["wall_s"]
            res["l1_ratio"] = b["l1_density"] / a["l1_density"]
        out[f"patch_{n}"] = res
    return out


def batch_l1(L, forests, seeds, lmaxs, t_end, levels, cache):
    """mean L1 density error over the instances whose finest level is in
    `levels`, each against the plain solution on its uniform finest level
    (computed once per instance and cached)"""
    errs = []
    for f, s, lm in zip(forests, seeds, lmaxs):
        if lm not in levels:
            continue
        if s not in cache:
            cache[s], _ = L.uniform_reference(s, lm, t_end)
        errs.append(L.l1_error(L.density_on_grid(f, lm), cache[s]))
    return float(np.mean(errs)), len(errs)


def batched_experiment(amrhip, args, model, plain_run, plain_forests, problem):
    """the same 64 problems run one after another on the plain path (8x8,
    every 10 steps) against the batch (plain_run: the headline's 8x8 batch
    run), per-instance final states compared bit for bit; then the batch with
    the learned tagger every --learned-interval steps.  Rates count the cells
    of instances that have not reached their stop; L1 errors are averaged
    over the instances with finest level 5-7 against their uniform plain
    solutions."""
    from amrhip import learned as L
    seeds, lmaxs = problem
    t_end = args.batch_t_end
    L.riemann_params(t_end, regrid=10)
    p = amrhip.get_params()
    insts = [amrhip.make_inst(p.coarse_grid_h, tuple(p.bc[:4]), p.cfl, p.dt, t_end) for _ in seeds]
    out = {"instances": len(seeds), "max_levels": lmaxs, "t_end": t_end}
    alone, wall_alone, cs_alone = [], 0.0, 0.0
    for fa in batch_forests(L, seeds, lmaxs, args.lmin, args.tol, 8):
        t0 = time.perf_counter()
        st = fa.evolve_device(args.tol, 10)
        wall_alone += time.perf_counter() - t0
        cs_alone += st.cell_stages
        alone.append(fa)
    out["sequential_plain_every_10"] = {"wall_s": round(wall_alone, 3), "cell_stages": cs_alone,
                                        "end_to_end_Mcell_upd_s": cs_alone / wall_alone / 1e6}
    log(f"batch: {len(seeds)} instances one by one: {wall_alone:.2f} s, {cs_alone / wall_alone / 1e6:.0f} Mcell-upd/s")
    same = [bool(a.quads() == b.quads() and np.array_equal(a.soln(), b.soln())) for a, b in zip(alone, plain_forests)]
    r = dict(plain_run)
    r["instances_bitwise_equal_to_alone"] = int(sum(same))
    out["batched_plain_every_10"] = r
    del alone
    levels = (5, 6, 7)
    cache = {}
    r["l1_density_mean_levels_5_7"], nl1 = batch_l1(L, plain_forests, seeds, lmaxs, t_end, levels, cache)
    ng = args.batch_groups
    if ng > 1:
        # the same batch split into groups on their own contexts and host
        # threads: one group's host regrid overlaps the others' steps
        fg = batch_forests(L, seeds, lmaxs, args.lmin, args.tol, 8)
        rg, _ = run_batch(amrhip, fg, insts, args.tol, 10, 8, groups=ng)
        rg["instances_bitwise_equal_to_one_context"] = int(sum(
            bool(a.quads() == b.quads() and np.array_equal(a.soln(), b.soln())) for a, b in zip(fg, plain_forests)))
        rg["speedup_vs_sequential_plain"] = rg["end_to_end_Mcell_upd_s"] / out["sequential_plain_every_10"][
            "end_to_end_Mcell_upd_s"]
        out[f"batched_plain_every_10_groups_{ng}"] = rg
        log(f"batched plain, {ng} groups: {rg['wall_s']:.2f} s, {rg['end_to_end_Mcell_upd_s']:.0f} Mcell-upd/s "
            f"end to end, {rg['instances_bitwise_equal_to_one_context']} of {len(seeds)} instances bitwise equal")
        del fg
    if model is not None:
        iv, lag = args.learned_interval, args.learned_lag
        fb = batch_forests(L, seeds, lmaxs, args.lmin, args.tol, 8)
        L.riemann_params(t_end, regrid=iv)
        rl, _ = run_batch(amrhip, fb, insts, args.tol, iv, 8, tagger=amrhip.MlpTagger(model, args.tol, True),
                          lag=lag)
        rl["l1_density_mean_levels_5_7"], _ = batch_l1(L, fb, seeds, lmaxs, t_end, levels, cache)
        out[f"batched_learned_every_{iv}_lag_{lag}"] = rl
        log(f"batched learned: {rl['wall_s']:.2f} s, {rl['end_to_end_Mcell_upd_s']:.0f} Mcell-upd/s end to end, "
            f"L1 {rl['l1_density_mean_levels_5_7']:.4e} vs plain {r['l1_density_mean_levels_5_7']:.4e}")
        rl["speedup_vs_sequential_plain"] = rl["end_to_end_Mcell_upd_s"] / out["sequential_plain_every_10"][
            "end_to_end_Mcell_upd_s"]
        rl["wall_time_ratio_sequential_plain_over_batched_learned"] = wall_alone / rl["wall_s"]
        if ng > 1:
            fb = batch_forests(L, seeds, lmaxs, args.lmin, args.tol, 8)
            rlg, _ = run_batch(amrhip, fb, insts, args.tol, iv, 8, tagger=amrhip.MlpTagger(model, args.tol, True),
                               lag=lag, groups=ng)
            rlg["l1_density_mean_levels_5_7"], _ = batch_l1(L, fb, seeds, lmaxs, t_end, levels, cache)
            rlg["speedup_vs_sequential_plain"] = rlg["end_to_end_Mcell_upd_s"] / out["sequential_plain_every_10"][
                "end_to_end_Mcell_upd_s"]
            rlg["wall_time_ratio_sequential_plain_over_batched_learned"] = wall_alone / rlg["wall_s"]
            out[f"batched_learned_every_{iv}_lag_{lag}_groups_{ng}"] = rlg
            log(f"batched learned, {ng} groups: {rlg['wall_s']:.2f} s, L1 {rlg['l1_density_mean_levels_5_7']:.4e}")
    out["l1_instances"] = nl1
    r["speedup_vs_sequential_plain"] = r["end_to_end_Mcell_upd_s"] / out["sequential_plain_every_10"][
        "end_to_end_Mcell_upd_s"]
    return out


# ---------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--lmin", type=int, default=3)
    ap.add_argument("--lmax", type=int, default=7)
    ap.add_argument("--n", type=int, default=8, help="patch size of the vortex / riemann4 workloads")
    ap.add_argument("--tol", type=float, default=0.1)
    ap.add_argument("--t-end", type=float, default=0.25)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-groups", type=int, default=2,
                    help="also run the 64-instance batch split into this many groups on their own contexts "
                         "and host threads (<= 1: not)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tile", type=str, default="", help="ppb,minw,design (skips the autotuner)")
    ap.add_argument("--workload", choices=["batch", "riemann4", "vortex"], default="batch",
                    help="batch: the 64-instance batch (headline); riemann4: the single adaptive Riemann "
                         "problem's 207k-cell mesh; vortex: uniform periodic isentropic vortex at --level")
    ap.add_argument("--level", type=int, default=4)
    ap.add_argument("--batch-instances", type=int, default=64)
    ap.add_argument("--batch-t-end", type=float, default=0.25)
    ap.add_argument("--batch-patches", type=str, default="8,16",
                    help="patch sizes the batch autotuner compares")
    ap.add_argument("--save-mesh", type=str, default="", help="batch: write the timed meshes here")
    ap.add_argument("--load-mesh", type=str, default="",
                    help="batch: time the meshes written by --save-mesh (no batch runs)")
    ap.add_argument("--load-patch", type=int, default=8, help="patch size of the meshes --load-mesh reads")
    ap.add_argument("--no-secondary", action="store_true", help="skip the 207k-cell Riemann line")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-tagger comparison")
    ap.add_argument("--no-equal-l1", action="store_true", help="skip the equal-L1 tolerance search")
    ap.add_argument("--learned-seeds", type=int, default=8, help="training runs (seeds 1..k)")
    ap.add_argument("--learned-lmax", type=int, default=9,
                    help="finest level of the deployed run (9: about 1M cells at peak)")
    ap.add_argument("--learned-interval", type=int, default=40)
    ap.add_argument("--learned-epochs", type=int, default=30)
    ap.add_argument("--learned-lag", type=int, default=10,
                    help="runs every --learned-interval steps: steps the GPU advances while the host regrids")
    ap.add_argument("--no-dmr", action="store_true", help="skip the double-Mach reflection comparison")
    ap.add_argument("--dmr-lmax", type=int, default=7)
    ap.add_argument("--dmr-t-end", type=float, default=0.2)
    ap.add_argument("--dmr-patches", type=str, default="8,16")
    ap.add_argument("--dmr-track", type=int, default=0,
                    help="only: run the uniform double-Mach forest of this level (8x8) with the state "
                         "inspected every 2048 steps, print the records as JSON and exit")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-mode comparison")
    args = ap.parse_args()
    args.batch_patches = [int(x) for x in args.batch_patches.split(",") if x]
    args.dmr_patches = [int(x) for x in args.dmr_patches.split(",") if x]
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the benchmark runs on exactly one GPU (the whole forest lives on one "
              "MI355X); sharding over several GPUs is designed but not implemented, so --gpus N "
              "with N > 1 is not run.",
              file=sys.stderr)
        sys.exit(2)

    import amrhip
    amrhip.lib()
    if args.dmr_track:
        from amrhip import learned as L
        recs, reached = L.dmr_track(args.dmr_track, args.dmr_t_end, every=2048, log=log)
        print(json.dumps({"dmr_track_level": args.dmr_track, "t_end": args.dmr_t_end, "reached": reached,
                          "records": recs}), flush=True)
        return
    tile_arg = tuple(int(x) for x in args.tile.split(",")) if args.tile else None
    batch_runs, problem, plain_forests = None, None, None
    if args.workload == "batch":
        if args.load_mesh:
            n = args.load_patch
            forests = load_forests(amrhip, args.load_mesh, args.batch_instances)
            log(f"loaded {len(forests)} meshes ({n}x{n}) from {args.load_mesh}")
        else:
            n, by_patch, batch_runs, problem = batch_headline(amrhip, args)
            forests, plain_forests = by_patch[n], by_patch.get(8)
            if args.save_mesh:
                save_forests(forests, args.save_mesh)
        arrays = amrhip.batch_external(forests)
        insts = [amrhip.make_inst(1.0, (0, 0, 0, 0), CFL_EFF, 1.0, 0.0) for _ in forests]
        nq = arrays[0].size // 3
        lv = np.concatenate([f.levels() for f in forests])
        workload = (f"{len(forests)} batched riemann4 problems (seeds 1000-{999 + len(forests)}, finest level "
                    f"from {{5,6,7,8}} at 8x8), {n}x{n} FP64 patches, SSP-RK3 CFL 0.4, meshes at t={args.batch_t_end}")
        wl_key, h_base = "batch64", 1.0
    elif args.workload == "riemann4":
        n = args.n
        f, st, _ = build_riemann_mesh(amrhip, args.seed, args.lmin, args.lmax, n, args.tol, args.t_end, 10)
        ci, fi, geom, soln = f.external()
        arrays = (ci, fi, geom, soln, None)
        insts = [amrhip.make_inst(1.0, (0, 0, 0, 0), CFL_EFF, 1.0, 0.0)]
        nq, lv = f.nquad, f.levels()
        workload = (f"riemann4 seed {args.seed}, adaptive levels {args.lmin}-{args.lmax}, {n}x{n} FP64 patches, "
                    f"SSP-RK3 CFL 0.4, mesh at t={st.t:.3f}")
        wl_key, h_base = "riemann4", 1.0
    else:
        n = args.n
        f = amrhip.Forest(1, 1, (1, 1), 0.0, 0.0, 10.0, n, args.level, args.level)
        f.apply_ic(amrhip.vortex_ic(5.0, 5.0, 5.0))
        ci, fi, geom, soln = f.external()
        arrays = (ci, fi, geom, soln, None)
        insts = [amrhip.make_inst(10.0, (0, 0, 0, 0), CFL_EFF, 1.0, 0.0)]
        nq, lv = f.nquad, f.levels()
        workload = (f"isentropic vortex strength 5, periodic [0,10]^2, uniform level {args.level}, "
                    f"{n}x{n} FP64 patches, SSP-RK3 CFL 0.4")
        wl_key, h_base = f"vortex_l{args.level}_n{n}", 10.0
    cells = nq * n * n
    log(f"timed mesh: {nq} quads / {cells} cells, levels {lv.min()}-{lv.max()}")
    g, tune, ms, stage_ms = time_steps(amrhip, n, insts, arrays, args.steps, args.warmup, tile_arg,
                                       not args.no_autotune)
    tile = g.tile()
    launch = "graph replay" if g.graphs() else "direct stream launches"
    g.close()
    ms_per_step = ms / args.steps
    value = cells * 3 * args.steps / (ms * 1e-3) / 1e6
    log(f"timed: {args.steps} steps, {ms:.2f} ms ({ms_per_step * 1e3:.1f} us/step); stage kernel "
        f"avg {[round(s * 1e3, 2) for s in stage_ms]} us; {value:.1f} Mcell-updates/s")
    try:
        probe = amrhip.copy_probe(1 << 30, 20)
    except Exception as e:   # noqa: BLE001
        log(f"copy probe failed: {e}")
        probe = None
    roof = roofline_of(cells, ms, args.steps, n, tile, launch, stage_ms)
    roof["copy_probe_GBs"] = round(probe, 1) if probe else None
    traffic = measured_traffic(wl_key, cells)
    if traffic is not None:
        roof["traffic"] = traffic[0]
        roof["traffic_tile"] = traffic[1]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "quads": nq, "cells": cells, "patch": n,
            "levels": [int(lv.min()), int(lv.max())], "parallelism": "1 GPU, every instance in one launch",
            "tile": {"patches_per_workgroup": tile[0], "min_waves_per_simd": tile[1],
                     "design": ["face task list", "paired faces", "persistent prefetch"][tile[2]],
                     "autotuned": tune is not None, "launch": launch},
        },
        "roofline": roof,
    }
    if batch_runs is not None:
        out["config"]["autotuner_patch_choice"] = {
            "rule": f"shorter batch run to t={args.batch_t_end} (tile autotuned in each run)",
            "chosen": n, "runs": {str(k): v for k, v in batch_runs.items()}}
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        lines = cpu_baseline(*arrays[:4], n, args.cpu_seconds, threads, h_base)
        if lines:
            best = max(lines, key=lambda k: lines[k]["value"])
            b = lines[best]
            out["cpu_baseline"] = {
                "value": b["value"], "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
                "sample": f"{b['steps']} SSP-RK3 steps of the same {cells}-cell mesh ({b['seconds']} s), "
                          f"C oracle {best} (faster build), {threads} threads on contiguous quadrant ranges",
                "builds": lines, "cpu_model": cpu_model(), "host_logical_cpus": os.cpu_count(),
                "threads_note": "one worker per core of the GPU's share of the host (at most 16)",
            }
            out["gpu_over_cpu"] = round(value / b["value"], 1)
    if not args.no_secondary and args.workload == "batch" and not args.load_mesh:
        out["riemann4_single"] = riemann_line(amrhip, args)
    model = None
    if not args.no_learned and not args.load_mesh:
        out["learned"], model = learned_experiment(amrhip, args)
    if not args.no_batch and plain_forests is not None:
        out["batched"] = batched_experiment(amrhip, args, model, batch_runs[8], plain_forests, problem)
    if not args.no_dmr and not args.load_mesh:
        out["double_mach"] = dmr_experiment(amrhip, args, model)
    if model is not None:
        model.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
