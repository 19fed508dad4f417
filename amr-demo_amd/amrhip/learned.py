"""Learned refinement tagger: data collection, training, recall and the
deployment comparison against the plain density-gradient indicator.

Everything runs through the library's C ABI (include/amr_tagger.h): the solver
produces the snapshots on the GPU, the network is trained on the GPU by the
library's own BF16-MFMA forward and backward kernels, and the deployed tagger
is called by the device evolve loop at each regrid.

Problem family: the four-quadrant Riemann problem on [0,1]^2 with seeded
states (riemann4_ic), 8x8 FP64 patches, SSP-RK3, effective CFL 0.3,
transmissive walls, run to a physical stop.
"""
from __future__ import annotations

import time

import numpy as np

from . import (Forest, Gpu, Mlp, MlpTagger, TAG_REFINE, Tagset, dmr_ic, dmr_inst, get_params, riemann4_ic,
               set_params, vortex_ic)

ROOT_LEN = 1 << 30


def riemann_params(t_end, regrid=10, n=8):
    set_params(nelem_subgrid=n, time_scheme=3, time_steps=10 ** 9, dt=1.0, cfl=0.25 * 0.4 * 3,
               coarse_grid_h=1.0, regrid_interval=regrid, checkpoint_interval=2 ** 31 - 1,
               visualization_interval=2 ** 31 - 1, bc=(0, 0, 0, 0, 1, 1), t_end=t_end, verbose=0)


def riemann_forest(seed, lmin, lmax, tol=0.1, n=8):
    ic = riemann4_ic(seed)
    f = Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, n, lmin, lmax)
    f.apply_ic(ic)
    f.regrid_initial(ic, tol)
    return f


def vortex_forest(lmin, lmax, tol=0.1, strength=5.0):
    """the isentropic vortex in the periodic [0,10]^2 box, adaptive"""
    ic = vortex_ic(strength=strength, xo=5.0, yo=5.0)
    f = Forest(1, 1, (1, 1), 0.0, 0.0, 10.0, 8, lmin, lmax)
    f.apply_ic(ic)
    f.regrid_initial(ic, tol)
    return f


def collect(seeds, lmin=3, lmax=7, t_end=0.25, tol=0.1, window=40, label_stride=5, vortex_runs=0,
            tagset=None, log=None):
    """labelled snapshots from seeded Riemann runs (and adaptive vortex runs)"""
    ts = tagset if tagset is not None else Tagset()
    for s in seeds:
        riemann_params(t_end)
        f = riemann_forest(s, lmin, lmax, tol)
        t0 = time.perf_counter()
        rows = f.collect(ts, tol, window=window, label_stride=label_stride)
        if log:
            log(f"collect riemann seed {s}: {rows} rows in {time.perf_counter() - t0:.2f} s")
    for k in range(vortex_runs):
        set_params(nelem_subgrid=8, time_scheme=3, time_steps=10 ** 9, dt=1.0, cfl=0.3, coarse_grid_h=10.0,
                   regrid_interval=10, checkpoint_interval=2 ** 31 - 1, visualization_interval=2 ** 31 - 1,
                   bc=(0, 0, 0, 0, 1, 1), t_end=2.0 + k, verbose=0)
        f = vortex_forest(2, 5, tol=0.02)
        rows = f.collect(ts, 0.02, window=window, label_stride=label_stride)
        if log:
            log(f"collect vortex run {k}: {rows} rows")
    return ts


def class_counts(y):
    return [int(np.sum(y == c)) for c in range(3)]


def recall_report(mlp, ts, tol=0.1):
    """tag recall of the network against the plain indicator's future flags
    (label refine), beside the indicator-now predictor on the same rows"""
    _, y, ind = ts.get(features=False)
    pred = mlp.predict(ts)
    fut = y == TAG_REFINE
    now = ind > tol
    net = pred == TAG_REFINE
    nf = max(int(fut.sum()), 1)
    rep = {
        "rows": int(y.size),
        "label_counts": class_counts(y),
        "refine_labels": int(fut.sum()),
        "recall_mlp": float(np.sum(net & fut) / nf),
        "recall_indicator_now": float(np.sum(now & fut) / nf),
        "recall_union": float(np.sum((net | now) & fut) / nf),
        "refine_fraction_mlp": float(net.mean()),
        "refine_fraction_indicator_now": float(now.mean()),
        "precision_mlp": float(np.sum(net & fut) / max(int(net.sum()), 1)),
        "precision_indicator_now": float(np.sum(now & fut) / max(int(now.sum()), 1)),
        "accuracy_3class": float(np.mean(pred == y)),
    }
    return rep


def train(ts, epochs=30, batch=4096, lr=2e-3, seed=0, cw=(1.0, 1.0, 2.0)):
    """Adam on the GPU; the refine class weighted 2x (missing a front costs
    accuracy, an extra refined patch only time)"""
    m = Mlp(seed)
    losses = m.train(ts, epochs=epochs, batch=batch, lr=lr, cw=cw, seed=seed)
    return m, losses


def recall_at_equal_fraction(mlp, ts, tol=0.1, biases=np.arange(0.0, 4.01, 0.25)):
    """the network's recall when its refine-logit bias is raised until it
    refines at least as many held-out patches as the indicator-now predictor
    (bias restored to 0 afterwards)"""
    _, y, ind = ts.get(features=False)
    frac_now = float(np.mean(ind > tol))
    out = None
    for b in biases:
        mlp.set_refine_bias(float(b))
        r = recall_report(mlp, ts, tol)
        out = {"bias": float(b), "recall_mlp": r["recall_mlp"], "refine_fraction_mlp": r["refine_fraction_mlp"],
               "recall_indicator_now": r["recall_indicator_now"], "refine_fraction_indicator_now": frac_now}
        if r["refine_fraction_mlp"] >= frac_now:
            break
    mlp.set_refine_bias(0.0)
    return out


def density_on_grid(f, level):
    """piecewise-constant density of the forest's patches sampled on the
    uniform grid of `level` quadrants (n x n cells each, n the forest's patch
    size): (N, ntx * N) array, N = n * 2**level, for a brick of ntx x 1 trees"""
    n = f.n
    N = n * (1 << level)
    out = np.zeros((N, f.ntx * N))
    q = f.soln()                         # (nquad, n, n, 4) [quad][j][i][field]
    quads = f.quads()
    lv = np.array([qq[1] for qq in quads])
    xs = np.array([qq[2] for qq in quads], np.int64)
    ys = np.array([qq[3] for qq in quads], np.int64)
    for L in np.unique(lv):
        if L > level:
            raise ValueError("forest finer than the sampling grid")
        idx = np.nonzero(lv == L)[0]
        r = 1 << (level - L)             # fine cells per patch cell
        trees = np.array([quads[i][0] for i in idx], np.int64)
        x0 = (xs[idx] >> (30 - level)) * n + trees * N
        y0 = (ys[idx] >> (30 - level)) * n
        rho = q[idx, :, :, 0]            # (k, n, n) [j][i]
        blk = np.repeat(np.repeat(rho, r, axis=1), r, axis=2)   # (k, nr, nr)
        for k in range(idx.size):
            out[y0[k]:y0[k] + n * r, x0[k]:x0[k] + n * r] = blk[k]
    return out


def uniform_reference(seed, level, t_end, log=None, n=8):
    """plain-method FP64 solution on the uniform forest of the finest level"""
    riemann_params(t_end, regrid=2 ** 31 - 1, n=n)
    ic = riemann4_ic(seed)
    f = Forest(1, 1, (0, 0), 0.0, 0.0, 1.0, n, level, level)
    f.apply_ic(ic)
    t0 = time.perf_counter()
    st = f.evolve_device(0.1, 0)
    if log:
        log(f"uniform level-{level} reference: {f.nquad * n * n} cells, {st.steps} steps, "
            f"{time.perf_counter() - t0:.1f} s")
    return density_on_grid(f, level), st


def l1_error(rho, ref):
    """mean |rho - ref| over the grid (L1 per unit area of a unit-size tree)"""
    return float(np.abs(rho - ref).mean())


# ---- double-Mach reflection (config 5): [0,4] x [0,1], four unit trees ----
def dmr_params(t_end, regrid=10, verbose=0, n=8):
    """verbose = 1: the device loop prints a progress line every 10 s"""
    set_params(nelem_subgrid=n, time_scheme=3, time_steps=10 ** 9, dt=1.0, cfl=0.25 * 0.4 * 3,
               coarse_grid_h=1.0, regrid_interval=regrid, checkpoint_interval=2 ** 31 - 1,
               visualization_interval=2 ** 31 - 1, bc=(0, 0, 0, 0, 1, 1), t_end=t_end, verbose=verbose)
    p = get_params()
    return dmr_inst(1.0, p.cfl, p.dt, t_end)


def dmr_forest(lmin, lmax, tol=0.1, n=8):
    ic = dmr_ic()
    f = Forest(4, 1, (0, 0), 0.0, 0.0, 1.0, n, lmin, lmax)
    f.apply_ic(ic)
    f.regrid_initial(ic, tol)
    return f


def dmr_uniform_reference(level, t_end, log=None, n=8):
    """the plain solution on the uniform forest of `level`, advanced on one
    device context in chunks of 16 steps (the instance freezes at t_end, as
    in the evolve loop), logging progress every 30 s"""
    inst = dmr_params(t_end, regrid=2 ** 31 - 1, n=n)
    f = Forest(4, 1, (0, 0), 0.0, 0.0, 1.0, n, level, level)
    f.apply_ic(dmr_ic())
    ci, fi, geom, soln = f.external()
    g = Gpu()
    g.setup(n, 3, [inst], ci, fi, geom, soln)
    t0 = last = time.perf_counter()
    steps = 0
    while True:
        g.run(16)
        steps += 16
        t = float(g.time()[0][0])
        if t >= t_end:
            break
        if log and time.perf_counter() - last > 30.0:
            log(f"  uniform level-{level} reference: {steps} steps, t = {t:.4f}")
            last = time.perf_counter()
    f.set_soln(np.asarray(g.soln()).reshape(-1))
    g.close()
    if log:
        log(f"uniform level-{level} double-Mach reference: {f.nquad * n * n} cells, {steps} steps, "
            f"{time.perf_counter() - t0:.1f} s")
    return density_on_grid(f, level), steps


def dmr_track(level, t_end, every=512, log=None, n=8):
    """the plain scheme on the uniform double-Mach forest of `level`, with
    the state inspected every `every` steps: time, minimum density and
    pressure, the largest flow speed and where it sits, and the speed of the
    flow behind the Mach stem for comparison.  Stops at t_end or at the
    first non-finite state (amrgpu_run's NaN check).  Returns the records
    and whether t_end was reached."""
    inst = dmr_params(t_end, regrid=2 ** 31 - 1, n=n)
    f = Forest(4, 1, (0, 0), 0.0, 0.0, 1.0, n, level, level)
    f.apply_ic(dmr_ic())
    ci, fi, geom, soln = f.external()
    g = Gpu()
    g.setup(n, 3, [inst], ci, fi, geom, soln)
    N = n << level
    h = 1.0 / N
    steps, recs, reached = 0, [], False
    while True:
        try:
            g.run(every)
        except FloatingPointError:
            recs.append({"steps": steps + every, "t": None, "nan": True})
            if log:
                log(f"  level {level}: non-finite residual within steps {steps}..{steps + every}")
            break
        steps += every
        t = float(g.time()[0][0])
        f.set_soln(np.asarray(g.soln()).reshape(-1))
        q = f.soln()
        rho = q[..., 0]
        u, v = q[..., 1] / rho, q[..., 2] / rho
        p = 0.4 * (q[..., 3] - 0.5 * rho * (u * u + v * v))
        speed = np.hypot(u, v)
        k = int(np.nanargmax(speed))
        kq, kj, ki = np.unravel_index(k, speed.shape)
        tree, lv, x, y = f.quads()[kq]
        size = 1.0 / (1 << lv)
        xs = tree + (x / ROOT_LEN) + (ki + 0.5) * size / n
        ys = (y / ROOT_LEN) + (kj + 0.5) * size / n
        rec = {"steps": steps, "t": t, "rho_min": float(np.nanmin(rho)), "p_min": float(np.nanmin(p)),
               "speed_max": float(speed.flat[k]), "at": [round(xs, 4), round(ys, 4)],
               "rho_at_speed_max": float(rho.flat[k]), "cells": f.nquad * n * n, "h": h}
        recs.append(rec)
        if log:
            log(f"  level {level} step {steps}: t={t:.4f} rho_min={rec['rho_min']:.3f} p_min={rec['p_min']:.3f} "
                f"|u|max={rec['speed_max']:.2f} at ({xs:.3f}, {ys:.4f}) rho there {rec['rho_at_speed_max']:.3f}")
        if not np.isfinite(rho).all():
            rec["nan"] = True
            break
        if t >= t_end:
            reached = True
            break
    g.close()
    return recs, reached


def dmr_collect(lmin, lmax, t_end, tol=0.1, window=40, label_stride=5, tagset=None, log=None, n=8):
    ts = tagset if tagset is not None else Tagset()
    inst = dmr_params(t_end, n=n)
    f = dmr_forest(lmin, lmax, tol, n=n)
    t0 = time.perf_counter()
    rows = f.collect(ts, tol, window=window, label_stride=label_stride, inst=inst)
    if log:
        log(f"collect double-Mach levels {lmin}-{lmax} to t={t_end}: {rows} rows in {time.perf_counter() - t0:.2f} s")
    return ts


def dmr_deploy(lmin, lmax, t_end, interval, tagger=None, tol=0.1, ref=None, ref_level=None, lag=0, verbose=0,
               n=8, tile=None):
    """ref_level: level of the sampling grid in quadrants of this run's patch
    size (the reference's grid has n * 2**ref_level cells per tree side)"""
    inst = dmr_params(t_end, regrid=interval, verbose=verbose, n=n)
    f = dmr_forest(lmin, lmax, tol, n=n)
    t0 = time.perf_counter()
    st = f.evolve_device(tol, interval, tagger=tagger, inst=inst, regrid_lag=lag, tile=tile)
    wall = time.perf_counter() - t0
    out = {
        "patch": n, "regrid_interval": interval, "regrid_lag": lag, "steps": st.steps, "regrids": st.regrids,
        "t": st.t, "peak_cells": st.peak_quads * n * n, "final_cells": f.nquad * n * n,
        "cell_stages": st.cell_stages,
        "wall_s": wall, "gpu_s": st.wall_gpu, "regrid_s": st.wall_regrid,
        "end_to_end_Mcell_upd_s": st.cell_stages / wall / 1e6,
    }
    if ref is not None:
        out["l1_density"] = l1_error(density_on_grid(f, ref_level), ref)
    if isinstance(tagger, MlpTagger):
        out["tagger"] = tagger.stats()
    return out, f


def deploy(seed, lmin, lmax, t_end, interval, tagger=None, tol=0.1, ref=None, ref_level=None, lag=0, n=8):
    """one adaptive run on the device loop; end-to-end rate and L1 error;
    lag > 0 overlaps the host regrid with `lag` steps on the old mesh"""
    riemann_params(t_end, regrid=interval, n=n)
    f = riemann_forest(seed, lmin, lmax, tol, n=n)
    t0 = time.perf_counter()
    st = f.evolve_device(tol, interval, tagger=tagger, regrid_lag=lag)
    wall = time.perf_counter() - t0
    out = {
        "regrid_interval": interval, "regrid_lag": lag, "tol": tol,
        "steps": st.steps, "regrids": st.regrids, "t": st.t,
        "peak_cells": st.peak_quads * n * n, "final_cells": f.nquad * n * n,
        "cell_stages": st.cell_stages, "wall_s": wall, "gpu_s": st.wall_gpu, "regrid_s": st.wall_regrid,
        "regrid_split_s": {"tags": st.wall_tags, "host_forest": st.wall_forest, "device_remesh": st.wall_remesh},
        "end_to_end_Mcell_upd_s": st.cell_stages / wall / 1e6,
    }
    if ref is not None:
        out["l1_density"] = l1_error(density_on_grid(f, ref_level), ref)
    if isinstance(tagger, MlpTagger):
        out["tagger"] = tagger.stats()
    return out, f
