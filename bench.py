#!/usr/bin/env python3
"""Benchmark of the batched patch-update path on one MI355X.

Workload (default): 2D shallow water radial dam break (g = 1, h = 2 inside
r < 0.2 about (0.5, 0.5), 1 outside, zero momentum, outflow boundaries) on the
AMR forest at levels 3-8 obtained by gradient tagging (|grad h| > 0.01) of the
initial data, patches mx = my = 32, mbc = 2, Roe solver with entropy fix, MC
limiter, order (2, 2), fixed dt at Courant number 0.9 on the finest level.

A step is one pass of the hot path over the whole forest: the batched update
of every leaf patch at every level, the device ghost fill and the
coarse/fine conservation fix-up (fclaw2d_advance_all_levels with
subcycle = F, time-sync = T).  The headline value is the
FP32-bulk form's cell-updates/s (sum over steps of patches x mx x my / wall
time); the FP64 form is timed the same way and reported beside it, together
with the FP32 form's L1/max error against the FP64 form and its mass change.

Prints ONE JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s, SWE AMR, FP32-bulk vs FP64 at stated L1 err; % HBM roofline"
MAX_CFL = 1.0   # fclaw_options max_cfl (src/fclaw_options.c:65)
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_cell(meqn, maux, mx, mbc=2, s=8):
    """s*[meqn*(h+1) + maux*h]: read q with its halo, read aux with its halo,
    write q; h = (mx+2mbc)^2/mx^2."""
    h = (mx + 2 * mbc) ** 2 / float(mx * mx)
    return s * (meqn * (h + 1.0) + maux * h)


def run_form(fa, prob, forest, q0, aux, precision, dt, steps, warmup, tile, use_graph,
             time_sync, dt_scale=None):
    sv = fa.Solver(prob.solver, prob.mx, prob.mx, precision=precision, mthbc=prob.mthbc,
                   grav=prob.grav, gamma=prob.gamma, tile=tile, max_patches=forest.npatch,
                   time_sync=time_sync)
    sv.set_forest(forest, prob.extent)
    sv.set_q(q0)
    if aux is not None:
        sv.set_aux(aux)
    if dt_scale is not None:
        sv.set_dt_scale(dt_scale)
    sv.ghost_fill()
    tuned = None
    if tile == 0:
        tuned = {"%d/%d/%d" % k: v for k, v in sv.autotune(dt, 8).items()}   # "tile/waves/layout"
    mass0 = sv.mass()
    if warmup > 0:
        sv.advance_async(dt, warmup)
    sv.sync()
    # kernel timing: an eager pass whose update launches carry HIP events bound
    # to each kernel's own start and end
    sv.advance_async(dt, steps)
    sv.sync()
    _, kernel_ms = sv.timing(steps)
    graph = use_graph and (steps % 2 == 0 or sv.tile == prob.mx)
    if graph:
        sv.capture_graph(dt, steps)
    sv.sync()
    t0 = time.perf_counter()
    if graph:
        sv.replay_graph()
    else:
        sv.advance_async(dt, steps)
    sv.sync()
    wall = time.perf_counter() - t0
    q = sv.get_q()
    mass1 = sv.mass()
    total_steps = warmup + 2 * steps
    return dict(wall=wall, kernel_ms=kernel_ms, q=q, mass0=mass0, mass1=mass1, tile=sv.tile,
                graph=graph, total_steps=total_steps, autotune=tuned, ghost_layout=sv.ghost_layout,
                waves=sv.waves)


def run_subcycle(fa, prob, forest, q0, aux, precision, dt_fine, ncoarse, warmup, use_graph,
                 time_sync):
    """The subcycled advance (fclaw2d_advance_all_levels with subcycle = T) on
    the same forest: level L takes 2^(L - minlevel) steps per coarse step of
    dt_fine * 2^(maxlevel - minlevel), with time interpolation, partial ghost
    fills and time syncs between the finest steps, all enqueued on the device
    (hipclaw_advance_subcycle) and timed as one captured graph of ncoarse
    coarse steps.  Cell updates count each level's own steps."""
    lv = forest.leaves()[0]
    lo, hi = int(lv.min()), int(lv.max())
    sv = fa.Solver(prob.solver, prob.mx, prob.mx, precision=precision, mthbc=prob.mthbc,
                   grav=prob.grav, gamma=prob.gamma, max_patches=forest.npatch,
                   time_sync=time_sync, subcycle=True)
    sv.set_forest(forest, prob.extent)
    sv.set_q(q0)
    if aux is not None:
        sv.set_aux(aux)
    dt = dt_fine * 2 ** (hi - lo)
    tuned = {"%d/%d/%d" % k: v for k, v in sv.autotune(dt, 1).items()}
    mass0 = sv.mass()
    if warmup > 0:
        sv.advance_subcycle_async(dt, warmup)
    sv.sync()
    if use_graph:
        sv.capture_subcycle_graph(dt, ncoarse)
        sv.sync()
    t0 = time.perf_counter()
    if use_graph:
        sv.replay_graph()
    else:
        sv.advance_subcycle_async(dt, ncoarse)
    sv.sync()
    wall = time.perf_counter() - t0
    mass1 = sv.mass()
    updates = int(np.sum(2 ** (lv.astype(np.int64) - lo))) * prob.mx * prob.mx * ncoarse
    return dict(wall=wall, updates=updates, ncoarse=ncoarse, dt=dt, tile=sv.tile, waves=sv.waves,
                ghost_layout=sv.ghost_layout, autotune=tuned, graph=use_graph,
                mass_change=float((mass1[0][0] - mass0[0][0]) / mass0[0][0]),
                fine_steps_per_coarse=2 ** (hi - lo))


def subcycle_summary(r, cells, value_nosub):
    """Throughput of the subcycled advance, and the simulated time it covers
    per second of wall time against the same forest advanced without
    subcycling (every level at the finest dt, value_nosub cell updates/s)."""
    ms_per_coarse = r["wall"] * 1e3 / r["ncoarse"]
    gain = None
    if value_nosub:
        gain = round(cells * r["fine_steps_per_coarse"] / value_nosub * 1e3 / ms_per_coarse, 3)
    return {"value": round(r["updates"] / r["wall"], 1), "unit": "cell-updates/s",
            "ms_per_coarse_step": round(ms_per_coarse, 4), "coarse_steps": r["ncoarse"],
            "cell_updates_per_coarse_step": r["updates"] // r["ncoarse"],
            "fine_steps_per_coarse_step": r["fine_steps_per_coarse"],
            "coarse_dt": r["dt"], "tile": r["tile"], "waves": r["waves"],
            "ghost_layout": r["ghost_layout"], "autotune_ms_per_coarse_step": r["autotune"],
            "timed_with_hip_graph": r["graph"], "mass_change": r["mass_change"],
            "simulated_time_per_s_vs_no_subcycle": gain}


def run_regrid(fa, prob, forest, q, repeats=5):
    """Cost of one regrid of the benchmark forest from the state q (the FP64
    form's state after the timed steps): device gradient tags, host forest
    adapt, device transfer onto the new leaves (copy / interpolate2fine /
    average2coarse) with the new ghost-fill and time-sync programs, and the
    ghost fill of every level.  The state is set up as the run loop sets it up
    (time sync on, slabs for twice the leaves) and put back before each of
    the `repeats` regrids: the first grows the tables and buffers a forest of
    that size needs (`first_ms`), the others reuse them, as the regrids of a
    run do.  Median over the repeats."""
    times = []
    run_maxlevel = prob.params.get("maxlevel") if forest.isolated else None
    sv = fa.Solver(prob.solver, prob.mx, prob.mx, precision=fa.FP64, mthbc=prob.mthbc,
                   grav=prob.grav, gamma=prob.gamma, max_patches=2 * forest.npatch, time_sync=True)
    for _ in range(repeats):
        f = fa.Forest(forest.nbx, forest.nby, forest.periodic, leaves=forest.leaves(),
                      isolated=forest.isolated)
        sv.set_forest(f, prob.extent)
        sv.set_q(q)
        sv.sync()
        t0 = time.perf_counter()
        er, ec = sv.tag(prob.tag_field, prob.tag_threshold, 0.5 * prob.tag_threshold)
        if run_maxlevel is not None:
            # an ensemble: each run refines up to its own maxlevel
            lv, gx, _ = f.leaves()
            er = er & (lv < run_maxlevel[gx >> lv]).astype(er.dtype)
        t1 = time.perf_counter()
        kind, src = f.adapt(er, (1 - ec).astype(np.int32), prob.minlevel, prob.maxlevel)
        t2 = time.perf_counter()
        parts = {}
        sv.regrid(f, kind, src, prob.extent, times=parts)
        t3 = time.perf_counter()
        sv.ghost_fill()
        sv.sync()
        t4 = time.perf_counter()
        times.append((t4 - t0, t1 - t0, t2 - t1, t4 - t2, f.npatch,
                      int(np.sum(kind == fa.REGRID_CHILD)), int(np.sum(kind == fa.REGRID_PARENT)),
                      parts["transfer"], parts["programs"], t4 - t3))
    del sv
    first = times[0][0]
    tot, tag, adapt, move, n, nchild, nparent, transfer, programs, fill = sorted(times)[len(times) // 2]
    return {"ms": round(tot * 1e3, 3), "tag_ms": round(tag * 1e3, 3),
            "adapt_ms": round(adapt * 1e3, 3), "transfer_and_ghost_fill_ms": round(move * 1e3, 3),
            "transfer_ms": round(transfer * 1e3, 3), "programs_ms": round(programs * 1e3, 3),
            "ghost_fill_ms": round(fill * 1e3, 3), "first_ms": round(first * 1e3, 3),
            "repeats": repeats,
            "patches_before": forest.npatch, "patches_after": n, "new_children": nchild,
            "new_parents": nparent, "refine_threshold": prob.tag_threshold,
            "coarsen_threshold": 0.5 * prob.tag_threshold}


def run_loop(fa, P, prob, forest, q0, aux, precision, regrid_interval, t_final=None, nsteps=None,
             stop_patches=None, t0=0.0, time_limit=None):
    """The workload as an application runs it (fclaw2d_run, src/fclaw2d_run.c:199-287)
    from the state (forest, q0): steps of all levels with dt set from the
    previous step's Courant number (desired prob.cfl), the conservation
    fix-up, and a regrid (device tags, host adapt, device transfer, ghost fill)
    every regrid_interval steps; until time t_final, nsteps accepted steps,
    or the first regrid after which the forest has stop_patches leaves.  Each
    step starts from a saved state (a device copy of every patch), and a step
    whose Courant number exceeds 1 is put back and retaken with the smaller
    dt.  One sync per step for the Courant number, as the reference's global
    reduction.  Returns the solver, the forest and the run's counters."""
    forest = fa.Forest(forest.nbx, forest.nby, forest.periodic, leaves=forest.leaves(),
                       isolated=forest.isolated)
    sv = fa.Solver(prob.solver, prob.mx, prob.mx, precision=precision, mthbc=prob.mthbc,
                   grav=prob.grav, gamma=prob.gamma, max_patches=2 * forest.npatch, time_sync=True)
    sv.set_forest(forest, prob.extent)
    sv.set_q(q0)
    if aux is not None:
        sv.set_aux(aux)
    sv.ghost_fill()
    sv.sync()
    dt = P.stable_dt(prob, forest, q0)
    t, steps, regrids, updates, t_regrid, maxp, retakes = t0, 0, 0, 0, 0.0, forest.npatch, 0
    w0 = time.perf_counter()
    while True:
        if t_final is not None and t >= t_final * (1 - 1e-12):
            break
        if nsteps is not None and steps >= nsteps:
            break
        if time_limit is not None and time.perf_counter() - w0 > time_limit:
            break
        step = dt if t_final is None else min(dt, t_final - t)
        sv.save_step()
        cfl = sv.advance(step, 1)
        if cfl > MAX_CFL:
            sv.restore_step()
            dt = step * prob.cfl / cfl
            retakes += 1
            continue
        t += step
        steps += 1
        updates += sv.npatch * prob.mx * prob.mx
        if cfl > 0:
            dt = step * prob.cfl / cfl
        if regrid_interval > 0 and steps % regrid_interval == 0 and (t_final is None or t < t_final):
            r0 = time.perf_counter()
            er, ec = sv.tag(prob.tag_field, prob.tag_threshold, 0.5 * prob.tag_threshold)
            kind, src = forest.adapt(er, (1 - ec).astype(np.int32), prob.minlevel, prob.maxlevel)
            if not (np.all(kind == fa.REGRID_SAME) and np.array_equal(src[:, 0], np.arange(len(kind)))):
                sv.regrid(forest, kind, src, prob.extent)
                if aux is not None:
                    sv.set_aux(P.initial_data(prob, forest)[1])
                sv.ghost_fill()
            sv.sync()
            t_regrid += time.perf_counter() - r0
            regrids += 1
            maxp = max(maxp, forest.npatch)
            if stop_patches is not None and forest.npatch >= stop_patches:
                break
    sv.sync()
    wall = time.perf_counter() - w0
    return sv, forest, dict(wall=wall, t=t, steps=steps, retakes=retakes, regrids=regrids,
                            updates=updates, regrid_s=t_regrid, max_patches=maxp)


def run_to(fa, P, prob, workload, t_final, regrid_interval, precision=None):
    """The whole run from t = 0 to t_final (run_loop), reported as its own line."""
    precision = fa.FP64 if precision is None else precision
    t0 = time.perf_counter()
    forest = P.build_forest(prob)
    q0, aux = P.initial_data(prob, forest)
    t_init = time.perf_counter() - t0
    sv, forest, r = run_loop(fa, P, prob, forest, q0, aux, precision, regrid_interval, t_final=t_final)
    mass = sv.mass()
    return {"metric": "time to t=%g, %s, regrid every %d steps" % (t_final, workload, regrid_interval),
            "value": round(r["wall"], 4), "unit": "s", "higher_is_better": False,
            "steps": r["steps"], "retaken_steps": r["retakes"], "regrids": r["regrids"],
            "final_patches": forest.npatch, "max_patches": r["max_patches"],
            "cell_updates": r["updates"], "cell_updates_per_s": round(r["updates"] / r["wall"], 1),
            "regrid_s": round(r["regrid_s"], 4), "init_s": round(t_init, 4),
            "final_time": r["t"], "mass": [float(m) for m in mass[0]],
            "dtype": "f64" if precision == fa.FP64 else "f32", "data": "synthetic",
            "config": {"workload": workload, "cfl": prob.cfl, "time_sync": True,
                       "refine_threshold": prob.tag_threshold,
                       "coarsen_threshold": 0.5 * prob.tag_threshold}}


GROW_PATCHES = 10000


def grow_forest(fa, P, prob, target=GROW_PATCHES, t_max=0.5, time_limit=120.0):
    """The benchmark state: the FP64 run (run_loop, regrid every 4 steps) from
    t = 0 until the first regrid after which the forest has `target` leaves
    (or t_max).  Returns the forest, its state with ghost cells filled, aux and
    a description of how it was reached."""
    forest = P.build_forest(prob)
    q0, aux = P.initial_data(prob, forest)
    n0 = forest.npatch
    sv, forest, r = run_loop(fa, P, prob, forest, q0, aux, fa.FP64, 4, t_final=t_max,
                             stop_patches=target, time_limit=time_limit)
    q = sv.get_q()
    del sv
    if aux is not None:
        aux = P.initial_data(prob, forest)[1]
    return forest, q, aux, {"grown_from_patches": n0, "patches": forest.npatch, "t": r["t"],
                            "steps": r["steps"], "regrids": r["regrids"],
                            "wall_s": round(r["wall"], 2), "target_patches": target}


def configured_run(fa, P, prob, forest, q0, aux, precision, nsteps=100, regrid_interval=4):
    """Throughput of the run as configured (regrid every regrid_interval steps)
    over nsteps accepted steps from the benchmark state: cell updates over the
    whole wall time, regrids included."""
    sv, f, r = run_loop(fa, P, prob, forest, q0, aux, precision, regrid_interval, nsteps=nsteps)
    del sv
    return {"value": round(r["updates"] / r["wall"], 1), "unit": "cell-updates/s",
            "steps": r["steps"], "retaken_steps": r["retakes"], "regrids": r["regrids"],
            "regrid_interval": regrid_interval, "wall_s": round(r["wall"], 4),
            "regrid_s": round(r["regrid_s"], 4), "patches_start": forest.npatch,
            "patches_end": f.npatch, "max_patches": r["max_patches"]}


def sweep_one(fa, P, minlevel, maxlevel, mx, args):
    """The dam break at one patch size: both forms with autotuned tiles, the
    FP32-bulk form's error against the FP64 form and both mass changes."""
    prob = P.dambreak_problem(minlevel, maxlevel, mx=mx)
    forest = P.build_forest(prob)
    q0, aux = P.initial_data(prob, forest)
    dt = P.stable_dt(prob, forest, q0)
    cells = forest.npatch * mx * mx
    r = {}
    for name, precision in (("fp64", fa.FP64), ("fp32", fa.FP32_BULK)):
        r[name] = run_form(fa, prob, forest, q0, aux, precision, dt, args.steps, args.warmup, 0,
                           not args.no_graph, not args.no_time_sync)
    _, _, dx, dy = forest.geometry(prob.extent, mx, mx)
    area = (dx * dy)[:, None, None, None]
    a, b = r["fp32"]["q"][:, :, 2:-2, 2:-2], r["fp64"]["q"][:, :, 2:-2, 2:-2]
    l1 = (np.abs(a - b) * area).sum(axis=(0, 2, 3)) / (np.abs(b) * area).sum(axis=(0, 2, 3))
    mx_err = np.abs(a - b).max(axis=(0, 2, 3)) / np.abs(b).max(axis=(0, 2, 3))
    B = algorithmic_bytes_per_cell(prob.meqn, prob.maux, mx)
    out = {"workload": "swe_dambreak_amr_L%d-%d_mx%d" % (minlevel, maxlevel, mx),
           "patches": forest.npatch, "cells": cells, "steps": args.steps}
    for name, x in r.items():
        out[name] = {"value": round(cells * args.steps / x["wall"], 1), "unit": "cell-updates/s",
                     "ms_per_step": round(x["wall"] * 1e3 / args.steps, 5), "tile": x["tile"],
                     "ghost_layout": x["ghost_layout"], "waves": x["waves"],
                     "autotune_ms_per_step": x["autotune"],
                     "kernel_avg_ms": round(x["kernel_ms"], 5),
                     "hbm_frac": round(B * cells / (x["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "mass_change": float((x["mass1"][0][0] - x["mass0"][0][0]) / x["mass0"][0][0])}
    out["fp32_l1_rel_vs_fp64"] = [float("%.3e" % v) for v in l1]
    out["fp32_max_rel_vs_fp64"] = [float("%.3e" % v) for v in mx_err]
    return out


PROFILE_STEPS = 6
COPY_BYTES = 256 << 20


def profile_run(fa, prob, forest, q0, aux, dt, tiles, time_sync, dt_scale=None):
    """The program profiled by rocprofv3 (kernel trace, then one counter per
    pass): PROFILE_STEPS steps of each form at the benchmark tiles, then three
    device-to-device copies of COPY_BYTES (the HIP runtime's copy kernel) that
    calibrate the byte counters.  Uses no torch."""
    import ctypes as C
    for precision, tile in ((fa.FP64, tiles[0]), (fa.FP32_BULK, tiles[1])):
        sv = fa.Solver(prob.solver, prob.mx, prob.mx, precision=precision, mthbc=prob.mthbc,
                       grav=prob.grav, gamma=prob.gamma, tile=tile, max_patches=forest.npatch,
                       time_sync=time_sync)
        sv.set_forest(forest, prob.extent)
        sv.set_q(q0)
        if dt_scale is not None:
            sv.set_dt_scale(dt_scale)
        sv.ghost_fill()
        sv.advance_async(dt, PROFILE_STEPS)
        sv.sync()
        print("profiled form %d tile %d: %d steps" % (precision, sv.tile, PROFILE_STEPS), flush=True)
        del sv
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(COPY_BYTES)) or hip.hipMalloc(C.byref(b), C.c_size_t(COPY_BYTES)):
        raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 0, C.c_size_t(COPY_BYTES)):
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(COPY_BYTES), 3):  # device to device
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print("calibration copies: 3 x %d bytes" % COPY_BYTES, flush=True)


def _profiled_rows(rows):
    """The dispatches of the profiled steps: those that start after the last
    regrid of the run that grew the forest (all of them for the t = 0 forest)."""
    cut = max((int(r["End_Timestamp"]) for r in rows if "regrid_remap" in r["Kernel_Name"]), default=-1)
    return [r for r in rows if int(r["Start_Timestamp"]) > cut]


def _csv_rows(d, suffix):
    import csv
    import glob as G
    rows = []
    for f in G.glob(os.path.join(d, "**", "*" + suffix), recursive=True):
        rows.extend(csv.DictReader(open(f)))
    return rows


def trace_summary(trace_dir):
    """Per-kernel dispatch count and average duration of the profiled steps of
    a rocprofv3 --kernel-trace run of --profile-run."""
    per = {}
    for r in _profiled_rows(_csv_rows(trace_dir, "kernel_trace.csv")):
        per.setdefault(r["Kernel_Name"].split("(")[0][:120], []).append(
            int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    lines = ["%-100s %6s %10s" % ("kernel", "calls", "avg_us")]
    for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        lines.append("%-100s %6d %10.2f" % (k, len(v), sum(v) / len(v) / 1e3))
    return "\n".join(lines)


def pmc_summary(fetch_dir, write_dir, cells):
    """Per-dispatch FETCH_SIZE / WRITE_SIZE (KiB) of every kernel in two
    rocprofv3 --pmc runs of --profile-run, calibrated against the copy kernel:
    factor = bytes the copy moves / bytes the counter reports for it."""
    import csv
    import glob as G

    def load(d, counter):
        out = {}
        for r in _profiled_rows(_csv_rows(d, "counter_collection.csv")):
            if r["Counter_Name"] != counter:
                continue
            out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return out

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    copy = [k for k in fetch if "copyBuffer" in k]
    if not copy:
        raise RuntimeError("no copy kernel in the counter runs")
    # the three calibration copies are the largest copy dispatches
    nf = sum(sorted(fetch[copy[0]])[-3:]) * 1024.0
    nw = sum(sorted(write[copy[0]])[-3:]) * 1024.0
    f_fetch = 3 * COPY_BYTES / nf
    f_write = 3 * COPY_BYTES / nw
    lines = ["calibration: 3 copies of %d bytes; FETCH_SIZE factor %.3f, WRITE_SIZE factor %.3f"
             % (COPY_BYTES, f_fetch, f_write)]
    for k in sorted(fetch):
        if not any(n in k for n in ("step_tiles", "ghost_", "physbc", "time_sync")):
            continue
        fk = sum(fetch[k]) / len(fetch[k])
        wk = sum(write.get(k, [0.0])) / max(len(write.get(k, [])), 1)
        traffic = fk * 1024 * f_fetch + wk * 1024 * f_write
        name = k.split("(")[0][:120]
        lines.append("kernel=%s dispatches=%d fetch_kib=%.1f write_kib=%.1f traffic_bytes=%.0f "
                     "bytes_per_cell=%.3f" % (name, len(fetch[k]), fk, wk, traffic,
                                              traffic / cells))
    return "\n".join(lines)


VALU_COUNTERS = ("SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_BUSY_CYCLES SQ_WAVE_CYCLES "
                 "SQ_WAIT_ANY SQ_LDS_BANK_CONFLICT")
CLOCK_GHZ = 2.4   # MI355X peak engine clock
SIMDS = 256 * 4


def valu_summary(pmc_dir, cells):
    """Per-dispatch instruction counters of the step kernels from one
    rocprofv3 --pmc run of --profile-run with VALU_COUNTERS: vector
    instructions per cell update and the share of the SIMDs' issue cycles
    they fill over the dispatch (wave64 VALU = 4 cycles of a 16-lane SIMD)."""
    import csv
    import glob as G
    per = {}
    span = {}
    for r in _profiled_rows(_csv_rows(pmc_dir, "counter_collection.csv")):
        k = r["Kernel_Name"]
        if "step_tiles" not in k:
            continue
        per.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
        span.setdefault(k, {})[r["Dispatch_Id"]] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    lines = ["counters: %s (one run); clock %.1f GHz, %d SIMDs" % (VALU_COUNTERS, CLOCK_GHZ, SIMDS)]
    for k in sorted(per):
        c = {n: sum(v) / len(v) for n, v in per[k].items()}
        dur_ns = sum(span[k].values()) / len(span[k])
        valu = c.get("SQ_INSTS_VALU", 0.0)
        util = valu * 4 / (SIMDS * dur_ns * CLOCK_GHZ)
        lines.append("kernel=%s dispatches=%d duration_us=%.1f waves=%.0f valu_per_cell=%.2f "
                     "valu_lane_instr_per_cell=%.0f salu_per_cell=%.2f lds_per_cell=%.2f "
                     "valu_issue_share=%.3f wait_any_share=%.3f lds_bank_conflict_per_cell=%.2f"
                     % (k.split("(")[0][:120], len(span[k]), dur_ns / 1e3, c.get("SQ_WAVES", 0),
                        valu / cells, 64 * valu / cells, c.get("SQ_INSTS_SALU", 0) / cells,
                        c.get("SQ_INSTS_LDS", 0) / cells, util,
                        c.get("SQ_WAIT_ANY", 0) / max(c.get("SQ_WAVE_CYCLES", 1), 1),
                        c.get("SQ_LDS_BANK_CONFLICT", 0) / cells))
    return "\n".join(lines)


def counter_summary(pmc_dir, cells):
    """Every counter of one rocprofv3 --pmc run of --profile-run, for each step
    kernel: the average per dispatch and per cell update; SQ_WAIT_* and
    SQ_ACTIVE_* also as a share of SQ_WAVE_CYCLES when the run holds it."""
    import csv
    import glob as G
    per = {}
    for r in _profiled_rows(_csv_rows(pmc_dir, "counter_collection.csv")):
        k = r["Kernel_Name"]
        if "step_tiles" in k:
            per.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    lines = []
    for k in sorted(per):
        c = {n: sum(v) / len(v) for n, v in per[k].items()}
        wc = c.get("SQ_WAVE_CYCLES")
        lines.append("kernel=%s" % k.split("(")[0][:120])
        for n in sorted(c):
            share = ""
            if wc and n != "SQ_WAVE_CYCLES" and (n.startswith("SQ_WAIT") or n.startswith("SQ_ACTIVE")):
                share = " share_of_wave_cycles=%.3f" % (c[n] / wc)
            lines.append("  %-26s per_dispatch=%.4g per_cell=%.3f%s" % (n, c[n], c[n] / cells, share))
    return "\n".join(lines)


def copy_peak_gbs(nbytes=1 << 30, reps=10):
    """Measured stream-copy peak: read + write bytes per second of the HIP
    runtime's device-to-device copy of nbytes, reps times, timed with HIP
    events on the null stream."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    n = C.c_size_t(nbytes)
    if hip.hipMalloc(C.byref(a), n) or hip.hipMalloc(C.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 0, n) or hip.hipEventCreate(C.byref(e0)) or hip.hipEventCreate(C.byref(e1)):
            raise RuntimeError("hipMemset / hipEventCreate failed")
        hip.hipMemcpyAsync(b, a, n, 3, None)      # warm-up, device to device
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            if hip.hipMemcpyAsync(b, a, n, 3, None):
                raise RuntimeError("hipMemcpyAsync failed")
        hip.hipEventRecord(e1, None)
        if hip.hipEventSynchronize(e1):
            raise RuntimeError("hipEventSynchronize failed")
        ms = C.c_float(0.0)
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def committed_traffic(workload, kernel_tag, cells):
    """HBM bytes per launch of the step kernel: the measured bytes per cell
    update of the committed counter summary of this workload, times the cells
    of this launch (None if there is no summary for this kernel)."""
    path = os.path.join(ROOT, "profiles", "step_tiles_%s_pmc.txt" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("kernel=") and kernel_tag in line:
            for tok in line.split():
                if tok.startswith("bytes_per_cell="):
                    return float(tok.split("=")[1]) * cells
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--problem", choices=("dambreak", "advection", "shockbubble", "ensemble"),
                    default="dambreak",
                    help="workload: shallow-water dam break (default), periodic advection of a "
                         "disk, Euler shock-bubble on a 4x1 brick, or an ensemble of --runs "
                         "independent dam breaks (one isolated block each, own dt)")
    ap.add_argument("--runs", type=int, default=64, help="ensemble size (--problem ensemble)")
    ap.add_argument("--mx", type=int, default=None, help="patch size (default 32; 16 for "
                                                          "advection and shock-bubble)")
    ap.add_argument("--minlevel", type=int, default=None)
    ap.add_argument("--maxlevel", type=int, default=None)
    ap.add_argument("--run-to", type=float, default=None, metavar="T",
                    help="instead of the step benchmark: run the workload to time T with the "
                         "CFL-controlled dt and a regrid every --regrid-interval steps, and "
                         "report the whole run (steps, regrids, time, cell updates/s)")
    ap.add_argument("--regrid-interval", type=int, default=4)
    ap.add_argument("--forest", choices=("grown", "initial"), default=None,
                    help="dam break: time the forest the run reaches when it first has %d "
                         "leaves (grown, the default) or the t = 0 forest (initial)" % GROW_PATCHES)
    ap.add_argument("--no-configured-run", action="store_true",
                    help="skip the run with its regrid cadence (configured_run)")
    ap.add_argument("--tile64", type=int, default=0, help="tile for the FP64 form (0 = autotune)")
    ap.add_argument("--tile32", type=int, default=0, help="tile for the FP32 form (0 = autotune)")
    ap.add_argument("--sweep", action="store_true",
                    help="the dam break at levels --minlevel..--maxlevel for mx = 8, 16, 32, "
                         "both forms autotuned: one JSON line per mx (no CPU baseline)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--subcycle", action="store_true",
                    help="headline the subcycled advance (each level at its own dt)")
    ap.add_argument("--no-subcycle", action="store_true",
                    help="leave the subcycled advance out of the default line")
    ap.add_argument("--subcycle-steps", type=int, default=8,
                    help="coarse steps timed for the subcycled advance")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-time-sync", action="store_true",
                    help="skip the coarse/fine conservation fix-up (on by default)")
    ap.add_argument("--profile-run", action="store_true",
                    help="run the workload's step kernels a few times for rocprofv3 (no JSON)")
    ap.add_argument("--valu-summary", metavar="PMC_DIR",
                    help="summarise a rocprofv3 --pmc run of --profile-run with VALU_COUNTERS")
    ap.add_argument("--counter-summary", metavar="PMC_DIR",
                    help="every counter of a rocprofv3 --pmc run of --profile-run, per step kernel")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs of --profile-run")
    ap.add_argument("--trace-summary", metavar="TRACE_DIR",
                    help="per-kernel averages of the profiled steps of a rocprofv3 --kernel-trace "
                         "run of --profile-run")
    ap.add_argument("--cells", type=int, default=None,
                    help="cells of the profiled forest, for the summaries (the --profile-run "
                         "output prints it)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: measures one GPU only (the batched patch update runs on a single "
              "MI355X; sharding the forest over several GPUs is not implemented)",
              file=sys.stderr)
        sys.exit(2)

    import forestclaw_amd as fa
    from forestclaw_amd import problems as P

    defaults = {"dambreak": (3, 8, 32), "advection": (4, 4, 16), "shockbubble": (2, 7, 16),
                "ensemble": (2, 7, 16)}
    dmin, dmax, dmx = defaults[args.problem]
    minlevel = dmin if args.minlevel is None else args.minlevel
    maxlevel = dmax if args.maxlevel is None else args.maxlevel
    mx = dmx if args.mx is None else args.mx
    ensemble = args.problem == "ensemble"
    if ensemble:
        # maxlevel of each run uniform in {4, ..., maxlevel}
        prob = P.ensemble_problem(args.runs, mx=mx, minlevel=minlevel,
                                  maxlevels=(min(4, maxlevel), maxlevel))
        workload = "swe_dambreak_ensemble%d_L%d-%d_mx%d" % (args.runs, minlevel, maxlevel, mx)
    else:
        make = {"dambreak": P.dambreak_problem, "advection": P.advection_problem,
                "shockbubble": P.shockbubble_problem}[args.problem]
        prob = make(minlevel, maxlevel, mx=mx)
        workload = {"dambreak": "swe_dambreak_amr", "advection": "advection_disk",
                    "shockbubble": "euler_shockbubble_amr"}[args.problem] + \
            "_L%d-%d_mx%d" % (minlevel, maxlevel, mx)
    if args.sweep:
        fa.initialize_device(0)
        for smx in (8, 16, 32):
            print(json.dumps(sweep_one(fa, P, minlevel, maxlevel, smx, args)), flush=True)
        return
    if args.run_to is not None:
        if ensemble:
            sys.exit("bench.py: --run-to is for single runs, not --problem ensemble")
        fa.initialize_device(0)
        print(json.dumps(run_to(fa, P, prob, workload, args.run_to, args.regrid_interval)))
        return
    dt_scale = None
    grown = None
    forest_kind = args.forest or ("grown" if args.problem == "dambreak" else "initial")
    if ensemble:
        forest = P.build_ensemble_forest(prob)
        q0, aux = P.initial_data(prob, forest)
        # each run at its own stable dt: step dt 1.0 times a per-slot scale
        dt_scale, dt_run = P.ensemble_dt(prob, forest, q0)
        dt = 1.0
    elif forest_kind == "grown" and (args.pmc_summary or args.valu_summary or args.counter_summary
                                     or args.trace_summary):
        if args.cells is None:
            sys.exit("bench.py: a summary of the grown forest needs --cells (printed by --profile-run)")
        forest = q0 = aux = None
        dt = 0.0
    elif forest_kind == "grown":
        fa.initialize_device(0)
        forest, q0, aux, grown = grow_forest(fa, P, prob)
        print("grown forest: %s" % json.dumps(grown), file=sys.stderr, flush=True)
        dt = P.stable_dt(prob, forest, q0)
    else:
        forest = P.build_forest(prob)
        q0, aux = P.initial_data(prob, forest)
        dt = P.stable_dt(prob, forest, q0)
    npatch = forest.npatch if forest is not None else args.cells // (prob.mx * prob.mx)
    cells = args.cells if args.cells is not None else npatch * prob.mx * prob.mx
    lv = forest.leaves()[0] if forest is not None else None
    if args.trace_summary:
        print("workload=%s cells=%d" % (workload, cells))
        print(trace_summary(args.trace_summary))
        return
    if args.pmc_summary:
        print("workload=%s cells=%d" % (workload, cells))
        print(pmc_summary(args.pmc_summary[0], args.pmc_summary[1], cells))
        return
    if args.valu_summary:
        print("workload=%s cells=%d" % (workload, cells))
        print(valu_summary(args.valu_summary, cells))
        return
    if args.counter_summary:
        print("workload=%s cells=%d" % (workload, cells))
        print(counter_summary(args.counter_summary, cells))
        return

    device = fa.initialize_device(0)
    if args.profile_run:
        print("workload=%s patches=%d cells=%d" % (workload, npatch, cells), flush=True)
        profile_run(fa, prob, forest, q0, aux, dt, (args.tile64, args.tile32), not args.no_time_sync,
                    dt_scale)
        return

    if args.subcycle:
        if ensemble:
            sys.exit("bench.py: --subcycle is for single runs, not --problem ensemble")
        r32 = run_form(fa, prob, forest, q0, aux, fa.FP32_BULK, dt, args.steps, args.warmup,
                       args.tile32, not args.no_graph, not args.no_time_sync)
        value32 = cells * args.steps / r32["wall"]
        subs = {name: run_subcycle(fa, prob, forest, q0, aux, precision, dt, args.subcycle_steps, 2,
                                   not args.no_graph, not args.no_time_sync)
                for name, precision in (("fp32", fa.FP32_BULK), ("fp64", fa.FP64))}
        s32 = subcycle_summary(subs["fp32"], cells, value32)
        s64 = subcycle_summary(subs["fp64"], cells, None)
        print(json.dumps({
            "metric": "cell-updates/s, %s, subcycled (each level at its own dt)" % workload,
            "value": s32["value"], "unit": "cell-updates/s", "n_gpus": 1,
            "steps": args.subcycle_steps, "warmup": 2, "ms_per_step": s32["ms_per_coarse_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": workload, "form": "fp32-bulk", "subcycle": True,
                       "step": "one coarse step", "patches": npatch, "cells": cells,
                       "patches_per_level": np.bincount(lv).tolist(), "device": device,
                       "time_sync": not args.no_time_sync,
                       "forest": grown if grown is not None else {"kind": "initial refinement at t = 0"}},
            "subcycle": {"fp32": s32, "fp64": s64},
            "no_subcycle_fp32": {"value": round(value32, 1), "ms_per_step": round(r32["wall"] * 1e3 / args.steps, 5)},
            "roofline": None, "cpu_baseline": None}))
        return

    r64 = run_form(fa, prob, forest, q0, aux, fa.FP64, dt, args.steps, args.warmup,
                   args.tile64, not args.no_graph, not args.no_time_sync, dt_scale)
    r32 = run_form(fa, prob, forest, q0, aux, fa.FP32_BULK, dt, args.steps, args.warmup,
                   args.tile32, not args.no_graph, not args.no_time_sync, dt_scale)
    subcycled = None
    if not args.no_subcycle and not ensemble:
        subs = {name: run_subcycle(fa, prob, forest, q0, aux, precision, dt, args.subcycle_steps, 2,
                                   not args.no_graph, not args.no_time_sync)
                for name, precision in (("fp32", fa.FP32_BULK), ("fp64", fa.FP64))}
        nosub = {"fp32": cells * args.steps / r32["wall"], "fp64": cells * args.steps / r64["wall"]}
        subcycled = {name: subcycle_summary(r, cells, nosub[name]) for name, r in subs.items()}

    regrid = run_regrid(fa, prob, forest, r64["q"])
    configured = None
    if not args.no_configured_run and not ensemble:
        configured = {name: configured_run(fa, P, prob, forest, q0, aux, precision)
                      for name, precision in (("fp32", fa.FP32_BULK), ("fp64", fa.FP64))}

    B = algorithmic_bytes_per_cell(prob.meqn, prob.maux, prob.mx)
    value32 = cells * args.steps / r32["wall"]
    value64 = cells * args.steps / r64["wall"]

    copy_peak = copy_peak_gbs()

    def roofline(r, rtype):
        achieved = B * cells / (r["kernel_ms"] * 1e-3) / 1e9
        traffic = committed_traffic(workload, "%s, %d," % (rtype, r["tile"]), cells)
        return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None if traffic is None else round(traffic),
                "kernel": "step_tiles", "kernel_avg_ms": round(r["kernel_ms"], 5),
                "bytes_per_cell_update": round(B, 4), "cell_updates_per_launch": cells,
                "copy_peak": round(copy_peak, 1), "frac_of_copy_peak": round(achieved / copy_peak, 4)}

    # accuracy of the FP32-bulk form against the FP64 form after the same steps
    _, _, dxs, dys = forest.geometry(prob.extent, prob.mx, prob.mx)
    area = (dxs * dys)[:, None, None, None]
    i64 = r64["q"][:, :, 2:-2, 2:-2]
    i32 = r32["q"][:, :, 2:-2, 2:-2]
    l1 = (np.abs(i32 - i64) * area).sum(axis=(0, 2, 3)) / (np.abs(i64) * area).sum(axis=(0, 2, 3))
    maxerr = np.abs(i32 - i64).max(axis=(0, 2, 3)) / np.abs(i64).max(axis=(0, 2, 3))
    mass_change = lambda r: float((r["mass1"][0][0] - r["mass0"][0][0]) / r["mass0"][0][0])
    runs = None
    if ensemble:
        # per-run accuracy of the FP32-bulk form (largest relative L1 error over the fields)
        lvs, gxs, _ = forest.leaves()
        blk = gxs >> lvs
        l1_run = []
        for b in range(prob.nbx):
            sel = blk == b
            num = (np.abs(i32[sel] - i64[sel]) * area[sel]).sum(axis=(0, 2, 3))
            den = (np.abs(i64[sel]) * area[sel]).sum(axis=(0, 2, 3))
            l1_run.append(float((num / den).max()))
        per = np.bincount(blk, minlength=prob.nbx)
        runs = {"runs": prob.nbx, "seed": prob.params["seed"],
                "patches_per_run": {"min": int(per.min()), "max": int(per.max())},
                "maxlevel_per_run": np.bincount(prob.params["maxlevel"]).tolist(),
                "dt_per_run": {"min": float(dt_run.min()), "max": float(dt_run.max())},
                "l1_rel_vs_fp64_per_run": {"max": float("%.3e" % max(l1_run)),
                                           "median": float("%.3e" % np.median(l1_run))}}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import cpu_baseline
        idx = np.linspace(0, npatch - 1, min(64, npatch)).astype(int)
        xl, yl, dx, dy = forest.geometry(prob.extent, prob.mx, prob.mx)
        cpu = cpu_baseline.measure(prob.solver, q0[idx], None if aux is None else aux[idx],
                                   dx[idx], dt if dt_scale is None else float(dt_scale.min()), prob.meqn,
                                   prob.maux, grav=prob.grav, target_seconds=args.cpu_seconds)

    out = {
        "metric": METRIC if args.problem == "dambreak" else "cell-updates/s, %s, FP32-bulk vs FP64" % workload,
        "value": round(value32, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["wall"] * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "form": "fp32-bulk (f32 Riemann/limiter/transverse; f64 q, fluxes, CFL, mass)",
            "patches": npatch, "patches_per_level": np.bincount(lv).tolist(),
            "cells": cells, "mx": prob.mx, "mbc": 2, "meqn": prob.meqn,
            "dt": dt, "tile_fp32": r32["tile"], "tile_fp64": r64["tile"],
            "ghost_layout_fp32": r32["ghost_layout"], "ghost_layout_fp64": r64["ghost_layout"],
            "waves_fp32": r32["waves"], "waves_fp64": r64["waves"],
            "autotune_ms_per_step": {"fp32": r32["autotune"], "fp64": r64["autotune"]},
            "timed_with_hip_graph": bool(r32["graph"]), "time_sync": not args.no_time_sync,
            "device": device,
            "forest": grown if grown is not None else {"kind": "initial refinement at t = 0"},
        },
        "roofline": roofline(r32, "RPShallow<float>, float"),
        "fp64": {
            "value": round(value64, 1), "unit": "cell-updates/s",
            "ms_per_step": round(r64["wall"] * 1e3 / args.steps, 5),
            "roofline": roofline(r64, "RPShallow<double>, double"),
        },
        "fp32_over_fp64": round(value32 / value64, 3),
        "accuracy": {
            "l1_rel_vs_fp64": [float("%.3e" % v) for v in l1],
            "max_rel_vs_fp64": [float("%.3e" % v) for v in maxerr],
            "steps_compared": r32["total_steps"],
            "mass_change_fp32": mass_change(r32),
            "mass_change_fp64": mass_change(r64),
        },
        "regrid": regrid,
        "configured_run": configured,
        "subcycle": subcycled,
        "cpu_baseline": cpu,
    }
    if runs is not None:
        out["ensemble"] = runs
        out["config"]["dt"] = "per run (ensemble.dt_per_run)"
    print(json.dumps(out))


if __name__ == "__main__":
    main()
