#!/usr/bin/env python3
"""Benchmark of the batched patch-update path on one MI355X.

Workload (default): 2D shallow water radial dam break (g = 1, h = 2 inside
r < 0.2 about (0.5, 0.5), 1 outside, zero momentum, outflow boundaries) on the
AMR forest at levels 3-8 obtained by gradient tagging (|grad h| > 0.01) of the
initial data, patches mx = my = 32, mbc = 2, Roe solver with entropy fix, MC
limiter, order (2, 2), fixed dt at Courant number 0.9 on the finest level.

A step is one pass of the hot path over the whole forest: the batched update
of every leaf patch at every level followed by the device ghost fill
(fclaw2d_advance_all_levels with subcycle = F).  The headline value is the
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
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_cell(meqn, maux, mx, mbc=2, s=8):
    """s*[meqn*(h+1) + maux*h]: read q with its halo, read aux with its halo,
    write q; h = (mx+2mbc)^2/mx^2."""
    h = (mx + 2 * mbc) ** 2 / float(mx * mx)
    return s * (meqn * (h + 1.0) + maux * h)


def run_form(fa, prob, forest, q0, aux, precision, dt, steps, warmup, tile, use_graph):
    sv = fa.Solver(prob.solver, prob.mx, prob.mx, precision=precision, mthbc=prob.mthbc,
                   grav=prob.grav, gamma=prob.gamma, tile=tile, max_patches=forest.npatch)
    sv.set_forest(forest, prob.extent)
    sv.set_q(q0)
    if aux is not None:
        sv.set_aux(aux)
    sv.ghost_fill()
    mass0 = sv.mass()
    if warmup > 0:
        sv.advance_async(dt, warmup)
    sv.sync()
    # kernel timing: an eager pass with HIP events around every update launch
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
                graph=graph, total_steps=total_steps)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mx", type=int, default=32)
    ap.add_argument("--minlevel", type=int, default=3)
    ap.add_argument("--maxlevel", type=int, default=8)
    ap.add_argument("--tile64", type=int, default=0, help="tile for the FP64 form (0 = tuned)")
    ap.add_argument("--tile32", type=int, default=0, help="tile for the FP32 form (0 = tuned)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (the batched patch update runs on "
              "a single MI355X; multi-GPU sharding is described in DESIGN.md, not run)",
              file=sys.stderr)
        sys.exit(2)

    import forestclaw_amd as fa
    from forestclaw_amd import problems as P

    device = fa.initialize_device(0)
    prob = P.dambreak_problem(args.minlevel, args.maxlevel, mx=args.mx)
    forest = P.build_forest(prob)
    q0, aux = P.initial_data(prob, forest)
    dt = P.stable_dt(prob, forest, q0)
    npatch = forest.npatch
    cells = npatch * prob.mx * prob.mx
    lv = forest.leaves()[0]

    r64 = run_form(fa, prob, forest, q0, aux, fa.FP64, dt, args.steps, args.warmup,
                   args.tile64, not args.no_graph)
    r32 = run_form(fa, prob, forest, q0, aux, fa.FP32_BULK, dt, args.steps, args.warmup,
                   args.tile32, not args.no_graph)

    B = algorithmic_bytes_per_cell(prob.meqn, prob.maux, prob.mx)
    value32 = cells * args.steps / r32["wall"]
    value64 = cells * args.steps / r64["wall"]

    def roofline(r):
        achieved = B * cells / (r["kernel_ms"] * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "step_tiles", "kernel_avg_ms": round(r["kernel_ms"], 5),
                "bytes_per_cell_update": round(B, 4), "cell_updates_per_launch": cells}

    # accuracy of the FP32-bulk form against the FP64 form after the same steps
    _, _, dxs, dys = forest.geometry(prob.extent, prob.mx, prob.mx)
    area = (dxs * dys)[:, None, None, None]
    i64 = r64["q"][:, :, 2:-2, 2:-2]
    i32 = r32["q"][:, :, 2:-2, 2:-2]
    l1 = (np.abs(i32 - i64) * area).sum(axis=(0, 2, 3)) / (np.abs(i64) * area).sum(axis=(0, 2, 3))
    maxerr = np.abs(i32 - i64).max(axis=(0, 2, 3)) / np.abs(i64).max(axis=(0, 2, 3))
    mass_change = lambda r: float((r["mass1"][0][0] - r["mass0"][0][0]) / r["mass0"][0][0])

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import cpu_baseline
        idx = np.linspace(0, npatch - 1, min(64, npatch)).astype(int)
        xl, yl, dx, dy = forest.geometry(prob.extent, prob.mx, prob.mx)
        cpu = cpu_baseline.measure(prob.solver, q0[idx], None, dx[idx], dt, prob.meqn,
                                   prob.maux, grav=prob.grav, target_seconds=args.cpu_seconds)

    out = {
        "metric": METRIC,
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
            "workload": "swe_dambreak_amr_L%d-%d_mx%d" % (args.minlevel, args.maxlevel, prob.mx),
            "form": "fp32-bulk (f32 Riemann/limiter/transverse; f64 q, fluxes, CFL, mass)",
            "patches": npatch, "patches_per_level": np.bincount(lv).tolist(),
            "cells": cells, "mx": prob.mx, "mbc": 2, "meqn": prob.meqn,
            "dt": dt, "tile_fp32": r32["tile"], "tile_fp64": r64["tile"],
            "timed_with_hip_graph": bool(r32["graph"]), "device": device,
        },
        "roofline": roofline(r32),
        "fp64": {
            "value": round(value64, 1), "unit": "cell-updates/s",
            "ms_per_step": round(r64["wall"] * 1e3 / args.steps, 5),
            "roofline": roofline(r64),
        },
        "fp32_over_fp64": round(value32 / value64, 3),
        "accuracy": {
            "l1_rel_vs_fp64": [float("%.3e" % v) for v in l1],
            "max_rel_vs_fp64": [float("%.3e" % v) for v in maxerr],
            "steps_compared": r32["total_steps"],
            "mass_change_fp32": mass_change(r32),
            "mass_change_fp64": mass_change(r64),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
