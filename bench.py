"""Benchmark: FP64 cell-updates per second per RK stage on one MI355X.

Workload: 3D Taylor-Green vortex, uniform 256^3 periodic box [0,2pi]^3 split
into 64 boxes of 64^3, rho0=1, Ma=0.1, Re=1600, Pr=0.71, scaled Sutherland
viscosity, KEEP4 convective flux + 2nd-order viscous flux, SSPRK(3,3), FP64,
fixed dt = 5e-5.  One "step" = one SSPRK(3,3) step = 3 x (ghost fill + fused
RHS/stage kernel).  The state is resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints ONE JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s
N = 256
BOX = 64
DT = 5.0e-5
# algorithmic HBM bytes per cell for the three SSPRK(3,3) stages of a fused
# stage kernel: stage 1 reads U^n and writes U1 (80 B); stages 2 and 3 read
# U^(s), U^n and write (120 B).
ALG_BYTES = [80.0, 120.0, 120.0]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run plain device copies of this many bytes (profiler calibration)")
    ap.add_argument("--tiling", type=str, default="auto",
                    help="fused-kernel tiling: auto (autotune), default, or TX,TY,TZ")
    ap.add_argument("--traffic", type=str, default=os.path.join(ROOT, "profiles", "traffic_tgv256_keep4v.json"))
    return ap.parse_args()


def cpu_baseline():
    """The CPU oracle (a restatement of the reference's unfused passes) on a
    bounded sample of the same workload: TGV 128^3 in 8 boxes of 64^3, same
    closure and scheme, 3 SSPRK(3,3) steps, one thread."""
    from amrsolver_amd import problems as P
    from oracle import oracle as O
    n = (128, 128, 128)
    cls = P.tgv_closures()
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    ol = O.OracleLevel(g, P.decompose(n, BOX))
    ol.set_state(P.tgv_state(cls, g, n))
    steps = 3
    t0 = time.perf_counter()
    for _ in range(steps):
        ol.advance(cls, sch, DT, nthreads=1)
    el = time.perf_counter() - t0
    return {"value": n[0] * n[1] * n[2] * 3 * steps / el, "unit": "cell-stage/s", "cores": 1, "kind": "port",
            "sample": f"TGV 128^3 (8 boxes of 64^3), KEEP4+viscous, {steps} SSPRK(3,3) steps, 1 thread, "
                      f"{el:.1f} s"}


def copy_calibration(nbytes):
    """Plain device-to-device copies of a known size, for scaling the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) of a profiler run."""
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(4):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 4}))
    return 0


def main():
    a = parse()
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        return 2
    import torch
    from amrsolver_amd import capi
    from amrsolver_amd import level as LV
    from amrsolver_amd import problems as P

    n = (a.n, a.n, a.n)
    cls = P.tgv_closures()
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    boxes = P.decompose(n, BOX)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    ctx = LV.Context(0)
    lev = LV.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout="level")
    W = [lev.new_state(layout="level") for _ in range(2)]
    S.upload(P.tgv_state(cls, g, n))
    ncell = n[0] * n[1] * n[2]
    stream = ctx.stream
    coefs = [capi.stage_coeffs(3, 3, DT, s)[1:] for s in range(3)]

    def step(events=None):
        srcs = [S, W[0], W[1]]
        outs = [W[0], W[1], S]
        for s in range(3):
            lev.fill_patch(srcs[s])
            if events is not None:
                events[s][0].record(stream)
            mode, ca, cb, cc, cd = coefs[s]
            lev.rk_stage(mode, ca, cb, cc, cd, S, srcs[s], outs[s])
            if events is not None:
                events[s][1].record(stream)

    lev.fill_patch(S)
    if a.tiling == "auto":
        tiling = lev.autotune(S)
    elif a.tiling == "default":
        tiling = None
    else:
        tiling = tuple(int(x) for x in a.tiling.split(","))
        lev.set_tiling(*tiling)
    t0 = np.array(lev.totals(S))
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    ev = [[[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(3)] for _ in range(a.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(a.steps):
        step(ev[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    # per-stage kernel durations
    dur = np.zeros(3)
    for k in range(a.steps):
        for s in range(3):
            dur[s] += ev[k][s][0].elapsed_time(ev[k][s][1]) * 1e-3
    dur /= a.steps
    t1 = np.array(lev.totals(S))
    drift = [abs(t1[c] - t0[c]) / abs(t0[c]) for c in (0, 4)]

    ms_per_step = elapsed / a.steps * 1e3
    value = ncell * 3 * a.steps / elapsed
    # roofline of the dominant kernel: the fused stage kernel, bytes averaged
    # over the three stages (106.7 B per cell-stage)
    alg_bytes = sum(ALG_BYTES) / 3 * ncell
    mean_dur = float(dur.mean())
    achieved = alg_bytes / mean_dur / 1e9
    traffic = None
    if os.path.exists(a.traffic):
        try:
            traffic = json.load(open(a.traffic)).get("bytes_per_launch")
        except Exception:
            traffic = None
    out = {
        "metric": "FP64 cell-updates/sec per RK stage",
        "value": value,
        "unit": "cell-stage/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic TGV initial state)",
        "config": {"workload": f"TGV {n[0]}^3 periodic, {len(boxes)} boxes of {BOX}^3, KEEP4+viscous, "
                               f"SSPRK(3,3), dt={DT}", "cells": ncell, "stages_per_step": 3},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "stage_kernel_ms": [float(x * 1e3) for x in dur],
        "tiling": tiling,
        "totals_drift": {"mass": drift[0], "energy": drift[1]},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
