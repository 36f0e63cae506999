"""Benchmark: FP64 cell-updates per second per RK stage on one MI355X.

Workload: 3D Taylor-Green vortex, uniform 256^3 periodic box [0,2pi]^3 split
into 64 boxes of 64^3, rho0=1, Ma=0.1, Re=1600, Pr=0.71, scaled Sutherland
viscosity, KEEP4 convective flux + 2nd-order viscous flux, SSPRK(3,3), FP64,
fixed dt = 5e-5.  One "step" = one SSPRK(3,3) step through amr_advance (the
CNS::advance replacement): three fused RHS/stage kernels; the periodic level
is one array, so the kernels read halo cells from their periodic images and
no ghost fill runs.  The state is resident in HBM before timing starts.

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
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (fused multiply-add counted as two)
# FP64 operations per cell-stage of the Sod workloads (4 boxes of 64x4x4,
# SSPRK(3,3), the Sod initial state; MUSCL + Riemann depends on the data
# through its branches), counted as the TGV constant is; the CPU test suite
# recounts both
FP64_OPS_SOD_W = 5222.2
FP64_OPS_SOD_R = 1307.0
# FP64 operations of the reference arithmetic per cell-stage for this
# workload (KEEP4 + Sutherland viscous, SSPRK(3,3), 64^3 boxes), counted by
# running the CPU restatement on an operation-counting scalar; the CPU test
# suite recounts it and checks this constant
FP64_OPS_PER_CELL_STAGE = 754.0
# the same count for KEEP6 + viscous
FP64_OPS_KEEP6V = 1282.0
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
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run plain device copies of this many bytes (profiler calibration)")
    ap.add_argument("--workload", type=str, default="tgv",
                    choices=["tgv", "sod-w", "sod-r", "riemann-batch-w", "riemann-batch-r", "adjoint", "amr",
                             "amr-hierarchy", "channel"])
    ap.add_argument("--tiling", type=str, default="auto",
                    help="fused-kernel tiling: auto (autotune), default, or TX,TY,TZ[,FORM] "
                         "(FORM 1: the self-loading form of 16x16 tiles)")
    ap.add_argument("--traffic", type=str, default=os.path.join(ROOT, "profiles", "traffic_tgv256_keep4v.json"))
    ap.add_argument("--keep", type=int, default=4, choices=[4, 6],
                    help="KEEP order of the TGV workload (the headline is order 4)")
    return ap.parse_args()


def box_layout_timing(a, ctx, cls, g, boxes, sch, level_ms):
    """The same TGV steps with one ghosted array per 64^3 box (the layout an
    AMReX MultiFab hands over), autotuned: whole amr_advance steps timed as
    the headline is; level_ms is the headline's time per step."""
    import torch
    from amrsolver_amd import level as LV
    from amrsolver_amd import problems as P
    lev = LV.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout="boxes")
    for b, (lo, hi) in enumerate(lev.boxes):
        S.upload_box(b, P.tgv_state_box(cls, g, lo, hi, 3))
    lev.fill_patch(S)
    tiling = lev.autotune(S)
    plan = lev.step_plan(S)
    for _ in range(a.warmup):
        lev.advance(S, 0.0, DT)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        lev.advance(S, 0.0, DT)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ncell = sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for lo, hi in lev.boxes)
    ms = el / a.steps * 1e3
    lev.close()
    del S
    torch.cuda.empty_cache()
    return {"value": ncell * 3 * a.steps / el, "unit": "cell-stage/s", "ms_per_step": ms,
            "vs_level_layout": ms / level_ms, "tiling": list(tiling), "tiles": plan["tiles"],
            "direct_neighbours": plan["direct_neighbours"]}


def cpu_baseline(keep=4):
    """The CPU oracle (a restatement of the reference's unfused passes) on a
    bounded sample of the same workload: TGV 128^3 in 8 boxes of 64^3, same
    closure and scheme, 12 SSPRK(3,3) steps (about 10 s at order 4), one
    thread."""
    from amrsolver_amd import problems as P
    from oracle import oracle as O
    n = (128, 128, 128)
    cls = P.tgv_closures()
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    sch = P.scheme(flux_euler=1, order_keep=keep, visc=True, order_rk=3, stages_rk=3)
    ol = O.OracleLevel(g, P.decompose(n, BOX))
    ol.set_state(P.tgv_state(cls, g, n))
    steps = 12 if keep == 4 else 8
    t0 = time.perf_counter()
    for _ in range(steps):
        ol.advance(cls, sch, DT, nthreads=1)
    el = time.perf_counter() - t0
    return {"value": n[0] * n[1] * n[2] * 3 * steps / el, "unit": "cell-stage/s", "cores": 1, "kind": "port",
            "sample": f"TGV 128^3 (8 boxes of 64^3), KEEP{keep}+viscous, {steps} SSPRK(3,3) steps, 1 thread, "
                      f"{el:.1f} s"}


def cpu_extra(keep=4):
    """The same oracle on all host cores the process may use (OpenMP over
    boxes, TGV 256^3 in 64 boxes of 64^3, 2 steps) and, one thread, built for
    x86-64-v3 (TGV 128^3, 6 steps); plus the host's CPU model."""
    from amrsolver_amd import problems as P
    from oracle import oracle as O
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=keep, visc=True, order_rk=3, stages_rk=3)
    threads = _host_threads()
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out = {"cpu_model": model, "threads_available": threads}
    n = (256, 256, 256)
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    ol = O.OracleLevel(g, P.decompose(n, BOX))
    ol.set_state(P.tgv_state(cls, g, n))
    t0 = time.perf_counter()
    for _ in range(2):
        ol.advance(cls, sch, DT, nthreads=threads)
    el = time.perf_counter() - t0
    out["all_cores"] = {"value": n[0] * n[1] * n[2] * 3 * 2 / el, "unit": "cell-stage/s", "cores": threads,
                        "sample": f"TGV 256^3 (64 boxes of 64^3), 2 SSPRK(3,3) steps, {el:.1f} s"}
    del ol
    n = (128, 128, 128)
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    ol = O.OracleLevel(g, P.decompose(n, BOX), lib=O.load("x86v3"))
    ol.set_state(P.tgv_state(cls, g, n))
    t0 = time.perf_counter()
    for _ in range(6):
        ol.advance(cls, sch, DT, nthreads=1)
    el = time.perf_counter() - t0
    out["x86_64_v3_one_thread"] = {"value": n[0] * n[1] * n[2] * 3 * 6 / el, "unit": "cell-stage/s", "cores": 1,
                                   "sample": f"TGV 128^3 (8 boxes of 64^3), 6 SSPRK(3,3) steps, {el:.1f} s"}
    return out


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


def base_line(metric_value, elapsed_ms_per_step, steps, warmup, workload, extra):
    out = {
        "metric": "FP64 cell-updates/sec per RK stage",
        "value": metric_value,
        "unit": "cell-stage/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": elapsed_ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload},
    }
    out.update(extra)
    return out


def _fp64(ops, rate):
    """Algorithmic FP64 rate (ops per cell-stage times cell-stages/s)
    against the FP64 vector peak."""
    ach = ops * rate / 1e12
    return {"ops_per_cell_stage": ops, "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / FP64_PEAK_TFLOPS}


def _oracle_cfl_run(cls, sch, g, boxes, u, stop, nthreads):
    """The CPU oracle to `stop` with the CFL 0.3 policy (computeNewDt on the
    host): (steps, seconds)."""
    from amrsolver_amd import level as LV
    from oracle import oracle as O
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    t, dtp, k = 0.0, None, 0
    t0 = time.perf_counter()
    while t < stop - 1e-14:
        d = LV.compute_new_dt(ol.est_dt(cls, 0.3), dtp, t, stop, dtp is None)
        dtp = d
        ol.advance(cls, sch, d, nthreads=nthreads)
        t += d
        k += 1
    return k, time.perf_counter() - t0


def _host_threads():
    """Host threads for the all-core CPU leg: the cores this process may
    run on, at most 16 (one GPU's share of a shared host)."""
    import os
    return min(16, len(os.sched_getaffinity(0)))


def run_sod(a, flux):
    """Sod tube 256x4x4 (4 boxes of 64x4x4), outflow in x, periodic y/z,
    SSPRK(3,3), CFL 0.3 with the 1.1x growth cap, to t = 0.2, with the time
    step chosen on the device (amr_advance_cfl_graph: captured chunks of
    steps, the clock read back once per chunk)."""
    import torch
    from amrsolver_amd import level as LV
    from amrsolver_amd import problems as P
    n = (256, 4, 4)
    cls = P.closures_reference()
    g = P.geom(n, (0, 0, 0), (1.0, 4 / 256, 4 / 256), (0, 1, 1), lo_bc=(2, 0, 0), hi_bc=(2, 0, 0))
    boxes = P.decompose(n, (64, 4, 4))
    sch = P.scheme(flux_euler=flux, visc=False)
    lev = LV.Level(g, boxes, cls, sch)
    S = lev.new_state(layout="boxes")
    u0 = P.sod_state(cls, g, n)
    chunk = 32

    def run():
        S.upload(u0)
        lev.clock_set(0.0)
        while True:
            lev.advance_cfl_graph(S, 0.3, 0.2, chunk)
            t, _, k = lev.clock_get()
            if t[0] >= 0.2 - 1e-14:
                break
        torch.cuda.synchronize()
        return k[0]

    for _ in range(max(1, a.warmup)):
        run()
    t0 = time.perf_counter()
    k = run()
    el = time.perf_counter() - t0
    rate = n[0] * n[1] * n[2] * 3 * k / el
    name = "Sod-W (WENO5-JS + LF)" if flux == 2 else "Sod-R (MUSCL + Riemann)"
    ops = FP64_OPS_SOD_W if flux == 2 else FP64_OPS_SOD_R
    extra = {"data": "synthetic (Sod initial state)", "fp64": _fp64(ops, rate)}
    if not a.no_cpu:
        # the whole run on the CPU oracle: one thread, then all host threads
        ncell = n[0] * n[1] * n[2]
        k1, e1 = _oracle_cfl_run(cls, sch, g, boxes, u0, 0.2, 1)
        extra["cpu_baseline"] = {"value": ncell * 3 * k1 / e1, "unit": "cell-stage/s", "cores": 1, "kind": "port",
                                 "sample": f"the whole run to t=0.2 ({k1} steps), 1 thread, {e1:.1f} s"}
        nt = _host_threads()
        k2, e2 = _oracle_cfl_run(cls, sch, g, boxes, u0, 0.2, nt)
        extra["cpu_all_cores"] = {"value": ncell * 3 * k2 / e2, "unit": "cell-stage/s", "cores": nt,
                                  "sample": f"the whole run to t=0.2 ({k2} steps), {nt} threads, {e2:.1f} s"}
    return base_line(rate, el / k * 1e3, k, a.warmup,
                     f"{name} 256x4x4, 4 boxes, SSPRK(3,3), CFL 0.3 on the device, t=0.2", extra)


def run_riemann_batch(a, flux):
    """512 independent 1D Riemann problems, L uniform in [64, 2048] cells (x4x4),
    seed 20261202, SSPRK(3,3), CFL 0.3 per instance, to t = 0.15; one launch
    per kernel for all instances, each problem's time step chosen on the
    device (amr_advance_cfl_graph, the clocks read back once per chunk)."""
    import torch
    from amrsolver_amd import level as LV
    from amrsolver_amd import problems as P
    cls = P.closures_reference()
    inst = P.riemann_batch_params()
    geoms, states = P.riemann_batch_problems(cls, inst)
    sch = P.scheme(flux_euler=flux, visc=False)
    bat = LV.Batch(geoms, cls, sch)
    S = bat.new_state()
    cells = [p["L"] * 16 for p in inst]
    stop = 0.15
    chunk = 32

    def run():
        S.upload_problems(states)
        bat.clock_set(0.0)
        while True:
            bat.advance_cfl_graph(S, 0.3, stop, chunk)
            t, _, k = bat.clock_get()
            if min(t) >= stop - 1e-14:
                break
        torch.cuda.synchronize()
        return k

    for _ in range(max(1, a.warmup)):
        run()
    t0 = time.perf_counter()
    k = run()
    el = time.perf_counter() - t0
    done = sum(c * 3 * kb for c, kb in zip(cells, k))
    name = "WENO5-JS + LF" if flux == 2 else "MUSCL + Riemann"
    ops = FP64_OPS_SOD_W if flux == 2 else FP64_OPS_SOD_R
    extra = {"data": "synthetic (seed 20261202, SplitMix64)", "cell_stages": done, "batch_steps": max(k),
             "fp64": _fp64(ops, done / el),
             "fp64_note": "ops per cell-stage counted on the Sod state with 64x4x4 boxes"}
    if not a.no_cpu:
        # the instances with L <= 160 solo on the CPU oracle: one after another
        # on one thread, then side by side on all host threads
        import concurrent.futures as cf
        sample = [b for b, p in enumerate(inst) if p["L"] <= 160]

        def solo(b):
            g = geoms[b]
            kk, ee = _oracle_cfl_run(cls, sch, g, [(tuple(g.dom_lo), tuple(g.dom_hi))], states[b], stop, 1)
            return cells[b] * 3 * kk, ee

        t0 = time.perf_counter()
        res = [solo(b) for b in sample]
        e1 = time.perf_counter() - t0
        work = sum(w for w, _ in res)
        extra["cpu_baseline"] = {"value": work / e1, "unit": "cell-stage/s", "cores": 1, "kind": "port",
                                 "sample": f"{len(sample)} instances with L <= 160 to t=0.15, one after another, "
                                           f"1 thread, {e1:.1f} s"}
        nt = _host_threads()
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=nt) as pool:
            list(pool.map(solo, sample))
        e2 = time.perf_counter() - t0
        extra["cpu_all_cores"] = {"value": work / e2, "unit": "cell-stage/s", "cores": nt,
                                  "sample": f"the same {len(sample)} instances side by side, {nt} threads, "
                                            f"{e2:.1f} s"}
    return base_line(done / el, el / max(k) * 1e3, max(k), a.warmup,
                     f"batched 1D Riemann, 512 instances L in [64,2048] x4x4, {name}, SSPRK(3,3), "
                     f"CFL 0.3 on the device, t=0.15", extra)


def adjoint_timing(a, fd_dirs=0):
    """Channel-like periodic box 128x64x64 (2 boxes of 64^3), KEEP4 + viscous,
    SSPRK(3,3), fixed dt from CFL 0.3, objective = kinetic energy after 20
    steps; gradient by the checkpointed adjoint (every 5 steps).  Returns the
    time ratio adjoint run / plain forward, and with fd_dirs > 0 the largest
    relative difference to central differences on seeded unit directions
    (step size per direction from a sweep)."""
    import torch
    from amrsolver_amd import level as LV
    from amrsolver_amd import problems as P
    n, nsteps, kcheck = (128, 64, 64), 20, 5
    cls, p0, _ = P.channel_closures()
    g = P.channel_geom(n)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    lev = LV.Level(g, P.decompose(n, BOX), cls, sch)
    u = P.channel_state(cls, g, n, p0)
    S0, S = lev.new_state(), lev.new_state()
    grad = lev.new_state(ng=0)
    S0.upload(u)
    lev.fill_patch(S0)
    dt = lev.est_time_step(S0, 0.3)
    tiling = lev.autotune(S0)

    def forward():
        for _ in range(nsteps):
            lev.advance(S, 0.0, dt)

    reps = 5
    S.upload(u)
    forward()
    lev.adjoint_run(S0, dt, nsteps, kcheck, grad)
    torch.cuda.synchronize()
    tf = []
    for _ in range(reps):
        S.upload(u)
        lev.synchronize()
        t0 = time.perf_counter()
        forward()
        lev.synchronize()
        tf.append(time.perf_counter() - t0)
    ta = []
    for _ in range(reps):
        t0 = time.perf_counter()
        J = lev.adjoint_run(S0, dt, nsteps, kcheck, grad)
        ta.append(time.perf_counter() - t0)
    t_fwd, t_adj = min(tf), min(ta)
    out = {"workload": f"channel {n[0]}x{n[1]}x{n[2]}, 2 boxes of 64^3, KEEP4+viscous, SSPRK(3,3), "
                       f"{nsteps} steps, checkpoint every {kcheck}", "ratio": t_adj / t_fwd,
           "forward_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3, "J": J, "dt": dt, "tiling": tiling}
    if fd_dirs:
        gr = grad.download()
        rng = np.random.default_rng(20261202)
        worst = 0.0
        steps_used = []

        def J_of(h, v):
            S.upload(u + h * v)
            forward()
            return P.kinetic_energy(S.download(), g)

        for _ in range(fd_dirs):
            v = rng.standard_normal(u.shape)
            v /= np.linalg.norm(v)
            got = float(np.sum(gr * v))
            # sweep: fourth-order central differences at decreasing steps; keep
            # the step whose estimate changes least against the next one
            hs = [0.08, 0.04, 0.02, 0.01, 0.005]
            ests = [(8.0 * (J_of(h, v) - J_of(-h, v)) - (J_of(2 * h, v) - J_of(-2 * h, v))) / (12.0 * h) for h in hs]
            k = int(np.argmin([abs(ests[i + 1] - ests[i]) for i in range(len(hs) - 1)]))
            fd = ests[k + 1]
            steps_used.append(hs[k + 1])
            worst = max(worst, abs(got - fd) / abs(fd))
        out["fd_directions"] = fd_dirs
        out["fd_max_rel_err"] = worst
        out["fd_steps"] = sorted(set(steps_used))
    return out


def channel_timing(a, steps=20, reps=3):
    """The reference's wall-bounded channel (exm/Channel: 96x128x64 in 4 boxes
    of at most 64^3, bc 6 isothermal no-slip walls in y, KEEP4 + viscous +
    body force, SSPRK(3,3)), dt from CFL 0.9 at the initial state; hipGraph
    of `steps` steps.  The walls route it through the pass-structured
    kernels."""
    from amrsolver_amd import level as LV
    from amrsolver_amd import problems as P
    n = (96, 128, 64)
    cls, g, sch, parm = P.wall_channel(n)
    lev = LV.Level(g, P.decompose(n, BOX), cls, sch)
    lev.set_problem(wall_T=parm["Tw"], body_force_x=parm["fx"])
    S = lev.new_state()
    S.upload(P.wall_channel_state(cls, g, n, parm))
    lev.fill_patch(S)
    dt = lev.est_time_step(S, 0.9)
    lev.advance_graph(S, dt, steps)
    lev.synchronize()
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        lev.advance_graph(S, dt, steps)
        lev.synchronize()
        best = min(best, time.perf_counter() - t0)
    ncell = n[0] * n[1] * n[2]
    u = S.download()
    return {"workload": "exm/Channel 96x128x64, 4 boxes, bc 6 walls in y, KEEP4+viscous+body force, SSPRK(3,3), "
                        f"dt from CFL 0.9 ({dt:.3e} s), hipGraph of {steps} steps",
            "value": ncell * 3 * steps / best, "unit": "cell-stage/s", "ms_per_step": best / steps * 1e3,
            "finite": bool(np.isfinite(u).all())}


def tgv_device(cls, g, ng=3):
    """The analytic TGV field on the whole ghosted level array, on the GPU
    (same formulas as problems.tgv_state; float64)."""
    import torch
    n = [g.dom_hi[d] + 1 for d in range(3)]
    dev = torch.device("cuda", 0)
    ax = [g.prob_lo[d] + (torch.arange(-ng, n[d] + ng, dtype=torch.float64, device=dev) + 0.5) * g.dx[d]
          for d in range(3)]
    Z, Y, X = torch.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    T0, mach, Re = 110.4, 0.1, 1600.0
    cs = math.sqrt(cls.gamma * cls.Rspec * T0)
    v0 = mach * cs
    mu = cls.visc_ref * T0 * math.sqrt(T0) / (cls.Tvisc_ref + T0)
    rho0 = Re * mu / v0
    p0 = rho0 * cls.Rspec * T0
    u0 = v0 * torch.sin(X) * torch.cos(Y) * torch.cos(Z)
    u1 = -v0 * torch.cos(X) * torch.sin(Y) * torch.cos(Z)
    p = p0 + rho0 * v0 * v0 / 16.0 * (torch.cos(2.0 * X) + torch.cos(2.0 * Y)) * (torch.cos(2.0 * Z) + 2.0)
    rho = p / (cls.Rspec * T0)
    out = torch.empty((5,) + tuple(X.shape), dtype=torch.float64, device=dev)
    out[0] = rho
    out[1] = rho * u0
    out[2] = rho * u1
    out[3] = 0.0
    out[4] = rho * (cls.cv * T0 + 0.5 * (u0 * u0 + u1 * u1))
    return out


def amr_timing(a, n_base=128, steps=5, reps=3):
    """Batched-AMR TGV levels (base n_base^3, 2 levels of ratio 2, tags where
    the analytic |omega| exceeds the level's 70th percentile, blocking factor
    8, boxes of 8..64 cells): every level is one box table (separate ghosted
    FABs), autotuned, and advanced by hipGraph-captured SSPRK(3,3) steps.
    Per-cell cost = step time / (3 * cells), against the uniform 256^3 level
    with the same scheme (level-wide array, and separate 64^3 FABs)."""
    import torch
    from amrsolver_amd import level as LV
    from amrsolver_amd import problems as P
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    ctx = LV.Context(0)

    def time_level(g, boxes, layout, dt):
        lev = LV.Level(g, boxes, cls, sch, ctx=ctx)
        S = lev.new_state(layout=layout)
        if layout == "boxes":
            for b, (lo, hi) in enumerate(lev.boxes):
                S.upload_box(b, P.tgv_state_box(cls, g, lo, hi, 3))
        else:
            S.upload_ghosted(tgv_device(cls, g))
        lev.fill_patch(S)
        tiling = lev.autotune(S)
        lev.advance_graph(S, dt, steps)
        lev.synchronize()
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            lev.advance_graph(S, dt, steps)
            lev.synchronize()
            best = min(best, time.perf_counter() - t0)
        ncell = sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for lo, hi in lev.boxes)
        ms = best / steps * 1e3
        out = {"boxes": len(boxes), "cells": ncell, "tiling": tiling, "plan": lev.step_plan(S), "ms_per_step": ms,
               "ns_per_cell_stage": ms * 1e6 / (3 * ncell)}
        lev.close()
        del S
        torch.cuda.empty_cache()
        return out

    n = (256, 256, 256)
    gu = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    uni_level = time_level(gu, P.decompose(n, BOX), "level", DT)
    uni_boxes = time_level(gu, P.decompose(n, BOX), "boxes", DT)
    levels = []
    for l, (g, boxes) in enumerate(P.tgv_amr_hierarchy(n_base=n_base, nlev=2)):
        if l == 0:
            continue
        for layout, ref in (("level", uni_level), ("boxes", uni_boxes)):
            r = time_level(g, boxes, layout, DT / 2 ** (l - 1))
            r["level"] = l
            r["layout"] = layout
            r["vs_uniform"] = r["ns_per_cell_stage"] / ref["ns_per_cell_stage"]
            levels.append(r)
    return {"workload": f"TGV AMR base {n_base}^3 + 2 levels (ratio 2, |omega| > 70th percentile, blocking 8, "
                        f"boxes 8..64), KEEP4+viscous, SSPRK(3,3), hipGraph of {steps} steps; level layout = one "
                        f"array per level (boxes are windows), boxes layout = one ghosted array per box; each "
                        f"compared with the uniform 256^3 level in the same layout",
            "uniform_256_level_layout": uni_level, "uniform_256_box_layout": uni_boxes, "levels": levels}


def cpu_hierarchy_baseline(n_base=64):
    """The CPU oracle's subcycled hierarchy step (two-level fills with
    conservative linear interpolation, average-down, the levels' steps with
    OpenMP over boxes on every host core the process may use) on a bounded
    sample of the hierarchy workload: the same generator at base n_base^3,
    one coarse step."""
    import os
    from amrsolver_amd import problems as P
    from oracle import oracle as O
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    levels = P.tgv_amr_hierarchy(n_base=n_base, nlev=2, nest=3, base_grid=32)  # 8 base boxes: the
    threads = _host_threads()                                     # oracle threads over boxes
    H = O.OracleHierarchy(levels)
    H.set_state([P.tgv_state(cls, g, tuple(g.dom_hi[d] + 1 for d in range(3))) for g, _ in levels])
    dt0 = DT * 256 / n_base
    t0 = time.perf_counter()
    H.advance(cls, sch, 0.0, dt0, nthreads=threads)
    el = time.perf_counter() - t0
    cells = [sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for lo, hi in b) for _, b in levels]
    cs = sum(c * 3 * 2 ** l for l, c in enumerate(cells))
    return {"value": cs / el, "unit": "cell-stage/s", "cores": threads, "kind": "port",
            "sample": f"TGV hierarchy base {n_base}^3 (boxes of 32^3) + 2 levels (nested 3), cells {cells}, one "
                      f"coarse step "
                      f"(1/2/4 level steps), {threads} threads, {el:.1f} s"}


def hierarchy_timing(a, layout, n_base=128, steps=3, reps=2, fill_reps=20):
    """The TGV AMR hierarchy with levels nested 3 coarse cells inside each
    other (base n_base^3, 2 levels of ratio 2, |omega| > 70th percentile,
    blocking 8), advanced by whole coarse steps with subcycling: every stage
    of every level fills its ghost cells (coarse-fine ones interpolated in
    time and space from the level below), levels 1 and 2 take 2 and 4 steps
    per coarse step, and each finer level is averaged down after its steps.
    Throughput = cell-stages of all levels per second.  Also reported: the
    same substeps of each level without coupling (amr_advance holding the
    coarse-fine ghost cells), and single calls of the coupling kernels."""
    import torch
    from amrsolver_amd import level as LV
    from amrsolver_amd import problems as P
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    ctx = LV.Context(0)
    levels = P.tgv_amr_hierarchy(n_base=n_base, nlev=2, nest=3)
    print("hierarchy: levels built", file=sys.stderr, flush=True)
    Ls = [LV.Level(g, boxes, cls, sch, ctx=ctx) for g, boxes in levels]
    H = LV.Hierarchy(Ls, layout=layout)

    def upload(S, l):
        g, boxes = levels[l]
        if layout == "level":
            S.upload_ghosted(tgv_device(cls, g))
        else:
            for b, (lo, hi) in enumerate(boxes):
                S.upload_box(b, P.tgv_state_box(cls, g, lo, hi, 3))

    for l in range(len(levels)):
        upload(H.cur[l], l)
    print("hierarchy: states uploaded", file=sys.stderr, flush=True)
    tilings = [L.autotune(H.cur[l]) for l, L in enumerate(Ls)]
    print("hierarchy: autotuned", file=sys.stderr, flush=True)
    dt0 = DT * 256 / n_base
    t = 0.0
    H.advance(t, dt0)
    t += dt0
    ctx.synchronize()
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        for _ in range(steps):
            H.advance(t, dt0)
            t += dt0
        ctx.synchronize()
        best = min(best, time.perf_counter() - t0)
    ms = best / steps * 1e3
    print(f"hierarchy: {ms:.2f} ms per coarse step", file=sys.stderr, flush=True)
    cells = [sum((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) for lo, hi in boxes)
             for _, boxes in levels]
    cell_stages = sum(c * 3 * 2 ** l for l, c in enumerate(cells))
    # the same substeps without coupling: amr_advance per level
    plain = []
    for l, L in enumerate(Ls):
        S = H.tmp[l]
        upload(S, l)
        L.fill_patch(S)
        L.advance(S, 0.0, dt0 / 2 ** l)
        ctx.synchronize()
        bl = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            for _ in range(steps * 2 ** l):
                L.advance(S, 0.0, dt0 / 2 ** l)
            ctx.synchronize()
            bl = min(bl, time.perf_counter() - t0)
        plain.append(bl / steps * 1e3)
    # single coupling calls on the finest level (time-blended fill; average-down)
    top, below = len(Ls) - 1, len(Ls) - 2
    fill_ms = avg_ms = 0.0
    for kind in ("fill", "avg"):
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(fill_reps):
            if kind == "fill":
                Ls[top].fill_two_levels(H.tmp[top], Ls[below], H.old[below], H.cur[below], 0.5, 0.5, 2,
                                        H.ngrow[top])
            else:
                Ls[top].average_down(H.cur[top], Ls[below], H.tmp[below], 2)
        ctx.synchronize()
        v = (time.perf_counter() - t0) / fill_reps * 1e3
        if kind == "fill":
            fill_ms = v
        else:
            avg_ms = v
    # a plotfile of the hierarchy (state + the five derived variables), timed
    # whole: plot variables on the device, packing, copy to the host, files
    import shutil
    import tempfile
    tmpd = tempfile.mkdtemp(prefix="amr_plt_")
    try:
        ctx.synchronize()
        t0 = time.perf_counter()
        names = LV.write_plotfile(os.path.join(tmpd, "plt"), Ls, H.cur, t, [steps + 1, 2 * (steps + 1),
                                                                            4 * (steps + 1)],
                                  derive=LV.DERIVED_NAMES)
        plot_s = time.perf_counter() - t0
        plot_bytes = sum(os.path.getsize(os.path.join(dp, f)) for dp, _, fs in os.walk(tmpd) for f in fs)
    finally:
        shutil.rmtree(tmpd, ignore_errors=True)
    return {"workload": f"TGV AMR hierarchy base {n_base}^3 + 2 levels (ratio 2, |omega| > 70th percentile, "
                        f"blocking 8, boxes 8..64, nested 3 coarse cells), KEEP4+viscous, SSPRK(3,3), subcycled "
                        f"(1/2/4 steps per coarse step), "
                        + ("one array per level (boxes are windows)" if layout == "level" else
                           "one ghosted array per box"),
            "metric": "FP64 cell-updates/sec per RK stage, all levels", "unit": "cell-stage/s",
            "value": cell_stages / (ms * 1e-3), "ms_per_coarse_step": ms, "steps": steps,
            "boxes": [len(b) for _, b in levels], "cells": cells, "tilings": tilings,
            "levels_without_coupling_ms": plain, "coupling_overhead": ms / sum(plain) - 1.0,
            "finest_fill_two_levels_ms": fill_ms, "finest_average_down_ms": avg_ms,
            "plotfile": {"variables": len(names), "bytes": plot_bytes, "seconds": plot_s,
                         "GB_per_s": plot_bytes / plot_s / 1e9}}


def main():
    a = parse()
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        return 2
    if a.workload == "amr":
        print(json.dumps(amr_timing(a)))
        return 0
    if a.workload == "amr-hierarchy":
        for layout in ("level", "boxes"):
            r = hierarchy_timing(a, layout)
            if layout == "level" and not a.no_cpu:
                r["cpu_baseline"] = cpu_hierarchy_baseline()
            print(json.dumps(r), flush=True)
            torch_mod = sys.modules.get("torch")
            if torch_mod is not None:
                torch_mod.cuda.empty_cache()
        return 0
    if a.workload == "channel":
        print(json.dumps(channel_timing(a)))
        return 0
    if a.workload == "adjoint":
        print(json.dumps(adjoint_timing(a, fd_dirs=32)))
        return 0
    if a.workload != "tgv":
        fn = {"sod-w": lambda: run_sod(a, 2), "sod-r": lambda: run_sod(a, 0),
              "riemann-batch-w": lambda: run_riemann_batch(a, 2),
              "riemann-batch-r": lambda: run_riemann_batch(a, 0)}[a.workload]
        print(json.dumps(fn()))
        return 0
    import torch
    from amrsolver_amd import capi
    from amrsolver_amd import level as LV
    from amrsolver_amd import problems as P

    n = (a.n, a.n, a.n)
    cls = P.tgv_closures()
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    boxes = P.decompose(n, BOX)
    sch = P.scheme(flux_euler=1, order_keep=a.keep, visc=True, order_rk=3, stages_rk=3)
    ctx = LV.Context(0)
    lev = LV.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout="level")
    W = [lev.new_state(layout="level") for _ in range(2)]
    S.upload(P.tgv_state(cls, g, n))
    ncell = n[0] * n[1] * n[2]
    stream = ctx.stream
    coefs = [capi.stage_coeffs(3, 3, DT, s)[1:] for s in range(3)]

    lev.fill_patch(S)
    if a.tiling == "auto":
        tiling = lev.autotune(S)
    elif a.tiling == "default":
        tiling = None
    else:
        tiling = tuple(int(x) for x in a.tiling.split(","))
        lev.set_tiling(*tiling)  # TX,TY,TZ[,FORM]
        tiling = tiling[:3]
    form = lev.step_plan(S)["form"]
    t0 = np.array(lev.totals(S))
    # timed region: whole CNS::advance steps (ghost fill + three fused stages)
    for _ in range(a.warmup):
        lev.advance(S, 0.0, DT)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(a.steps):
        lev.advance(S, 0.0, DT)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    # stage-kernel durations: the three stages again, with HIP events on the
    # library's stream around each stage launch
    reps = max(3, min(a.steps, 20))
    ev = [[[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(3)] for _ in range(reps)]
    srcs = [S, W[0], W[1]]
    outs = [W[0], W[1], W[0]]
    for k in range(reps):
        for s in range(3):
            lev.fill_patch(srcs[s])
            ev[k][s][0].record(stream)
            mode, ca, cb, cc, cd = coefs[s]
            lev.rk_stage(mode, ca, cb, cc, cd, S, srcs[s], outs[s])
            ev[k][s][1].record(stream)
    torch.cuda.synchronize()
    dur = np.zeros(3)
    for k in range(reps):
        for s in range(3):
            dur[s] += ev[k][s][0].elapsed_time(ev[k][s][1]) * 1e-3
    dur /= reps
    t1 = np.array(lev.totals(S))
    drift = [abs(t1[c] - t0[c]) / abs(t0[c]) for c in (0, 4)]

    ms_per_step = elapsed / a.steps * 1e3
    value = ncell * 3 * a.steps / elapsed
    # roofline of the dominant kernel: the fused stage kernel, bytes averaged
    # over the three stages (106.7 B per cell-stage)
    alg_bytes = sum(ALG_BYTES) / 3 * ncell
    mean_dur = float(dur.mean())
    achieved = alg_bytes / mean_dur / 1e9
    ops = FP64_OPS_PER_CELL_STAGE if a.keep == 4 else FP64_OPS_KEEP6V
    fp64_tflops = ops * ncell / mean_dur / 1e12
    # HBM bytes per launch measured with FETCH_SIZE / WRITE_SIZE for this
    # tiling (null where no measurement of the tiling in force exists)
    traffic = None
    if os.path.exists(a.traffic) and tiling and a.keep == 4:
        try:
            key = ",".join(str(x) for x in tiling) + (f",{form}" if form else "")
            entry = json.load(open(a.traffic)).get("by_tiling", {}).get(key)
            traffic = entry["bytes_per_launch"] if entry else None
        except (OSError, ValueError, KeyError):
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
        "config": {"workload": f"TGV {n[0]}^3 periodic, {len(boxes)} boxes of {BOX}^3, KEEP{a.keep}+viscous, "
                               f"SSPRK(3,3), dt={DT}", "cells": ncell, "stages_per_step": 3},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "fp64": {"ops_per_cell_stage": ops, "achieved": fp64_tflops,
                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": fp64_tflops / FP64_PEAK_TFLOPS},
        "stage_kernel_ms": [float(x * 1e3) for x in dur],
        "tiling": tiling,
        "kernel_form": form,
        "totals_drift": {"mass": drift[0], "energy": drift[1]},
    }
    if a.keep == 4:
        # the drop-in layout: one ghosted array per box
        out["box_layout"] = box_layout_timing(a, ctx, cls, g, boxes, sch, ms_per_step)
    if not a.no_adjoint and a.keep == 4:
        out["adjoint"] = adjoint_timing(a)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.keep)
        out["cpu_other"] = cpu_extra(a.keep)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
