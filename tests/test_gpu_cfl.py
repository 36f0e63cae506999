"""Device-side time-step control (amr_advance_cfl, amr_advance_cfl_graph,
amr_new_dt_levels) against the host policy (level.compute_new_dt, the
restatement of CNS::computeInitialDt / computeNewDt, src/CNS.cpp:282-377)
whose runs the oracle parity tests hold: the same time steps bit for bit,
step by step, and the same final states, eager and graph-captured, on the
Sod tube (256x4x4 to t = 0.2) and the 512-instance 1D Riemann batch (to
t = 0.15); the multi-level rule on a three-level hierarchy."""
import math

import numpy as np
import pytest

from amrsolver_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _sod(ctx, flux):
    from amrsolver_amd import level as L
    n = (256, 4, 4)
    cls = P.closures_reference()
    g = P.geom(n, (0, 0, 0), (1.0, 4 / 256, 4 / 256), (0, 1, 1), lo_bc=(2, 0, 0), hi_bc=(2, 0, 0))
    lev = L.Level(g, P.decompose(n, (64, 4, 4)), cls, P.scheme(flux_euler=flux, visc=False), ctx=ctx)
    return lev, P.sod_state(cls, g, n)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("flux", [2, 0], ids=["weno", "muscl"])
def test_sod_device_policy_matches_host_policy(ctx, flux):
    from amrsolver_amd import level as L
    lev, u0 = _sod(ctx, flux)
    stop, cfl = 0.2, 0.3
    print(f"case sod device cfl flux {flux}", flush=True)
    # host policy
    H = lev.new_state(layout="boxes")
    H.upload(u0)
    t, dt, dts = 0.0, None, []
    while t < stop - 1e-14:
        dt = L.compute_new_dt(lev.est_time_step(H, cfl), dt, t, stop, dt is None)
        lev.advance(H, t, dt)
        t += dt
        dts.append(dt)
    ref = H.download()
    # device policy, eager, one step at a time (clock read back to compare)
    D = lev.new_state(layout="boxes")
    D.upload(u0)
    lev.clock_set(0.0)
    got = []
    for k in range(len(dts) + 3):
        lev.advance_cfl(D, cfl, stop, 1)
        tk, dk, nk = lev.clock_get()
        if nk[0] > len(got):
            got.append(dk[0])
    assert got == dts, "device time steps differ from the host policy's"
    assert tk[0] == t and nk[0] == len(dts)
    assert np.array_equal(D.download(), ref)
    # device policy, graph-captured, 32 steps per replay (extra steps past
    # stop_time leave the state untouched)
    G = lev.new_state(layout="boxes")
    G.upload(u0)
    lev.clock_set(0.0)
    reps = 0
    while True:
        lev.advance_cfl_graph(G, cfl, stop, 32)
        reps += 1
        tg, dg, ng = lev.clock_get()
        if tg[0] >= stop - 1e-14:
            break
        assert reps < 100
    assert ng[0] == len(dts) and tg[0] == t and dg[0] == dts[-1]
    assert np.array_equal(G.download(), ref)
    lev.close()


@pytest.mark.timeout(600)
def test_riemann_batch_device_policy_graph(ctx):
    """All 512 instances to t = 0.15: every instance's time-step sequence
    and final state under the device policy (eager, then graph-captured in
    replays of 64 steps) equal the host policy's."""
    from amrsolver_amd import level as L
    cls = P.closures_reference()
    inst = P.riemann_batch_params()
    geoms, states = P.riemann_batch_problems(cls, inst)
    sch = P.scheme(flux_euler=2, visc=False)
    bat = L.Batch(geoms, cls, sch, ctx=ctx)
    stop, cfl = 0.15, 0.3
    nb = len(inst)
    print("case batch512 device cfl", flush=True)
    H = bat.new_state()
    H.upload_problems(states)
    t = [0.0] * nb
    dtp = [None] * nb
    seq = [[] for _ in range(nb)]
    while True:
        est = bat.est_time_steps(H, cfl)
        dts = []
        for b in range(nb):
            if t[b] >= stop - 1e-14:
                dts.append(0.0)
                continue
            d = L.compute_new_dt(est[b], dtp[b], t[b], stop, dtp[b] is None)
            dtp[b] = d
            dts.append(d)
            seq[b].append(d)
        if not any(dts):
            break
        bat.advance_dt(H, dts)
        for b in range(nb):
            t[b] += dts[b]
    ref = H.download_problems()
    nmax = max(len(s) for s in seq)
    print(f"  host policy: {nmax} batch steps", flush=True)
    # eager device policy, clocks read every step
    D = bat.new_state()
    D.upload_problems(states)
    bat.clock_set(0.0)
    got = [[] for _ in range(nb)]
    for k in range(nmax):
        bat.advance_cfl(D, cfl, stop, 1)
        _, dk, nk = bat.clock_get()
        for b in range(nb):
            if nk[b] > len(got[b]):
                got[b].append(dk[b])
    assert got == seq
    out = D.download_problems()
    assert all(np.array_equal(a, r) for a, r in zip(out, ref))
    # graph-captured device policy
    G = bat.new_state()
    G.upload_problems(states)
    bat.clock_set(0.0)
    reps = 0
    while True:
        bat.advance_cfl_graph(G, cfl, stop, 64)
        reps += 1
        tg, _, ng = bat.clock_get()
        if min(tg) >= stop - 1e-14:
            break
        assert reps < 200
    assert tg == t and ng == [len(s) for s in seq]
    out = G.download_problems()
    assert all(np.array_equal(a, r) for a, r in zip(out, ref))
    bat.close()


def _host_levels(est, n_cycle, cur, stop, dt_prev, first):
    """computeNewDt / computeInitialDt, src/CNS.cpp:283-377, on host values."""
    dt0 = float("inf")
    nf = 1
    for l, e in enumerate(est):
        d = e if first else min(e, 1.1 * dt_prev[l])
        nf *= n_cycle[l]
        dt0 = min(dt0, nf * d)
    if not first:
        eps = 0.001 * dt0
        if stop >= 0.0 and cur + dt0 > stop - eps:
            dt0 = stop - cur
    out, nf = [], 1
    for l in range(len(est)):
        nf *= n_cycle[l]
        out.append(dt0 / nf)
    return out


def test_hierarchy_new_dt_levels(ctx):
    from amrsolver_amd import level as L
    H = P.tgv_amr_hierarchy(n_base=32, nlev=2)
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    levs, sts = [], []
    for g, boxes in H:
        lev = L.Level(g, boxes, cls, sch, ctx=ctx)
        S = lev.new_state(layout="boxes")
        for b, (lo, hi) in enumerate(lev.boxes):
            S.upload_box(b, P.tgv_state_box(cls, g, lo, hi, 3))
        levs.append(lev)
        sts.append(S)
    cfl, n_cycle = 0.5, [1, 2, 2]
    est = [lv.est_time_step(S, cfl) for lv, S in zip(levs, sts)]
    first = L.new_dt_levels(levs, sts, n_cycle, cfl, 0.0, 1.0, [None] * 3, True)
    assert first == _host_levels(est, n_cycle, 0.0, 1.0, None, True)
    # a later step: previous steps small enough that the 1.1x cap binds on
    # one level, and a stop time close enough to cut dt0
    prev = [first[0] * 0.5, first[1] * 2.0, first[2] * 0.95]
    for cur, stop in ((0.0, 1.0), (0.3, 0.3 + 0.6 * first[0]), (0.0, -1.0)):
        got = L.new_dt_levels(levs, sts, n_cycle, cfl, cur, stop, prev, False)
        assert got == _host_levels(est, n_cycle, cur, stop, prev, False), (cur, stop)
    for lv in levs:
        lv.close()
