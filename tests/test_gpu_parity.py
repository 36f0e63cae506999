"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Tolerance (the workload's parity bar): per component
|gpu - oracle| <= 1e-12 * max(|oracle|, max_cells |oracle_n|).
Device code is compiled without FMA contraction in the reference operation
order, so in practice the results are expected to agree to the last bit.
"""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _level_mod():
    from amrsolver_amd import level
    return level


@pytest.fixture(scope="module")
def ctx():
    L = _level_mod()
    return L.Context(0)


def oracle_mod():
    from oracle import oracle
    return oracle


SCHEMES = [
    ("keep2", dict(flux_euler=1, order_keep=2, visc=False)),
    ("keep4", dict(flux_euler=1, order_keep=4, visc=False)),
    ("keep6", dict(flux_euler=1, order_keep=6, visc=False)),
    ("keep2v", dict(flux_euler=1, order_keep=2, visc=True)),
    ("keep4v", dict(flux_euler=1, order_keep=4, visc=True)),
    ("keep6v", dict(flux_euler=1, order_keep=6, visc=True)),
    ("weno", dict(flux_euler=2, visc=False)),
    ("wenov", dict(flux_euler=2, visc=True)),
    ("riemann", dict(flux_euler=0, visc=False)),
    ("riemannv", dict(flux_euler=0, visc=True)),
]


def periodic_setup(n=(24, 20, 16), mg=(12, 10, 8), seed=1):
    cls = P.closures_reference()
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.8, 0.7), (1, 1, 1))
    boxes = P.decompose(n, mg)
    u = P.smooth_random_state(cls, n, seed)
    return cls, g, boxes, u


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_rhs_parity(ctx, name, kw, layout):
    L = _level_mod()
    O = oracle_mod()
    cls, g, boxes, u = periodic_setup()
    sch = P.scheme(**kw)
    print(f"case rhs {name} {layout}", flush=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout=layout)
    S.upload(u)
    lev.fill_patch(S)
    D = lev.new_state(layout=layout, ng=0)
    lev.compute_rhs(S, D)
    got = D.download()
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ref = ol.rhs(cls, sch)
    err = rel_err(got, ref)
    assert err <= TOL, f"{name}: max normalised error {err:.3e}"


def test_fill_patch_sod_boundaries(ctx):
    """Outflow in x (foextrap), reflecting walls in y (slip wall code 4),
    periodic z; ghost cells against the oracle's FillPatch."""
    L = _level_mod()
    O = oracle_mod()
    n = (16, 12, 8)
    cls = P.closures_reference()
    g = P.geom(n, (0, 0, 0), (1, 1, 1), (0, 0, 1), lo_bc=(2, 4, 0), hi_bc=(2, 5, 0))
    boxes = P.decompose(n, (8, 6, 8))
    u = P.smooth_random_state(cls, n, 3)
    lev = L.Level(g, boxes, cls, P.scheme(), ctx=ctx)
    for layout in ("level", "boxes"):
        S = lev.new_state(layout=layout)
        S.upload(u)
        lev.fill_patch(S)
        ol = O.OracleLevel(g, boxes)
        ol.set_state(u)
        ol.fill_patch()
        for b in range(len(boxes)):
            got = S.download_box(b)
            assert np.array_equal(got, ol.fabs[b]), f"box {b} ({layout}) ghost cells differ"


ADV = [
    ("ssprk33", 3, 3),
    ("ssprk43", 3, 4),
    ("ssprk32", 2, 3),
    ("euler", 1, 1),
    ("heun", -2, 2),
    ("rhs_only", 0, 0),
]


@pytest.mark.parametrize("name,order_rk,stages_rk", ADV, ids=[a[0] for a in ADV])
def test_advance_parity(ctx, name, order_rk, stages_rk):
    L = _level_mod()
    O = oracle_mod()
    cls, g, boxes, u = periodic_setup(seed=7)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=order_rk, stages_rk=stages_rk)
    print(f"case advance {name}", flush=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state()
    S.upload(u)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    dt = 2e-6
    nsteps = 1 if order_rk == 0 else 3
    for _ in range(nsteps):
        lev.advance(S, 0.0, dt)
        ol.advance(cls, sch, dt)
    err = rel_err(S.download(), ol.get_state())
    assert err <= TOL, f"{name}: {err:.3e}"


@pytest.mark.parametrize("flux", [2, 0], ids=["weno", "riemann"])
def test_sod_steps(ctx, flux):
    """Sod tube, outflow in x, CFL time step and 1.1x growth cap, 5 steps."""
    L = _level_mod()
    O = oracle_mod()
    n = (64, 4, 4)
    cls = P.closures_reference()
    g = P.geom(n, (0, 0, 0), (1.0, 1.0 / 16, 1.0 / 16), (0, 1, 1), lo_bc=(2, 0, 0), hi_bc=(2, 0, 0))
    boxes = P.decompose(n, (32, 4, 4))
    u = P.sod_state(cls, g, n)
    sch = P.scheme(flux_euler=flux, visc=False)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout="boxes")
    S.upload(u)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    t, dt = 0.0, None
    for step in range(5):
        est_g = lev.est_time_step(S, 0.3)
        est_o = ol.est_dt(cls, 0.3)
        assert est_g == est_o
        dt = L.compute_new_dt(est_g, dt, t, 0.2, step == 0)
        lev.advance(S, t, dt)
        ol.advance(cls, sch, dt)
        t += dt
    err = rel_err(S.download(), ol.get_state())
    assert err <= TOL, f"{err:.3e}"
    tg = lev.totals(S)
    to = ol.totals()
    assert np.allclose(tg, to, rtol=1e-13, atol=1e-13 * np.max(np.abs(to)))


def test_graph_replay_matches_eager(ctx):
    L = _level_mod()
    cls, g, boxes, u = periodic_setup(seed=11)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    A = lev.new_state()
    B = lev.new_state()
    A.upload(u)
    B.upload(u)
    dt = 2e-6
    for rep in range(3):
        lev.advance_graph(A, dt, 2)
        lev.synchronize()
        for _ in range(2):
            lev.advance(B, 0.0, dt)
        lev.synchronize()
        assert np.array_equal(A.download(), B.download()), f"replay {rep} differs from eager"


def test_graph_replay_after_workspace_realloc(ctx):
    """A step on a state of another layout reallocates the box table's
    workspaces, which a graph captured on the first state writes: the graph
    must be dropped and captured again, so every replay still equals eager
    steps (a stale replay would write freed memory)."""
    L = _level_mod()
    cls, g, boxes, u = periodic_setup(seed=12)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    A = lev.new_state()                 # level layout, graph-captured
    E = lev.new_state()                 # level layout, eager twin of A
    B = lev.new_state(layout="boxes")   # per-box arrays: other workspaces
    for S in (A, E, B):
        S.upload(u)
    dt = 2e-6
    for rep in range(3):
        print(f"case graph after realloc rep {rep}", flush=True)
        lev.advance_graph(A, dt, 2)
        for _ in range(2):
            lev.advance(B, 0.0, dt)
        for _ in range(2):
            lev.advance(E, 0.0, dt)
        lev.synchronize()
        a = A.download()
        assert np.array_equal(a, E.download()), f"replay {rep} differs from eager"
        err = rel_err(B.download(), a)
        assert err <= TOL, f"box layout {err:.3e}"


def test_tgv_small_conserves(ctx):
    """TGV 32^3 (KEEP4 + viscous, SSPRK3): oracle parity after 3 steps and
    conserved totals to round-off."""
    L = _level_mod()
    O = oracle_mod()
    n = (32, 32, 32)
    cls = P.tgv_closures()
    two_pi = 2.0 * np.pi
    g = P.geom(n, (0, 0, 0), (two_pi,) * 3, (1, 1, 1))
    boxes = P.decompose(n, 16)
    u = P.tgv_state(cls, g, n)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state()
    S.upload(u)
    t0 = np.array(lev.totals(S))
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    for _ in range(3):
        lev.advance(S, 0.0, 4e-4)
        ol.advance(cls, sch, 4e-4)
    err = rel_err(S.download(), ol.get_state())
    assert err <= TOL
    t1 = np.array(lev.totals(S))
    for c in (0, 4):
        assert abs(t1[c] - t0[c]) <= 1e-13 * abs(t0[c])
