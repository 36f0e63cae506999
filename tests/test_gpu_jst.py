"""JST artificial dissipation (art_diss = 1) on the GPU against the oracle:
right-hand sides and whole steps on a level of several boxes with a pressure
jump (sensor active; the sensor and spectral radius are zero outside each box,
so box boundaries matter), both layouts, every Euler scheme, with and without
viscous fluxes, and next to bc 6 walls.  Tolerance 1e-12 normalised per
component."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12

SCHEMES = [
    ("keep4", dict(flux_euler=1, order_keep=4, visc=False)),
    ("keep2v", dict(flux_euler=1, order_keep=2, visc=True)),
    ("keep6v", dict(flux_euler=1, order_keep=6, visc=True)),
    ("weno", dict(flux_euler=2, visc=False)),
    ("riemannv", dict(flux_euler=0, visc=True)),
]


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _setup(seed=4):
    cls = P.closures_scaled_sutherland(0.01, T0=300.0)
    n = (24, 16, 12)
    g = P.geom(n, (0.0, 0.0, 0.0), (1.2, 0.8, 0.6), (1, 1, 1))
    boxes = P.decompose(n, (8, 8, 12))
    u = P.smooth_random_state(cls, n, seed)
    u[4, :, :, 10:] *= 2.5  # pressure jump inside the second column of boxes
    return cls, g, boxes, u


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_jst_rhs(ctx, name, kw, layout):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case jst rhs {name} {layout}", flush=True)
    cls, g, boxes, u = _setup()
    sch = P.scheme(jst=True, **kw)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout=layout)
    S.upload(u)
    lev.fill_patch(S)
    D = lev.new_state(layout=layout, ng=0)
    lev.compute_rhs(S, D)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ref = ol.rhs(cls, sch)
    plain = ol.rhs(cls, P.scheme(**kw))
    assert np.abs(ref - plain).max() > 0.0  # the dissipation acts
    err = rel_err(D.download(), ref)
    assert err <= TOL, f"{name} {layout}: {err:.3e}"


@pytest.mark.parametrize("name,kw", SCHEMES[:2] + SCHEMES[3:4], ids=["keep4", "keep2v", "weno"])
def test_jst_steps(ctx, name, kw):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case jst steps {name}", flush=True)
    cls, g, boxes, u = _setup(7)
    sch = P.scheme(jst=True, order_rk=3, stages_rk=3, **kw)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout="level")
    S.upload(u)
    dt = 2e-7
    for _ in range(2):
        lev.advance(S, 0.0, dt)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    for _ in range(2):
        ol.advance(cls, sch, dt, nthreads=8)
    err = rel_err(S.download(), ol.get_state())
    assert err <= TOL, f"{name}: {err:.3e}"


@pytest.mark.parametrize("layout", ["level", "boxes"])
def test_jst_next_to_walls(ctx, layout):
    """The bc 6 channel (KEEP4 order reduction at the walls, wall viscous
    fluxes, body force) with JST."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case jst walls {layout}", flush=True)
    N, MG = (32, 24, 16), (16, 12, 16)
    cls, g, _, parm = P.wall_channel(N)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=0, stages_rk=3, source=True, jst=True)
    lev = L.Level(g, P.decompose(N, MG), cls, sch, ctx=ctx)
    lev.set_problem(wall_T=parm["Tw"], body_force_x=parm["fx"])
    u = P.wall_channel_state(cls, g, N, parm)
    u[4, :, :, N[0] // 2:] *= 1.3
    S = lev.new_state(layout=layout)
    D = lev.new_state(layout=layout, ng=0)
    S.upload(u)
    lev.fill_patch(S)
    lev.compute_rhs(S, D)
    O.set_problem(cls, parm["Tw"], parm["fx"])  # process-wide in the oracle: cleared below
    try:
        ol = O.OracleLevel(g, lev.boxes)
        ol.set_state(u)
        ol.advance(cls, sch, 1e-7)  # order_rk 0: the new state is dSdt
    finally:
        O.set_problem(None)
    err = rel_err(D.download(), ol.get_state())
    assert err <= TOL, f"walls {layout}: {err:.3e}"


def test_jst_rejected_by_the_adjoint(ctx):
    from amrsolver_amd import level as L
    cls, g, boxes, u = _setup()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=False, jst=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout="level")
    S.upload(u)
    lev.fill_patch(S)
    Lm, G = lev.new_state(layout="level", ng=0), lev.new_state(layout="level", ng=0)
    with pytest.raises(RuntimeError, match="JST"):
        lev.vjp_rhs(S, Lm, G)
