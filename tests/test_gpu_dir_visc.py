"""Viscous WENO5-JS and MUSCL + Riemann schemes on the directional flux
passes (the cell-centred viscous fluxes computed first, each pass adding
their face averages to its Euler faces, CNS_hydro_K.h:254-261): steps in
both layouts, the channel's bc 6 walls (one-sided viscous wall fluxes) and
a batch of independent problems against the oracle (1e-12 normalised per
component), and graph replays equal to eager steps bit for bit.  The
right-hand side also equals, bit for bit, the assembly of the
pass-structured face fluxes (tests/test_gpu_face_fluxes.py)."""
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12
SCHEMES = [("wenov", 2), ("riemannv", 0)]


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


@pytest.fixture(scope="module", autouse=True)
def _clear_oracle_problem():
    yield
    from oracle import oracle as O
    O.set_problem(None)


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("name,flux", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_dir_visc_steps_match_oracle(ctx, name, flux, layout):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case dir visc steps {name} {layout}", flush=True)
    n = (40, 24, 20)
    cls = P.closures_reference()
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.6, 0.5), (1, 1, 1))
    boxes = P.decompose(n, (20, 12, 10))
    u = P.smooth_random_state(cls, n, 61)
    sch = P.scheme(flux_euler=flux, visc=True, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout=layout)
    S.upload(u)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    for _ in range(3):
        lev.advance(S, 0.0, 2e-6)
        ol.advance(cls, sch, 2e-6)
    err = rel_err(S.download(), ol.get_state())
    print(f"  max normalised error {err:.3e}", flush=True)
    lev.close()
    assert err <= TOL, f"{name} {layout}: {err:.3e}"


@pytest.mark.parametrize("name,flux", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_dir_visc_walls_match_oracle(ctx, name, flux):
    """Channel with bc 6 walls in y: the wall rows' one-sided viscous fluxes
    replace the cell-centred ones before the passes average them."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case dir visc walls {name}", flush=True)
    N, MG = (24, 32, 16), (12, 16, 16)
    cls, g, _, parm = P.wall_channel(N)
    sch = P.scheme(flux_euler=flux, visc=True, order_rk=3, stages_rk=3, source=True)
    lev = L.Level(g, P.decompose(N, MG), cls, sch, ctx=ctx)
    lev.set_problem(wall_T=parm["Tw"], body_force_x=parm["fx"])
    O.set_problem(cls, parm["Tw"], parm["fx"])
    u = P.wall_channel_state(cls, g, N, parm)
    S = lev.new_state(layout="boxes")
    S.upload(u)
    lev.fill_patch(S)
    dt = lev.est_time_step(S, 0.3)
    ol = O.OracleLevel(g, lev.boxes)
    ol.set_state(u)
    for _ in range(2):
        lev.advance(S, 0.0, dt)
        ol.advance(cls, sch, dt, nthreads=8)
    err = rel_err(S.download(), ol.get_state())
    print(f"  max normalised error {err:.3e}", flush=True)
    lev.close()
    O.set_problem(None)
    assert err <= TOL, f"{name}: {err:.3e}"


def test_dir_visc_batch_matches_solo_oracle(ctx):
    """Independent periodic viscous WENO problems of different sizes and
    spacings advanced in one batch."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print("case dir visc batch", flush=True)
    cls = P.closures_reference()
    sizes = [(16, 8, 8), (24, 8, 4), (8, 16, 8)]
    geoms, states = [], []
    for s, n in enumerate(sizes):
        g = P.geom(n, (0, 0, 0), (0.5 * n[0] / 16, 0.4 * n[1] / 16, 0.3 * n[2] / 8), (1, 1, 1))
        geoms.append(g)
        states.append(P.smooth_random_state(cls, n, 70 + s))
    sch = P.scheme(flux_euler=2, visc=True)
    bat = L.Batch(geoms, cls, sch, ctx=ctx)
    S = bat.new_state()
    S.upload_problems(states)
    dts = [1e-6, 2e-6, 1.5e-6]
    for _ in range(2):
        bat.advance_dt(S, dts)
    got = S.download_problems()
    bat.close()
    for b, (g, u) in enumerate(zip(geoms, states)):
        ol = O.OracleLevel(g, [(tuple(g.dom_lo), tuple(g.dom_hi))])
        ol.set_state(u)
        for _ in range(2):
            ol.advance(cls, sch, dts[b])
        err = rel_err(got[b], ol.get_state())
        assert err <= TOL, f"problem {b}: {err:.3e}"


@pytest.mark.parametrize("name,flux", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_dir_visc_graph_replay(ctx, name, flux):
    import torch
    from amrsolver_amd import level as L
    print(f"case dir visc graph {name}", flush=True)
    n = (32, 16, 16)
    cls = P.closures_reference()
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.5, 0.5), (1, 1, 1))
    boxes = P.decompose(n, 16)
    u = P.smooth_random_state(cls, n, 63)
    sch = P.scheme(flux_euler=flux, visc=True, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    A, B = lev.new_state(), lev.new_state()
    A.upload(u)
    B.upload(u)
    for _ in range(3):
        lev.advance_graph(A, 1e-6, 2)
        for _ in range(2):
            lev.advance(B, 0.0, 1e-6)
        lev.synchronize()
        assert torch.equal(A.download_tensor(), B.download_tensor())
    lev.close()
