"""bc 6 walls of exm/Channel on the GPU (pass-structured kernels: the wall
ghost fill, KEEP order reduction next to the walls, one-sided viscous wall
fluxes, body-force source) against the oracle on the same inputs.  Ghost
fills must agree bit for bit; right-hand sides and steps to 1e-12
normalised per component."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu
N = (24, 32, 16)
MG = (12, 16, 16)


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


@pytest.fixture(scope="module", autouse=True)
def _clear_oracle_problem():
    """The oracle's problem parameters are process-wide: clear them after
    this module so that later oracle calls see no walls or body force."""
    yield
    from oracle import oracle as O
    O.set_problem(None)


def _setup(ctx, sch=None):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    cls, g, sch0, parm = P.wall_channel(N)
    sch = sch0 if sch is None else sch
    lev = L.Level(g, P.decompose(N, MG), cls, sch, ctx=ctx)
    lev.set_problem(wall_T=parm["Tw"], body_force_x=parm["fx"])
    O.set_problem(cls, parm["Tw"], parm["fx"])
    return lev, cls, g, sch, parm, P.wall_channel_state(cls, g, N, parm)


@pytest.mark.parametrize("layout", ["level", "boxes"])
def test_wall_fill_matches_oracle(ctx, layout):
    from oracle import oracle as O
    print(f"case wall fill {layout}", flush=True)
    lev, cls, g, sch, parm, u = _setup(ctx)
    S = lev.new_state(layout=layout)
    S.upload(u)
    lev.fill_patch(S)
    ol = O.OracleLevel(g, lev.boxes)
    ol.set_state(u)
    ol.fill_patch()
    for b in range(len(lev.boxes)):
        assert np.array_equal(S.download_box(b), ol.fabs[b]), f"box {b}"


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("order,visc", [(2, True), (4, True), (6, True), (4, False)],
                         ids=["keep2v", "keep4v", "keep6v", "keep4"])
def test_wall_rhs_matches_oracle(ctx, layout, order, visc):
    from oracle import oracle as O
    print(f"case wall rhs keep{order} visc={visc} {layout}", flush=True)
    sch = P.scheme(flux_euler=1, order_keep=order, visc=visc, order_rk=0, stages_rk=3, source=True)
    lev, cls, g, sch, parm, u = _setup(ctx, sch)
    S = lev.new_state(layout=layout)
    D = lev.new_state(layout=layout, ng=0)
    S.upload(u)
    lev.fill_patch(S)
    lev.compute_rhs(S, D)
    ol = O.OracleLevel(g, lev.boxes)
    ol.set_state(u)
    ol.advance(cls, sch, 1e-7)  # order_rk 0: the new state is dSdt
    err = rel_err(D.download(), ol.get_state())
    assert err <= 1e-12, f"keep{order}: {err:.3e}"


@pytest.mark.parametrize("layout", ["level", "boxes"])
def test_wall_channel_steps_match_oracle(ctx, layout):
    from oracle import oracle as O
    print(f"case wall channel steps {layout}", flush=True)
    lev, cls, g, sch, parm, u = _setup(ctx)
    S = lev.new_state(layout=layout)
    S.upload(u)
    lev.fill_patch(S)
    dt = lev.est_time_step(S, 0.3)
    ol = O.OracleLevel(g, lev.boxes)
    ol.set_state(u)
    for _ in range(3):
        lev.advance(S, 0.0, dt)
        ol.advance(cls, sch, dt, nthreads=8)
    err = rel_err(S.download(), ol.get_state())
    assert err <= 1e-12, f"{layout}: {err:.3e}"


def test_wall_channel_graph_replay(ctx):
    print("case wall channel graph", flush=True)
    lev, cls, g, sch, parm, u = _setup(ctx)
    A, B = lev.new_state(), lev.new_state()
    A.upload(u)
    B.upload(u)
    lev.fill_patch(A)
    dt = lev.est_time_step(A, 0.3)
    for rep in range(3):
        lev.advance_graph(A, dt, 2)
        for _ in range(2):
            lev.advance(B, 0.0, dt)
        assert np.array_equal(A.download(), B.download()), f"replay {rep}"


def test_walls_rejections(ctx):
    from amrsolver_amd import level as L
    from amrsolver_amd.capi import AmrError
    print("case wall rejections", flush=True)
    cls, g, sch, parm = P.wall_channel(N)
    # walls in x are not part of the reference's wall path
    gx = P.geom(N, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (0, 1, 1), lo_bc=(6, 0, 0), hi_bc=(6, 0, 0))
    with pytest.raises(AmrError, match="y only"):
        L.Level(gx, P.decompose(N, MG), cls, sch, ctx=ctx)
    lev = L.Level(g, P.decompose(N, MG), cls, sch, ctx=ctx)
    S = lev.new_state()
    S.upload(P.wall_channel_state(cls, g, N, parm))
    with pytest.raises(AmrError, match="amr_set_problem"):
        lev.fill_patch(S)
    lev.set_problem(wall_T=parm["Tw"], body_force_x=parm["fx"])
    lev.fill_patch(S)
    Lm, G = lev.new_state(ng=0), lev.new_state(ng=0)
    with pytest.raises(AmrError, match="walls"):
        lev.vjp_rhs(S, Lm, G)
