"""GPU parity of the fused KEEP + viscous stage kernel (every compiled tile
shape, partial z runs, all stage-combination modes) against the CPU oracle
and against the pass-structured kernels.  Tolerance: 1e-12 normalised per
component (the kernels round identically, so agreement is expected to be
exact)."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12

TILINGS = [(16, 16, 64), (16, 16, 7), (8, 8, 5), (32, 8, 16)]
KEEP = [("keep2", 2, False), ("keep4", 4, False), ("keep6", 6, False), ("keep2v", 2, True), ("keep4v", 4, True),
        ("keep6v", 6, True)]


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def setup(seed):
    n = (64, 32, 20)
    cls = P.closures_reference()
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.5, 0.3), (1, 1, 1))
    boxes = P.decompose(n, (32, 16, 10))
    u = P.smooth_random_state(cls, n, seed)
    return cls, g, boxes, u


@pytest.mark.parametrize("tiling", TILINGS, ids=["x".join(map(str, t)) for t in TILINGS])
@pytest.mark.parametrize("name,order,visc", KEEP, ids=[k[0] for k in KEEP])
def test_fused_rhs(ctx, tiling, name, order, visc):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    cls, g, boxes, u = setup(5)
    sch = P.scheme(flux_euler=1, order_keep=order, visc=visc)
    print(f"case fused rhs {name} {tiling}", flush=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    lev.set_tiling(*tiling)
    S = lev.new_state()
    S.upload(u)
    lev.fill_patch(S)
    D = lev.new_state(ng=0)
    lev.compute_rhs(S, D)
    got = D.download()
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ref = ol.rhs(cls, sch)
    lev.set_tiling(0, 0, 0)
    assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("order", [4, 6], ids=["keep4v", "keep6v"])
@pytest.mark.parametrize("order_rk,stages_rk", [(3, 3), (3, 4), (2, 3)], ids=["ssprk33", "ssprk43", "ssprk32"])
@pytest.mark.parametrize("tiling", TILINGS[:3], ids=["x".join(map(str, t)) for t in TILINGS[:3]])
def test_fused_advance(ctx, tiling, order_rk, stages_rk, order):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    cls, g, boxes, u = setup(9)
    sch = P.scheme(flux_euler=1, order_keep=order, visc=True, order_rk=order_rk, stages_rk=stages_rk)
    print(f"case fused advance keep{order} {tiling} {order_rk},{stages_rk}", flush=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    lev.set_tiling(*tiling)
    A = lev.new_state()
    A.upload(u)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    for _ in range(2):
        lev.advance(A, 0.0, 1e-6)
        ol.advance(cls, sch, 1e-6)
    got = A.download()
    # the pass-structured kernels on the same input
    lev.set_tiling(-1)
    B = lev.new_state()
    B.upload(u)
    for _ in range(2):
        lev.advance(B, 0.0, 1e-6)
    lev.set_tiling(0, 0, 0)
    assert rel_err(got, ol.get_state()) <= TOL
    assert np.array_equal(got, B.download())


def test_autotune_picks_a_fitting_tiling(ctx):
    from amrsolver_amd import level as L
    cls, g, boxes, u = setup(2)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state()
    S.upload(u)
    lev.fill_patch(S)
    before = S.download()
    t = lev.autotune(S)
    assert t[0] in (8, 16, 32) and t[2] > 0
    assert np.array_equal(before, S.download())
    lev.set_tiling(0, 0, 0)


@pytest.mark.parametrize("order", [4, 6], ids=["keep4v", "keep6v"])
@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("tiling", [(16, 16, 64), (32, 8, 8), (8, 8, 16)], ids=["16x16", "32x8", "8x8"])
def test_fused_partial_tiles(ctx, tiling, layout, order):
    """Boxes whose x/y extents are not multiples of the tile (tiles overhang
    the box and are masked), including boxes narrower than one tile."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    n = (40, 28, 12)
    cls = P.closures_reference()
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.7, 0.3), (1, 1, 1))
    boxes = [((0, 0, 0), (23, 11, 11)), ((24, 0, 0), (39, 11, 11)), ((0, 12, 0), (7, 27, 11)),
             ((8, 12, 0), (39, 27, 5)), ((8, 12, 6), (39, 27, 11))]
    u = P.smooth_random_state(cls, n, 17)
    sch = P.scheme(flux_euler=1, order_keep=order, visc=True)
    print(f"case fused partial keep{order} {tiling} {layout}", flush=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    lev.set_tiling(*tiling)
    S = lev.new_state(layout=layout)
    S.upload(u)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    for _ in range(2):
        lev.advance(S, 0.0, 1e-6)
        ol.advance(cls, sch, 1e-6)
    lev.set_tiling(0, 0, 0)
    assert rel_err(S.download(), ol.get_state()) <= TOL


@pytest.mark.parametrize("order", [4, 6], ids=["keep4v", "keep6v"])
def test_fused_periodic_level_reads_halo_images(ctx, order):
    """A periodic level held as one array: the fused kernel reads halo cells
    from their periodic images, so the result does not depend on the ghost
    cells (left unfilled here)."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    cls, g, boxes, u = setup(23)
    sch = P.scheme(flux_euler=1, order_keep=order, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    lev.set_tiling(16, 16, 16)
    S = lev.new_state()
    S.upload(u)  # ghost cells stay zero
    D = lev.new_state(ng=0)
    lev.compute_rhs(S, D)
    got = D.download()
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ref = ol.rhs(cls, sch)
    lev.set_tiling(0, 0, 0)
    assert rel_err(got, ref) <= TOL
