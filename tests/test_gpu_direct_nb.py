"""Fused stages with tiles per box that read the cells other boxes hold
straight from those boxes' arrays (an owner map of the level's blocks):
a step fills no same-level ghost cells, so the result must not depend on
what those ghost cells hold, must equal the pass-structured kernels' (which
read filled ghost cells) bit for bit and the oracle to 1e-12; a level with
a physical boundary next to its boxes keeps the fills;
amr_fill_two_levels_cf fills exactly the coarse-fine ghost cells
amr_fill_two_levels fills and leaves the others alone."""
import math

import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _tgv(n, mg):
    cls = P.tgv_closures()
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    return cls, g, P.decompose(n, mg)


def _upload_boxes(lev, S, cls, g, ghost=None):
    for b, (lo, hi) in enumerate(lev.boxes):
        a = P.tgv_state_box(cls, g, lo, hi, 3)
        if ghost is not None:
            inner = a[:, 3:-3, 3:-3, 3:-3].copy()
            a[...] = ghost
            a[:, 3:-3, 3:-3, 3:-3] = inner
        S.upload_box(b, a)


@pytest.mark.parametrize("order_rk,stages_rk", [(3, 3), (3, 4)], ids=["ssprk33", "ssprk43"])
def test_direct_reads_match_passes_and_oracle(ctx, order_rk, stages_rk):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case direct nb steps {order_rk},{stages_rk}", flush=True)
    n = (48, 40, 32)
    cls, g, boxes = _tgv(n, (16, 20, 16))
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=order_rk, stages_rk=stages_rk)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    A = lev.new_state(layout="boxes")
    _upload_boxes(lev, A, cls, g)
    lev.set_tiling(16, 16, 16)
    assert lev.step_plan(A)["direct_neighbours"]
    dt = 2e-4
    for _ in range(2):
        lev.advance(A, 0.0, dt)
    got = A.download()
    # the pass-structured kernels on ghost-filled per-box arrays
    B = lev.new_state(layout="boxes")
    _upload_boxes(lev, B, cls, g)
    lev.set_tiling(-1)
    assert lev.step_plan(B)["tiles"] == "passes"
    for _ in range(2):
        lev.advance(B, 0.0, dt)
    assert np.array_equal(got, B.download())
    ol = O.OracleLevel(g, boxes)
    ol.set_state(P.tgv_state(cls, g, n))
    for _ in range(2):
        ol.advance(cls, sch, dt, nthreads=8)
    err = rel_err(got, ol.get_state())
    print(f"  max normalised error {err:.3e}", flush=True)
    lev.set_tiling(0, 0, 0)
    lev.close()
    assert err <= 1e-12


@pytest.mark.parametrize("tiling", [(16, 16, 16, 0), (16, 16, 32, 1), (8, 8, 16, 0), (32, 8, 16, 0)],
                         ids=["16x16", "16x16self", "8x8", "32x8"])
def test_direct_reads_ignore_ghost_cells(ctx, tiling):
    """NaN in every ghost cell of the per-box arrays: the steps read none of
    them (every halo cell is another box's or a periodic image's)."""
    from amrsolver_amd import level as L
    print(f"case direct nb nan ghosts {tiling}", flush=True)
    n = (40, 32, 24)
    cls, g, boxes = _tgv(n, (20, 16, 12))
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    lev.set_tiling(*tiling[:3], form=tiling[3])
    A, B = lev.new_state(layout="boxes"), lev.new_state(layout="boxes")
    _upload_boxes(lev, A, cls, g, ghost=np.nan)
    _upload_boxes(lev, B, cls, g)
    lev.fill_patch(B)
    assert lev.step_plan(A)["direct_neighbours"]
    for _ in range(2):
        lev.advance(A, 0.0, 2e-4)
        lev.advance(B, 0.0, 2e-4)
    a, b = A.download(), B.download()
    lev.set_tiling(0, 0, 0)
    lev.close()
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


def test_physical_boundary_keeps_the_fills(ctx):
    """Outflow in x with boxes on the boundary: no owner map, the stages
    fill their ghost cells as before."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print("case direct nb physical boundary", flush=True)
    n = (32, 24, 16)
    cls = P.closures_reference()
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.75, 0.5), (0, 1, 1), lo_bc=(2, 0, 0), hi_bc=(2, 0, 0))
    boxes = P.decompose(n, (16, 12, 16))
    u = P.smooth_random_state(cls, n, 81)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout="boxes")
    S.upload(u)
    plan = lev.step_plan(S)
    assert plan["tiles"] == "boxes" and not plan["direct_neighbours"]
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    for _ in range(2):
        lev.advance(S, 0.0, 1e-6)
        ol.advance(cls, sch, 1e-6)
    err = rel_err(S.download(), ol.get_state())
    lev.close()
    assert err <= 1e-12


def test_fill_two_levels_cf_fills_only_coarse_fine_cells(ctx):
    from amrsolver_amd import level as L
    print("case fill two levels cf", flush=True)
    Lx = 2 * math.pi
    gc = P.geom((16, 16, 16), (0, 0, 0), (Lx,) * 3, (1, 1, 1))
    gf = P.geom((32, 32, 32), (0, 0, 0), (Lx,) * 3, (1, 1, 1))
    cboxes = P.decompose((16, 16, 16), 8)
    fboxes = [((8, 8, 8), (15, 23, 23)), ((16, 8, 8), (23, 23, 23))]
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    C = L.Level(gc, cboxes, cls, sch, ctx=ctx)
    F = L.Level(gf, fboxes, cls, sch, ctx=ctx)
    Sc = C.new_state(layout="boxes")
    for b, (lo, hi) in enumerate(C.boxes):
        Sc.upload_box(b, P.tgv_state_box(cls, gc, lo, hi, 3))
    C.fill_patch(Sc)
    full, cf = F.new_state(layout="boxes"), F.new_state(layout="boxes")
    _upload_boxes(F, full, cls, gf, ghost=np.nan)
    _upload_boxes(F, cf, cls, gf, ghost=np.nan)
    F.fill_two_levels(full, C, Sc, None, 1.0, 0.0, 2, 3)
    F.fill_two_levels(cf, C, Sc, None, 1.0, 0.0, 2, 3, same_level=False)
    assert F.step_plan(cf)["direct_neighbours"]
    for b, (lo, hi) in enumerate(fboxes):
        a, c = full.download_box(b), cf.download_box(b)
        # ghost cells another fine box holds: copied by the full fill only
        other = np.zeros(a.shape[1:], dtype=bool)
        for ob, (olo, ohi) in enumerate(fboxes):
            if ob == b:
                continue
            sl = [slice(max(olo[d] - lo[d] + 3, 0), max(min(ohi[d] - lo[d] + 4, a.shape[3 - d]), 0)) for d in range(3)]
            other[sl[2], sl[1], sl[0]] = True
        assert np.isfinite(a).all()
        assert np.isnan(c[:, other]).all()
        assert np.array_equal(a[:, ~other], c[:, ~other])
    C.close()
    F.close()
