"""Error estimation and regridding on the GPU against the oracle:
amr_tag_cells (CNS::errorEst, src/CNS.cpp:469-507) bit for bit on evolved
Sod and TGV states and on the channel's wall band; amr_tag_blocks against a
reduction of the cell tags; amr_regrid_fill (CNS::init(old),
src/CNS.cpp:155-179) bit for bit in both layouts; and a two-level TGV run
that regrids its fine level from the device's tags between steps, equal to
the oracle hierarchy that regrids onto the same boxes (1e-12 normalised per
component)."""
import math

import numpy as np
import pytest

from amrsolver_amd import capi
from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _tagger(kind, threshold=0.0, h=0.0, factor=0.0):
    return capi.Tagger(kind=kind, threshold=threshold, h=h, factor=factor)


def test_dengrad_tags_of_an_evolved_sod_state(ctx):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print("case tags sod", flush=True)
    cls = P.closures_reference()
    n = (256, 4, 4)
    g = P.geom(n, (0.0, 0.0, 0.0), (10.0, 10.0, 10.0), (0, 1, 1), lo_bc=(2, 0, 0), hi_bc=(2, 0, 0))
    boxes = P.decompose(n, (64, 4, 4))
    sch = P.scheme(flux_euler=2, visc=False, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state()
    S.upload(P.sod_state(cls, g, n))
    for _ in range(40):
        lev.advance(S, 0.0, 4e-3)
    lev.fill_patch(S)
    tags = lev.tag_cells(S, _tagger(capi.AMR_TAG_DENGRAD, 0.5)).cpu().numpy()
    ol = O.OracleLevel(g, boxes)
    ol.set_state(S.download())
    ol.fill_patch()
    want = ol.tag_cells(1, 0.5)
    assert want.any() and not want.all()
    assert np.array_equal(tags, want)
    # blocks of 4 cells (tags reduced on the device)
    blk = lev.tag_blocks(lev.tag_cells(S, _tagger(capi.AMR_TAG_DENGRAD, 0.5)), 4)
    red = want.reshape(n[2] // 4, 4, n[1] // 4, 4, n[0] // 4, 4).any(axis=(1, 3, 5)).transpose(2, 1, 0)
    assert np.array_equal(blk, red)
    lev.close()


@pytest.mark.parametrize("layout", ["level", "boxes"])
def test_vorticity_tags_of_an_evolved_tgv_state(ctx, layout):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case tags tgv {layout}", flush=True)
    cls = P.tgv_closures()
    n = (48, 48, 48)
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    boxes = P.decompose(n, 16)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout=layout)
    S.upload(P.tgv_state(cls, g, n))
    for _ in range(3):
        lev.advance(S, 0.0, 1e-3)
    lev.fill_patch(S)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(S.download())
    ol.fill_patch()
    for thr in (12.0, 26.0):  # |omega| of the TGV state spans about 0 .. 41
        got = lev.tag_cells(S, _tagger(capi.AMR_TAG_VORT, thr)).cpu().numpy()
        want = ol.tag_cells(3, thr)
        assert want.any() and not want.all()
        assert np.array_equal(got, want), thr
    lev.close()


def test_wallband_tags(ctx):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print("case tags wall band", flush=True)
    cls = P.closures_reference()
    n = (16, 32, 8)
    g = P.geom(n, (0.0, 0.0, 0.0), (2.0, 2.0, 1.0), (1, 1, 1))
    boxes = P.decompose(n, (16, 16, 8))
    lev = L.Level(g, boxes, cls, P.scheme(), ctx=ctx)
    S = lev.new_state(ng=0)
    got = lev.tag_cells(S, _tagger(capi.AMR_TAG_WALLBAND, 0.0, 1.0, 0.5)).cpu().numpy()
    want = O.OracleLevel(g, boxes, ng=0).tag_cells(2, 0.0, 1.0, 0.5)
    assert np.array_equal(got, want) and want.any() and not want.all()
    lev.close()


def _tgv_pair():
    """Coarse 32^3 periodic level (8 boxes), an old and a new fine box set
    that overlap in part (the new one also reaches across the periodic
    seam's neighbourhood)."""
    L = 2 * math.pi
    gc = P.geom((32, 32, 32), (0, 0, 0), (L, L, L), (1, 1, 1))
    gf = P.geom((64, 64, 64), (0, 0, 0), (L, L, L), (1, 1, 1))
    old = [((8, 8, 8), (39, 39, 23)), ((40, 8, 8), (55, 23, 23))]
    new = [((0, 40, 0), (15, 55, 15)), ((16, 16, 16), (47, 31, 47)), ((48, 16, 16), (63, 31, 31))]
    return (gc, P.decompose((32, 32, 32), 16)), gf, old, new


@pytest.mark.parametrize("layout", ["level", "boxes"])
def test_regrid_fill_matches_the_oracle(ctx, layout):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case regrid fill {layout}", flush=True)
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    (gc, cboxes), gf, old_boxes, new_boxes = _tgv_pair()
    C = L.Level(gc, cboxes, cls, sch, ctx=ctx)
    F0 = L.Level(gf, old_boxes, cls, sch, ctx=ctx)
    F1 = L.Level(gf, new_boxes, cls, sch, ctx=ctx)
    nc = [gc.dom_hi[d] + 1 for d in range(3)]
    nf = [gf.dom_hi[d] + 1 for d in range(3)]
    Sc = C.new_state(layout=layout)
    Sc.upload(P.tgv_state(cls, gc, nc))
    C.fill_patch(Sc)
    rng = np.random.default_rng(11)
    uf = P.tgv_state(cls, gf, nf) * (1.0 + 0.01 * rng.random((5,) + tuple(nf[::-1])))
    S0 = F0.new_state(layout=layout)
    S0.upload(uf)
    S1 = F1.new_state(layout=layout)
    F1.regrid_fill(S1, F0, S0, C, None, Sc, 0.0, 1.0, 2)
    oc = O.OracleLevel(gc, cboxes)
    oc.set_state(Sc.download())
    oc.fill_patch()
    of0 = O.OracleLevel(gf, old_boxes)
    of0.set_state(uf)
    of1 = of0.regridded(new_boxes, oc, None, oc.fabs, 0.0, 1.0, 2)
    got, want = S1.download(), of1.get_state()
    assert np.array_equal(got, want)
    # a new level from the coarse data alone (CNS::init())
    S2 = F1.new_state(layout=layout)
    F1.regrid_fill(S2, None, None, C, Sc, None, 1.0, 0.0, 2)
    of2 = O.OracleLevel(gf, []).regridded(new_boxes, oc, oc.fabs, None, 1.0, 0.0, 2)
    assert np.array_equal(S2.download(), of2.get_state())
    for lev in (C, F0, F1):
        lev.close()


def test_hierarchy_regrids_between_steps(ctx):
    """Two coarse steps, a regrid of level 1 from the device's vorticity
    tags on level 0, two more steps; the oracle hierarchy regrids onto the
    boxes its own tags give (asserted equal)."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print("case hierarchy regrid", flush=True)
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    levels = P.tgv_amr_hierarchy(n_base=32, nlev=1, pct=80.0, nest=3)
    us = [P.tgv_state(cls, g, tuple(g.dom_hi[d] + 1 for d in range(3))) for g, _ in levels]
    H = L.Hierarchy([L.Level(g, b, cls, sch, ctx=ctx) for g, b in levels])
    H.upload(us)
    OH = O.OracleHierarchy(levels)
    OH.set_state(us)
    dt, t = 1e-3, 0.0
    for _ in range(2):
        H.advance(t, dt)
        OH.advance(cls, sch, t, dt, nthreads=8)
        t += dt
    thr = 30.0  # above the 90th percentile of |omega|
    boxes = H.regrid(0, _tagger(capi.AMR_TAG_VORT, thr), blocking=8, max_grid=32, eff=0.7, nest=3)
    assert boxes != levels[1][1]
    # the oracle's tags on level 0 give the same boxes
    OH._fill(0, t)
    c0 = OH.levels[0]
    c0.fill_patch()
    tags = c0.tag_cells(3, thr)
    n0 = c0.n
    blk = tags.reshape(n0[2] // 4, 4, n0[1] // 4, 4, n0[0] // 4, 4).any(axis=(1, 3, 5)).transpose(2, 1, 0)
    assert P.boxes_from_block_tags(blk, levels[0][1], n0, 4, 2, 4, 0.7, 3) == boxes
    OH.regrid(0, boxes)
    for lev in range(2):
        got, want = H.download(lev), OH.get_state(lev)
        assert rel_err(got, want) <= 1e-12, lev
    for _ in range(2):
        H.advance(t, dt)
        OH.advance(cls, sch, t, dt, nthreads=8)
        t += dt
    for lev, bl in enumerate([levels[0][1], boxes]):
        got, want = H.download(lev), OH.get_state(lev)
        mask = np.zeros(got.shape[1:], dtype=bool)
        for lo, hi in bl:
            mask[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
        err = rel_err(got[:, mask], want[:, mask])
        assert err <= 1e-12, f"level {lev}: {err:.3e}"
