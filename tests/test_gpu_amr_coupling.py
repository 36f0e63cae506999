"""Coarse-fine coupling on the GPU against the oracle: average-down (ratios 2
and 4), the two-level ghost fill (old state, new state, blend; periodic and
physical boundaries; both layouts) bit for bit, and whole subcycled steps of
a three-level hierarchy to 1e-12 per cell (normalised per component)."""
import math

import numpy as np
import pytest

from amrsolver_amd import capi
from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _pair(kind):
    """(coarse geom, coarse boxes), (fine geom, fine boxes) of a test case."""
    if kind == "periodic":
        L = 2 * math.pi
        gc = P.geom((16, 16, 16), (0, 0, 0), (L, L, L), (1, 1, 1))
        gf = P.geom((32, 32, 32), (0, 0, 0), (L, L, L), (1, 1, 1))
        # the third fine box sits in the corner: its ghost cells wrap around
        return (gc, P.decompose((16, 16, 16), 8)), (gf, [((8, 8, 4), (23, 15, 27)), ((8, 16, 4), (23, 23, 27)),
                                                          ((0, 0, 0), (5, 7, 7))])
    # foextrap in x, slip walls (reflect) in y, periodic z
    gc = P.geom((16, 12, 8), (0, 0, 0), (1.6, 1.2, 0.8), (0, 0, 1), lo_bc=(2, 4, 0), hi_bc=(2, 4, 0))
    gf = P.geom((32, 24, 16), (0, 0, 0), (1.6, 1.2, 0.8), (0, 0, 1), lo_bc=(2, 4, 0), hi_bc=(2, 4, 0))
    return (gc, P.decompose((16, 12, 8), (8, 6, 8))), (gf, [((0, 0, 0), (9, 11, 15)), ((10, 4, 0), (19, 15, 7))])


def _rand(g, seed, scale=1.0):
    n = [g.dom_hi[d] + 1 for d in range(3)]
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((5, n[2], n[1], n[0])) * scale
    u[0] += 3.0
    u[4] += 10.0
    return u


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("ratio", [2, 4])
def test_average_down(ctx, ratio, layout):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case avgdown r{ratio} {layout}", flush=True)
    nc = (8, 8, 8)
    gc = P.geom(nc, (0, 0, 0), (1.0, 1.0, 1.0), (1, 1, 1))
    gf = P.geom(tuple(n * ratio for n in nc), (0, 0, 0), (1.0, 1.0, 1.0), (1, 1, 1))
    cboxes = P.decompose(nc, 4)
    fboxes = [((ratio * 1, ratio * 2, 0), (ratio * 6 - 1, ratio * 5 - 1, ratio * 3 - 1)),
              ((0, 0, ratio * 5), (ratio * 2 - 1, ratio * 8 - 1, ratio * 8 - 1))]
    cls = P.tgv_closures()
    sch = P.scheme()
    C = L.Level(gc, cboxes, cls, sch, ctx=ctx)
    F = L.Level(gf, fboxes, cls, sch, ctx=ctx)
    uc, uf = _rand(gc, 1), _rand(gf, 2)
    Sc, Sf = C.new_state(layout=layout), F.new_state(layout=layout)
    Sc.upload(uc)
    Sf.upload(uf)
    F.average_down(Sf, C, Sc, ratio)
    oc, of = O.OracleLevel(gc, cboxes), O.OracleLevel(gf, fboxes)
    oc.set_state(uc)
    of.set_state(uf)
    of.average_down(oc, ratio)
    assert np.array_equal(Sc.download(), oc.get_state())


def _fill_case(ctx, kind, weights, layout, seed=5, ngrow=3):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    (gc, cboxes), (gf, fboxes) = _pair(kind)
    cls = P.tgv_closures()
    sch = P.scheme()
    C = L.Level(gc, cboxes, cls, sch, ctx=ctx)
    F = L.Level(gf, fboxes, cls, sch, ctx=ctx)
    u0, u1, uf = _rand(gc, seed), _rand(gc, seed + 1), _rand(gf, seed + 2)
    S0, S1, Sf = C.new_state(layout=layout), C.new_state(layout=layout), F.new_state(layout=layout)
    S0.upload(u0)
    S1.upload(u1)
    C.fill_patch(S0)
    C.fill_patch(S1)
    Sf.upload(uf)
    a, b = weights
    F.fill_two_levels(Sf, C, S0, S1, a, b, 2, ngrow)
    o0, o1, of = O.OracleLevel(gc, cboxes), O.OracleLevel(gc, cboxes), O.OracleLevel(gf, fboxes)
    o0.set_state(u0)
    o1.set_state(u1)
    o0.fill_patch()
    o1.fill_patch()
    of.set_state(uf)
    of.fill_two_levels(o0, o0.fabs, o1.fabs, a, b, 2)
    return F, Sf, of


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("weights", [(1.0, 0.0), (0.0, 1.0), (0.25, 0.75)], ids=["old", "new", "blend"])
@pytest.mark.parametrize("kind", ["periodic", "bcs"])
def test_fill_two_levels(ctx, kind, weights, layout):
    print(f"case fill2 {kind} {weights} {layout}", flush=True)
    F, Sf, of = _fill_case(ctx, kind, weights, layout)
    # every ghost cell of every box (in the level layout each box's window of
    # the one array, whose ghost cells inside other boxes are their data)
    for b in range(len(F.boxes)):
        assert np.array_equal(Sf.download_box(b), of.fabs[b]), f"box {b}"


@pytest.mark.parametrize("ngrow", [1, 2])
@pytest.mark.parametrize("kind", ["periodic", "bcs"])
def test_fill_two_levels_inner_layers(ctx, kind, ngrow):
    """ngrow < 3 (separate arrays per box): the inner layers equal the full
    fill's, the outer ones are left as they were (zero)."""
    print(f"case fill2 {kind} ngrow {ngrow}", flush=True)
    F, Sf, of = _fill_case(ctx, kind, (0.25, 0.75), "boxes", ngrow=ngrow)
    k = 3 - ngrow
    for b in range(len(F.boxes)):
        got, want = Sf.download_box(b), of.fabs[b]
        inner = (slice(None), slice(k, got.shape[1] - k), slice(k, got.shape[2] - k), slice(k, got.shape[3] - k))
        assert np.array_equal(got[inner], want[inner]), f"box {b}"
        outer = np.ones(got.shape[1:], dtype=bool)
        outer[inner[1:]] = False
        assert np.all(got[:, outer] == 0.0), f"box {b}"


@pytest.mark.parametrize("layout", ["level", "boxes"])
def test_fill_two_levels_ratio4(ctx, layout):
    """Ratio 4 (offsets -3/8 .. 3/8), periodic in x only, blended in time."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case fill2 ratio4 {layout}", flush=True)
    gc = P.geom((12, 8, 8), (0, 0, 0), (1.2, 0.8, 0.8), (1, 0, 0), lo_bc=(0, 2, 3), hi_bc=(0, 2, 3))
    gf = P.geom((48, 32, 32), (0, 0, 0), (1.2, 0.8, 0.8), (1, 0, 0), lo_bc=(0, 2, 3), hi_bc=(0, 2, 3))
    cboxes = P.decompose((12, 8, 8), (6, 4, 8))
    fboxes = [((0, 4, 8), (15, 23, 27)), ((36, 0, 0), (47, 15, 15))]
    cls, sch = P.tgv_closures(), P.scheme()
    C = L.Level(gc, cboxes, cls, sch, ctx=ctx)
    F = L.Level(gf, fboxes, cls, sch, ctx=ctx)
    u0, u1, uf = _rand(gc, 21), _rand(gc, 22), _rand(gf, 23)
    S0, S1, Sf = C.new_state(layout=layout), C.new_state(layout=layout), F.new_state(layout=layout)
    S0.upload(u0)
    S1.upload(u1)
    C.fill_patch(S0)
    C.fill_patch(S1)
    Sf.upload(uf)
    F.fill_two_levels(Sf, C, S0, S1, 0.625, 0.375, 4)
    o0, o1, of = O.OracleLevel(gc, cboxes), O.OracleLevel(gc, cboxes), O.OracleLevel(gf, fboxes)
    o0.set_state(u0)
    o1.set_state(u1)
    o0.fill_patch()
    o1.fill_patch()
    of.set_state(uf)
    of.fill_two_levels(o0, o0.fabs, o1.fabs, 0.625, 0.375, 4)
    for b in range(len(fboxes)):
        assert np.array_equal(Sf.download_box(b), of.fabs[b]), f"box {b}"


def test_fill_two_levels_rejects_bad_nesting(ctx):
    from amrsolver_amd import level as L
    (gc, cboxes), (gf, fboxes) = _pair("periodic")
    cls, sch = P.tgv_closures(), P.scheme()
    holey = [b for b in cboxes if b[0] != (0, 8, 0)]
    C = L.Level(gc, holey, cls, sch, ctx=ctx)
    F = L.Level(gf, fboxes, cls, sch, ctx=ctx)
    Sc, Sf = C.new_state(layout="boxes"), F.new_state(layout="boxes")
    with pytest.raises(capi.AmrError, match="nested"):
        F.fill_two_levels(Sf, C, Sc, None, 1.0, 0.0, 2)
    with pytest.raises(capi.AmrError, match="ratio"):
        F.fill_two_levels(Sf, C, Sc, None, 1.0, 0.0, 3)
    with pytest.raises(capi.AmrError, match="ngrow"):
        F.fill_two_levels(Sf, C, Sc, None, 1.0, 0.0, 2, 4)


def _nested_levels():
    L = 2 * math.pi
    g0 = P.geom((16, 16, 16), (0, 0, 0), (L, L, L), (1, 1, 1))
    g1 = P.geom((32, 32, 32), (0, 0, 0), (L, L, L), (1, 1, 1))
    g2 = P.geom((64, 64, 64), (0, 0, 0), (L, L, L), (1, 1, 1))
    return [(g0, P.decompose((16, 16, 16), 8)),
            (g1, [((8, 8, 4), (23, 15, 27)), ((8, 16, 4), (23, 23, 27))]),
            (g2, [((24, 24, 16), (39, 39, 47))])]


SCHEMES = [
    ("keep4v_ssprk33", dict(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)),
    ("keep4v_ssprk43", dict(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=4)),
    ("weno_rk2m3", dict(flux_euler=2, visc=False, order_rk=2, stages_rk=3)),
]


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_hierarchy_step(ctx, name, kw, layout):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case hierarchy {name} {layout}", flush=True)
    cls = P.tgv_closures()
    sch = P.scheme(**kw)
    levels = _nested_levels()
    us = [P.tgv_state(cls, g, tuple(g.dom_hi[d] + 1 for d in range(3))) for g, _ in levels]
    H = L.Hierarchy([L.Level(g, b, cls, sch, ctx=ctx) for g, b in levels], layout=layout)
    H.upload(us)
    OH = O.OracleHierarchy(levels)
    OH.set_state(us)
    dt, t = 2e-3, 0.0
    for _ in range(2):
        H.advance(t, dt)
        OH.advance(cls, sch, t, dt, nthreads=8)
        t += dt
    for lev, (g, boxes) in enumerate(levels):
        got, want = H.download(lev), OH.get_state(lev)
        mask = np.zeros(got.shape[1:], dtype=bool)
        for lo, hi in boxes:
            mask[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
        err = rel_err(got[:, mask], want[:, mask])
        assert err <= TOL, f"level {lev}: {err:.3e}"
