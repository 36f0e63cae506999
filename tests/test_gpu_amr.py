"""Batched AMR levels: every box of a refined TGV level (uneven boxes of
8 to 64 cells a side, blocking factor 8) advanced in one launch per kernel,
against the oracle on the same ghosted boxes.  Coarse-fine ghost cells hold
the analytic initial field and are left untouched by the single-level fill
on both sides.  Tolerance 1e-12 normalised per component."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


@pytest.fixture(scope="module")
def hierarchy():
    return P.tgv_amr_hierarchy(n_base=32, nlev=2)


def _boxes_state(lev, S, cls, g):
    """Analytic state on every ghosted box (the ghost cells no box covers
    keep these values); returns the per-box arrays."""
    arrs = [P.tgv_state_box(cls, g, lo, hi, 3) for lo, hi in lev.boxes]
    if S.layout == "boxes":
        for b, a in enumerate(arrs):
            S.upload_box(b, a)
    else:
        n = [g.dom_hi[d] + 1 for d in range(3)]
        S.upload_ghosted(P.tgv_state_box(cls, g, (0, 0, 0), tuple(x - 1 for x in n), 3))
    return arrs


def _covered(u, boxes):
    """u with the cells outside every box set to zero (the oracle's view)."""
    out = np.zeros_like(u)
    for lo, hi in boxes:
        sl = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
        out[sl] = u[sl]
    return out


@pytest.mark.parametrize("layout", ["boxes", "level"])
@pytest.mark.parametrize("lev_idx", [1, 2])
@pytest.mark.parametrize("tiling", ["auto", "passes"])
def test_amr_level_step_matches_oracle(ctx, hierarchy, lev_idx, tiling, layout):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    g, boxes = hierarchy[lev_idx]
    print(f"case amr level {lev_idx} boxes {len(boxes)} tiling {tiling} layout {layout}", flush=True)
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout=layout)
    arrs = _boxes_state(lev, S, cls, g)
    lev.fill_patch(S)
    if tiling == "auto":
        lev.autotune(S)
    else:
        lev.set_tiling(-1)
    dt = 2e-4 / (2 ** lev_idx)
    for _ in range(2):
        lev.advance(S, 0.0, dt)
    ol = O.OracleLevel(g, boxes)
    for b, a in enumerate(arrs):
        ol.fabs[b][...] = a
    for _ in range(2):
        ol.advance(cls, sch, dt, nthreads=8)
    err = rel_err(_covered(S.download(), boxes), ol.get_state())
    lev.set_tiling(0)
    assert err <= 1e-12, f"level {lev_idx} {tiling}: {err:.3e}"


@pytest.mark.parametrize("layout", ["boxes", "level"])
def test_amr_level_graph_replay(ctx, hierarchy, layout):
    from amrsolver_amd import level as L
    g, boxes = hierarchy[2]
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    A, B = lev.new_state(layout=layout), lev.new_state(layout=layout)
    _boxes_state(lev, A, cls, g)
    _boxes_state(lev, B, cls, g)
    lev.fill_patch(A)
    lev.autotune(A)
    dt = 5e-5
    for rep in range(3):
        lev.advance_graph(A, dt, 2)
        for _ in range(2):
            lev.advance(B, 0.0, dt)
        assert np.array_equal(A.download(), B.download()), f"replay {rep}"


def test_hierarchy_properties(hierarchy):
    """Blocking factor 8, boxes of 8..64 cells a side, proper nesting, no overlap."""
    for l in range(1, len(hierarchy)):
        g, boxes = hierarchy[l]
        gc, cboxes = hierarchy[l - 1]
        n = [g.dom_hi[d] + 1 for d in range(3)]
        occ = np.zeros(n, dtype=np.int8)
        for lo, hi in boxes:
            for d in range(3):
                assert lo[d] % 8 == 0 and (hi[d] + 1) % 8 == 0
                assert 8 <= hi[d] - lo[d] + 1 <= 64
            clo = [x // 2 for x in lo]
            chi = [x // 2 for x in hi]
            assert any(all(B[0][d] <= clo[d] and chi[d] <= B[1][d] for d in range(3)) for B in cboxes)
            occ[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] += 1
        assert occ.max() == 1


def _irregular_boxes(n, seed=7, keep=0.7):
    """Disjoint boxes with odd boundaries (uneven cuts per direction) and holes
    (about 30% of the pieces dropped): the covered-cell bitmap of level tiling
    then has one-cell blocks and a partly covered bounding box."""
    cuts = ([0, 7, 19, 31, n[0]], [0, 11, 23, n[1]], [0, 5, 17, 30, n[2]])
    rng = np.random.default_rng(seed)
    boxes = []
    for z in range(len(cuts[2]) - 1):
        for y in range(len(cuts[1]) - 1):
            for x in range(len(cuts[0]) - 1):
                if rng.random() < keep:
                    boxes.append(((cuts[0][x], cuts[1][y], cuts[2][z]),
                                  (cuts[0][x + 1] - 1, cuts[1][y + 1] - 1, cuts[2][z + 1] - 1)))
    return boxes


@pytest.mark.parametrize("order", [4, 6], ids=["keep4v", "keep6v"])
@pytest.mark.parametrize("case", ["irregular", "hierarchy"])
def test_level_tiling_matches_oracle(ctx, hierarchy, case, order):
    """Level layout, fused kernel with tiles laid over the union of the boxes
    (crossing box boundaries; only covered cells written) and with tiles per
    box, for several tile shapes, the stages after the first reading the
    ghost cells no box covers from U^n where their tiles allow (else a copy
    pass): every tiling agrees with the oracle to 1e-12 and bit for bit with
    every other tiling."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    if case == "irregular":
        n = (40, 36, 44)
        g = P.geom(n, (0, 0, 0), (2 * np.pi,) * 3, (1, 1, 1))
        boxes = _irregular_boxes(n)
        dt = 2e-4
    else:
        g, boxes = hierarchy[2]
        dt = 5e-5
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=order, visc=True, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    ol = O.OracleLevel(g, boxes)
    S = lev.new_state(layout="level")
    arrs = _boxes_state(lev, S, cls, g)
    for b, a in enumerate(arrs):
        ol.fabs[b][...] = a
    for _ in range(2):
        ol.advance(cls, sch, dt, nthreads=8)
    want = ol.get_state()
    results, in_kernel = [], set()
    for kind in ("level", "boxes"):
        for tiling in [(16, 16, 16), (16, 16, 128), (32, 8, 32), (8, 8, 64), (8, 8, 16)]:
            print(f"case level tiling keep{order} {case} boxes {len(boxes)} tiles {kind} tiling {tiling}", flush=True)
            _boxes_state(lev, S, cls, g)
            lev.fill_patch(S)
            if kind == "level":
                lev.set_tiling(*tiling)
            else:
                lev.set_box_tiling(*tiling)
            plan = lev.step_plan(S)
            if plan["tiles"] == "passes":  # tiles per box too idle for the fused kernel
                assert kind == "boxes"
                continue
            assert plan["tiles"] == kind and plan["tiling"] == tiling, plan
            in_kernel.add(plan["uncovered_from_un"])
            for _ in range(2):
                lev.advance(S, 0.0, dt)
            got = _covered(S.download(), boxes)
            err = rel_err(got, want)
            assert err <= 1e-12, f"{case} {kind} {tiling}: {err:.3e}"
            results.append(got)
    lev.set_tiling(0)
    for r in results[1:]:
        assert np.array_equal(r, results[0])
    assert True in in_kernel
