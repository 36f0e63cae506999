"""Pins of the oracle's coarse-fine coupling (AMReX FillPatchTwoLevels with
CellConservativeLinear interpolation, average_down, and the subcycled
multi-level step), against what the mathematics fixes.  CPU only.

- time weights: the old or new state exactly at (and within 1e-3 of the
  interval of) its time, linear weights in between;
- average_down: exact means of integer data for ratios 2 and 4, cells not
  under a fine box untouched;
- interpolation: constants and linear data reproduced exactly, the mean of a
  coarse cell's interpolated children equal to the coarse value, the MC clip
  and the one slope factor shared by all components, extrema flattened;
- the hierarchy step: a uniform state stays uniform to the last bits; a fine level
  covering the whole periodic domain steps as the single-level step does
  (twice, dt / 2) and its average lands on the coarse level;
- improperly nested levels and ext_dir stencils are rejected.
"""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from oracle import oracle as O


def test_cf_weights():
    assert O.cf_weights(1.0, 3.0, 1.0) == (1.0, 0.0)
    assert O.cf_weights(1.0, 3.0, 3.0) == (0.0, 1.0)
    assert O.cf_weights(1.0, 3.0, 2.0) == (0.5, 0.5)
    assert O.cf_weights(1.0, 3.0, 1.5) == (0.75, 0.25)
    assert O.cf_weights(1.0, 3.0, 1.0 + 1e-3) == (1.0, 0.0)   # within teps = 2e-3
    assert O.cf_weights(1.0, 3.0, 3.0 - 1e-3) == (0.0, 1.0)
    a, b = O.cf_weights(1.0, 3.0, 1.0 + 3e-3)
    assert a != 1.0 and abs(a + b - 1.0) < 1e-15
    assert O.cf_weights(2.0, 2.0, 2.0) == (0.0, 1.0)


def _two_levels(ratio=2, nc=(16, 12, 8), fine_c=((4, 2, 2), (11, 9, 5)), periodic=(0, 0, 0), bc=2, cgrid=(8, 6, 8)):
    """Coarse level over nc (boxes of cgrid), one fine box over the coarse
    cells fine_c refined."""
    gc = P.geom(nc, (0, 0, 0), (1.6, 1.2, 0.8), periodic, lo_bc=(bc,) * 3, hi_bc=(bc,) * 3)
    nf = tuple(n * ratio for n in nc)
    gf = P.geom(nf, (0, 0, 0), (1.6, 1.2, 0.8), periodic, lo_bc=(bc,) * 3, hi_bc=(bc,) * 3)
    lo, hi = fine_c
    fb = [(tuple(x * ratio for x in lo), tuple((x + 1) * ratio - 1 for x in hi))]
    return (gc, P.decompose(nc, cgrid)), (gf, fb)


def _coarse_level(gc, boxes, u):
    C = O.OracleLevel(gc, boxes)
    C.set_state(u)
    C.fill_patch()
    return C


def _cf_cells(F):
    """(box, i, j, k) of fine ghost cells inside the domain that no fine box covers."""
    out = []
    n = F.n
    for b, (lo, hi) in enumerate(F.boxes):
        for k in range(lo[2] - 3, hi[2] + 4):
            for j in range(lo[1] - 3, hi[1] + 4):
                for i in range(lo[0] - 3, hi[0] + 4):
                    c = (i, j, k)
                    if any(c[d] < 0 or c[d] >= n[d] for d in range(3)):
                        continue
                    if any(all(bl[d] <= c[d] <= bh[d] for d in range(3)) for bl, bh in F.boxes):
                        continue
                    out.append((b, i, j, k))
    return out


def _ghost(F, b, i, j, k):
    lo = F.boxes[b][0]
    return F.fabs[b][:, k - lo[2] + 3, j - lo[1] + 3, i - lo[0] + 3]


@pytest.mark.parametrize("ratio", [2, 4])
def test_average_down_exact_means(ratio):
    (gc, cboxes), (gf, fboxes) = _two_levels(ratio, fine_c=((2, 1, 1), (9, 6, 5)))  # spans two coarse boxes
    rng = np.random.default_rng(3)
    uf = rng.integers(0, 1000, size=(5,) + tuple(gf.dom_hi[d] + 1 for d in (2, 1, 0))).astype(np.float64)
    F = O.OracleLevel(gf, fboxes)
    F.set_state(uf)
    C = O.OracleLevel(gc, cboxes)
    sentinel = np.full((5,) + tuple(gc.dom_hi[d] + 1 for d in (2, 1, 0)), -7.0)
    C.set_state(sentinel)
    F.average_down(C, ratio)
    got = C.get_state()
    (flo, fhi), = fboxes
    r = ratio
    sub = uf[:, flo[2]:fhi[2] + 1, flo[1]:fhi[1] + 1, flo[0]:fhi[0] + 1]
    s = sub.shape
    mean = sub.reshape(5, s[1] // r, r, s[2] // r, r, s[3] // r, r).sum(axis=(2, 4, 6)) / r ** 3
    cl = [flo[d] // r for d in range(3)]
    ch = [(fhi[d] + 1) // r for d in range(3)]
    assert np.array_equal(got[:, cl[2]:ch[2], cl[1]:ch[1], cl[0]:ch[0]], mean)
    mask = np.ones(got.shape, dtype=bool)
    mask[:, cl[2]:ch[2], cl[1]:ch[1], cl[0]:ch[0]] = False
    assert np.all(got[mask] == -7.0)


def test_average_down_rejects_uncoarsenable_boxes():
    (gc, cboxes), (gf, _) = _two_levels(2)
    F = O.OracleLevel(gf, [((1, 0, 0), (8, 7, 7))])
    C = O.OracleLevel(gc, cboxes)
    with pytest.raises(RuntimeError, match="coarsenable"):
        F.average_down(C, 2)


def _fill_fine(ratio, uc_fn, uf_fn, fine_c=((4, 2, 2), (11, 9, 5)), weights=(1.0, 0.0), uc1_fn=None, **kw):
    (gc, cboxes), (gf, fboxes) = _two_levels(ratio, fine_c=fine_c, **kw)
    nc = [gc.dom_hi[d] + 1 for d in range(3)]
    C0 = _coarse_level(gc, cboxes, uc_fn(nc))
    crse1 = None
    if uc1_fn is not None:
        C1 = _coarse_level(gc, cboxes, uc1_fn(nc))
        crse1 = C1.fabs
    F = O.OracleLevel(gf, fboxes)
    nf = [gf.dom_hi[d] + 1 for d in range(3)]
    F.set_state(uf_fn(nf))
    for f in F.fabs:  # ghost cells start as NaN: every filled one must be written
        g = f.copy()
        f[...] = np.nan
        f[:, 3:-3, 3:-3, 3:-3] = g[:, 3:-3, 3:-3, 3:-3]
    F.fill_two_levels(C0, C0.fabs, crse1, weights[0], weights[1], ratio)
    return C0, F


def _grid(n):
    K, J, I = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    return I.astype(float), J.astype(float), K.astype(float)


@pytest.mark.parametrize("ratio", [2, 4])
def test_interp_constant_and_linear_exact(ratio):
    coef = np.array([[7.0, 3.0, -2.0, 5.0], [1.0, 0.5, 0.25, -1.0], [-3.0, 0.0, 2.0, 1.0],
                     [2.0, -1.0, -1.0, 0.5], [40.0, 4.0, 8.0, -2.0]])

    def lin_c(n):
        I, J, K = _grid(n)
        return np.stack([c[0] + c[1] * I + c[2] * J + c[3] * K for c in coef])

    C0, F = _fill_fine(ratio, lin_c, lambda n: np.zeros((5, n[2], n[1], n[0])))
    nc = [C0.n[d] for d in range(3)]
    # stencils that reach the foextrap ghost cells see a kink: not linear there
    cells = [c for c in _cf_cells(F) if all(1 <= c[1 + d] // ratio <= nc[d] - 2 for d in range(3))]
    assert len(cells) > 100
    for b, i, j, k in cells:
        x = [(v + 0.5) / ratio - 0.5 for v in (i, j, k)]  # fine centre in coarse index units
        want = np.array([c[0] + c[1] * x[0] + c[2] * x[1] + c[3] * x[2] for c in coef])
        assert np.array_equal(_ghost(F, b, i, j, k), want), (i, j, k)
    # constants
    C0, F = _fill_fine(ratio, lambda n: np.ones((5, n[2], n[1], n[0])) * np.arange(1.0, 6.0)[:, None, None, None],
                       lambda n: np.zeros((5, n[2], n[1], n[0])))
    for b, i, j, k in _cf_cells(F):
        assert np.array_equal(_ghost(F, b, i, j, k), np.arange(1.0, 6.0))


def test_interp_children_mean_is_the_coarse_value():
    rng = np.random.default_rng(11)

    def rough(n):
        return rng.standard_normal((5, n[2], n[1], n[0])) + 3.0

    C0, F = _fill_fine(2, rough, lambda n: np.zeros((5, n[2], n[1], n[0])))
    cells = {(i, j, k): (b, i, j, k) for b, i, j, k in _cf_cells(F)}
    checked = 0
    cv = C0.get_state()
    for (i, j, k) in list(cells):
        if i % 2 or j % 2 or k % 2:
            continue
        kids = [(i + a, j + bb, k + c) for c in (0, 1) for bb in (0, 1) for a in (0, 1)]
        if not all(q in cells for q in kids):
            continue
        # the 8 children may sit in different fine boxes' ghost shells; any copy will do
        vals = np.array([_ghost(F, *cells[q]) for q in kids])
        ic = (i // 2, j // 2, k // 2)
        assert np.allclose(vals.mean(axis=0), cv[:, ic[2], ic[1], ic[0]], rtol=0, atol=1e-14 * 8)
        checked += 1
    assert checked > 10


def test_interp_shared_limiter_factor_and_extrema():
    """Along x, component 0 reads 0, 1, 1.2 around coarse cell 1 (MC clip:
    central 0.6, limited 0.4) and has a maximum at coarse cell 9; component 1
    is linear in x.  The clip scales component 1's x slope by the same 2/3, and
    at the maximum every component's x slope is zero."""
    def data(n):
        I, J, K = _grid(n)
        u = np.zeros((5, n[2], n[1], n[0]))
        prof = np.array([0.0, 1.0, 1.2, 1.3, 1.4, 1.5, 1.6, 1.7, 1.8, 2.5, 1.0, 0.5, 0.2, 0.1, 0.05, 0.0])
        u[0] = 1.0 + prof[I.astype(int)]
        u[1] = 10.0 + I
        u[2] = 2.0 + 0.5 * J
        u[3] = 1.0
        u[4] = 5.0 + K
        return u

    C0, F = _fill_fine(2, data, lambda n: np.zeros((5, n[2], n[1], n[0])), fine_c=((2, 2, 2), (8, 9, 5)))
    cv = C0.get_state()
    cells = _cf_cells(F)
    seen_clip = seen_max = 0
    for b, i, j, k in cells:
        ic, jc, kc = i // 2, j // 2, k // 2
        g = _ghost(F, b, i, j, k)
        offx = (i - 2 * ic + 0.5) / 2 - 0.5
        if ic == 1 and 2 <= jc <= 9 and 2 <= kc <= 5:
            # the shared factor: component 1's x increment is 2/3 of its central slope
            ratio1 = (g[1] - cv[1, kc, jc, ic]) / offx
            assert abs(ratio1 - 2.0 / 3.0) < 1e-14
            ratio0 = (g[0] - cv[0, kc, jc, ic]) / offx
            assert abs(ratio0 - 0.4) < 1e-14
            seen_clip += 1
        if ic == 9 and 2 <= jc <= 9 and 2 <= kc <= 5:
            assert g[0] == cv[0, kc, jc, ic] and g[1] == cv[1, kc, jc, ic]
            seen_max += 1
        # bounds: no new extrema along x for component 0
        lo = min(cv[0, kc, jc, max(ic - 1, 0)], cv[0, kc, jc, ic], cv[0, kc, jc, min(ic + 1, 15)])
        hi = max(cv[0, kc, jc, max(ic - 1, 0)], cv[0, kc, jc, ic], cv[0, kc, jc, min(ic + 1, 15)])
        assert lo - 1e-14 <= g[0] <= hi + 1e-14
    assert seen_clip > 0 and seen_max > 0


def test_interp_time_weights():
    """Old only, new only (NaN in the unused state) and the blend, which is
    linear in the two states."""
    def lin(n, s=1.0):
        I, J, K = _grid(n)
        return np.stack([s * (1.0 + I + 2 * J + 3 * K) + c for c in range(5)])

    zf = lambda n: np.zeros((5, n[2], n[1], n[0]))
    _, Fa = _fill_fine(2, lin, zf, weights=(1.0, 0.0), uc1_fn=lambda n: np.full((5, n[2], n[1], n[0]), np.nan))
    _, Fb = _fill_fine(2, lambda n: np.full((5, n[2], n[1], n[0]), np.nan), zf, weights=(0.0, 1.0),
                       uc1_fn=lambda n: lin(n, 2.0))
    _, Fm = _fill_fine(2, lin, zf, weights=(0.5, 0.5), uc1_fn=lambda n: lin(n, 2.0))
    for b, i, j, k in _cf_cells(Fa):
        ga, gb, gm = _ghost(Fa, b, i, j, k), _ghost(Fb, b, i, j, k), _ghost(Fm, b, i, j, k)
        assert np.all(np.isfinite(ga)) and np.all(np.isfinite(gb))
        assert np.allclose(gm, 0.5 * ga + 0.5 * gb, rtol=1e-15, atol=1e-13)


def test_interp_periodic_images_and_bcs():
    """A fine box at the corner of a periodic-in-x, foextrap-in-y/z domain:
    ghost cells across the periodic boundary are interpolated from the far
    side's coarse cells, those outside in y/z take the BCs after the
    interpolation."""
    def lin(n):
        I, J, K = _grid(n)
        return np.stack([1.0 + 0.0 * I + 2 * J + 3 * K + c for c in range(5)])

    zf = lambda n: np.zeros((5, n[2], n[1], n[0]))
    C0, F = _fill_fine(2, lin, zf, fine_c=((0, 0, 3), (5, 8, 5)), periodic=(1, 0, 0))
    f = F.fabs[0]
    assert np.all(np.isfinite(f))
    # x ghost cells at i = -1 (periodic image of the last fine column) are
    # interpolated: j, k linear parts only (rows whose y stencil stays inside)
    lo = F.boxes[0][0]
    for k in range(6, 12):
        for j in range(2, 18):
            g = f[:, k - lo[2] + 3, j - lo[1] + 3, -1 - lo[0] + 3]
            x = [(j + 0.5) / 2 - 0.5, (k + 0.5) / 2 - 0.5]
            assert np.array_equal(g, np.array([1.0 + 2 * x[0] + 3 * x[1] + c for c in range(5)]))
    # y ghost below the domain (foextrap) equals the first interior row, also where that row is interpolated
    assert np.array_equal(f[:, :, 2, :], f[:, :, 3, :])


def test_interp_rejects_bad_nesting_and_ext_dir():
    (gc, cboxes), (gf, fboxes) = _two_levels(2, fine_c=((4, 2, 2), (11, 9, 5)))
    nc = [gc.dom_hi[d] + 1 for d in range(3)]
    u = np.ones((5, nc[2], nc[1], nc[0]))
    # coarse level with a hole next to the fine box
    holey = [b for b in cboxes if not (b[0][0] == 0 and b[0][1] == 0 and b[0][2] == 0)]
    C = O.OracleLevel(gc, holey)
    C.set_state(u)
    C.fill_patch()
    F = O.OracleLevel(gf, fboxes)
    F.set_state(np.ones((5, 2 * nc[2], 2 * nc[1], 2 * nc[0])))
    with pytest.raises(RuntimeError, match="nested"):
        F.fill_two_levels(C, C.fabs, None, 1.0, 0.0, 2)
    # slip wall (bc 4) is reflect: fine; bc 1 (ext_dir) next to the stencil: rejected
    (gc, cboxes), (gf, fboxes) = _two_levels(2, fine_c=((0, 2, 2), (5, 9, 5)), bc=1)
    C = O.OracleLevel(gc, cboxes)
    C.set_state(u)
    F = O.OracleLevel(gf, fboxes)
    F.set_state(np.ones((5, 2 * nc[2], 2 * nc[1], 2 * nc[0])))
    with pytest.raises(RuntimeError, match="ext_dir"):
        F.fill_two_levels(C, C.fabs, None, 1.0, 0.0, 2)


SCH = dict(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)


def test_hierarchy_uniform_state_is_steady():
    cls = P.tgv_closures()
    levels = _nested_levels()
    H = O.OracleHierarchy(levels)
    u0 = np.array([1.2, 0.3, -0.2, 0.1, 250000.0])
    us = [np.ones((5, g.dom_hi[2] + 1, g.dom_hi[1] + 1, g.dom_hi[0] + 1)) * u0[:, None, None, None]
          for g, _ in levels]
    H.set_state(us)
    H.advance(cls, P.scheme(**SCH), 0.0, 1e-4)
    for lev, (g, boxes) in enumerate(levels):
        got = H.get_state(lev)
        for lo, hi in boxes:
            # the stage weights 1/3, 2/3 round: equal to the last bit or two
            v = got[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
            assert np.all(np.abs(v - u0[:, None, None, None]) <= 4e-16 * np.abs(u0)[:, None, None, None])


def _nested_levels():
    """Periodic [0, 2 pi]^3: level 0 16^3 (boxes of 8^3), level 1 over coarse
    cells 4..11 in x and y, 2..13 in z (two boxes), level 2 over level-1 cells
    12..19 x 12..19 x 8..23 — each nested at least 3 coarse cells inside the
    one below."""
    import math
    L = 2 * math.pi
    g0 = P.geom((16, 16, 16), (0, 0, 0), (L, L, L), (1, 1, 1))
    g1 = P.geom((32, 32, 32), (0, 0, 0), (L, L, L), (1, 1, 1))
    g2 = P.geom((64, 64, 64), (0, 0, 0), (L, L, L), (1, 1, 1))
    b0 = P.decompose((16, 16, 16), 8)
    b1 = [((8, 8, 4), (23, 15, 27)), ((8, 16, 4), (23, 23, 27))]
    b2 = [((24, 24, 16), (39, 39, 47))]
    return [(g0, b0), (g1, b1), (g2, b2)]


def test_hierarchy_full_coverage_equals_single_level():
    """A fine level covering the periodic domain has no coarse-fine ghost
    cells: it steps as the single-level step does, and level 0 becomes its
    average."""
    import math
    cls = P.tgv_closures()
    L = 2 * math.pi
    g0 = P.geom((8, 8, 8), (0, 0, 0), (L, L, L), (1, 1, 1))
    g1 = P.geom((16, 16, 16), (0, 0, 0), (L, L, L), (1, 1, 1))
    b0, b1 = P.decompose((8, 8, 8), 4), P.decompose((16, 16, 16), 8)
    u1 = P.tgv_state(cls, g1, (16, 16, 16))
    u0 = u1.reshape(5, 8, 2, 8, 2, 8, 2).mean(axis=(2, 4, 6))
    for rk in [(3, 3), (3, 4), (2, 3)]:
        sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=rk[0], stages_rk=rk[1])
        H = O.OracleHierarchy([(g0, b0), (g1, b1)])
        H.set_state([u0, u1])
        H.advance(cls, sch, 0.0, 2e-2)
        S = O.OracleLevel(g1, b1)
        S.set_state(u1)
        S.advance(cls, sch, 1e-2)
        S.advance(cls, sch, 1e-2)
        assert np.array_equal(H.get_state(1), S.get_state())
        F = O.OracleLevel(g1, b1)
        F.set_state(S.get_state())
        C = O.OracleLevel(g0, b0)
        F.average_down(C, 2)
        assert np.array_equal(H.get_state(0), C.get_state())


def test_hierarchy_three_levels_runs_and_stays_finite():
    cls = P.tgv_closures()
    levels = _nested_levels()
    H = O.OracleHierarchy(levels)
    H.set_state([P.tgv_state(cls, g, tuple(g.dom_hi[d] + 1 for d in range(3))) for g, _ in levels])
    H.advance(cls, P.scheme(**SCH), 0.0, 2e-3, nthreads=8)
    for lev in range(3):
        assert np.all(np.isfinite(H.get_state(lev)))


def test_hierarchy_generator_nesting():
    """With nest=3 every coarse stencil of every fine ghost cell lies in the
    coarse level (what the interpolation needs); without it level 2 of the
    TGV hierarchy is not nested that deeply."""
    def dilate(m, w):
        out = m.copy()
        for d in range(3):
            acc = out.copy()
            for s in range(1, w + 1):
                acc |= np.roll(out, s, axis=d) | np.roll(out, -s, axis=d)
            out = acc
        return out

    def cover(g, boxes):
        m = np.zeros([g.dom_hi[d] + 1 for d in range(3)], dtype=bool)
        for lo, hi in boxes:
            m[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = True
        return m

    def nested(levels):
        out = []
        for lev in range(1, len(levels)):
            C, F = cover(*levels[lev - 1]), cover(*levels[lev])
            ghost = dilate(F, 3) & ~F
            n = ghost.shape
            need = dilate(ghost.reshape(n[0] // 2, 2, n[1] // 2, 2, n[2] // 2, 2).any(axis=(1, 3, 5)), 1)
            out.append(not (need & ~C).any())
        return out

    assert nested(P.tgv_amr_hierarchy(64, 2, nest=3)) == [True, True]
    assert nested(P.tgv_amr_hierarchy(64, 2)) != [True, True]


@pytest.mark.parametrize("t", [0.0, 0.5, 1.0, 0.9999, 1e-4, 0.25, 0.75, 2.0 / 3.0])
def test_host_weights_match_the_oracle(t):
    """The host's coarse-level weights (level.cf_weights, used by the
    hierarchy step) are the oracle's, bit for bit."""
    from amrsolver_amd import level as LV
    for t0, t1 in [(0.0, 1.0), (0.1, 0.3), (2.0, 2.0)]:
        tt = t0 + t * (t1 - t0)
        assert LV.cf_weights(t0, t1, tt) == O.cf_weights(t0, t1, tt)


def test_host_fill_times_follow_the_stage_sequence():
    """level.fill_times: one FillPatch time per RHS evaluation of each
    scheme, the times CNS::advance passes (first stage at t, last at the
    scheme's final abscissa)."""
    from amrsolver_amd import capi
    from amrsolver_amd import level as LV
    t, dt = 0.5, 0.125
    cases = {(-2, 0): [t, t + dt], (0, 0): [t], (1, 0): [t], (2, 2): [t, t + dt], (2, 3): [t, t + dt / 2, t + dt],
             (2, 5): [t, t + dt / 4, t + dt / 2, t + 3 * dt / 4, t + dt], (3, 3): [t, t + dt, t + dt / 2],
             (3, 4): [t, t + dt / 2, t + dt, t + dt / 2]}
    for (o, s), want in cases.items():
        got = LV.fill_times(o, s, t, dt)
        assert got == want, (o, s, got)
        assert capi.stage_coeffs(o, s, dt, 0)[0] == len(got)  # host-only entry point


def test_hierarchy_generator_boxes_are_disjoint_and_coarsenable():
    """Every generated level's boxes are disjoint, inside the domain and
    coarsenable by the ratio (average-down needs it), and each fine box lies
    over the level below."""
    levels = P.tgv_amr_hierarchy(64, 2, nest=3)
    for lev, (g, boxes) in enumerate(levels):
        n = [g.dom_hi[d] + 1 for d in range(3)]
        m = np.zeros(n, dtype=np.int32)
        for lo, hi in boxes:
            assert all(0 <= lo[d] <= hi[d] < n[d] for d in range(3))
            assert all(lo[d] % 2 == 0 and (hi[d] + 1) % 2 == 0 for d in range(3))
            m[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] += 1
        assert m.max() == 1
        if lev:
            gc, cb = levels[lev - 1]
            c = np.zeros([gc.dom_hi[d] + 1 for d in range(3)], dtype=bool)
            for lo, hi in cb:
                c[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = True
            for lo, hi in boxes:
                assert c[lo[0] // 2:(hi[0] + 1) // 2, lo[1] // 2:(hi[1] + 1) // 2, lo[2] // 2:(hi[2] + 1) // 2].all()


def test_driver_rejects_bad_arguments():
    """The C++ driver checks its arguments before touching the GPU."""
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "amrsolver_amd", "lib", "cns_driver")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    for args in (["--stages", "5"], ["--n", "4"], ["--bogus"], ["--steps"]):
        r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.returncode, r.stderr)
