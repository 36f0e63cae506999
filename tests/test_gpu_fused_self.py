"""The self-loading form of the fused KEEP + viscous stage kernel (16x16
tiles; each thread converts the cells of the plane above that it reads
itself, which takes one workgroup barrier per plane out): bit for bit equal
to the first form, and within 1e-12 (normalised per component) of the
oracle, over partial z runs, tiles overhanging their boxes, periodic images,
a user_source array, a body force and the ghost cells no box covers."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


@pytest.fixture(scope="module", autouse=True)
def _clear_oracle_problem():
    yield
    from oracle import oracle as O
    O.set_problem(None)


def _both_forms(lev, tiling, run):
    """run() under the first and the self-loading form of `tiling`; returns
    both results after checking that each form ran."""
    out = []
    for form in (0, 1):
        lev.set_tiling(*tiling, form=form)
        out.append(run(form))
    lev.set_tiling(0)
    return out


def _setup(seed, n=(64, 32, 20), mg=(32, 16, 10)):
    cls = P.closures_reference()
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.5, 0.3), (1, 1, 1))
    return cls, g, P.decompose(n, mg), P.smooth_random_state(cls, n, seed)


@pytest.mark.parametrize("tz", [64, 7, 1])
@pytest.mark.parametrize("order", [2, 4], ids=["keep2v", "keep4v"])
def test_self_rhs(ctx, order, tz):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    cls, g, boxes, u = _setup(5)
    sch = P.scheme(flux_euler=1, order_keep=order, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state()
    S.upload(u)
    lev.fill_patch(S)

    def run(form):
        print(f"case self rhs keep{order} tz {tz} form {form}", flush=True)
        assert lev.step_plan(S)["form"] == form
        D = lev.new_state(ng=0)
        lev.compute_rhs(S, D)
        return D.download()

    a, b = _both_forms(lev, (16, 16, tz), run)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    assert np.array_equal(a, b)
    assert rel_err(b, ol.rhs(cls, sch)) <= TOL
    lev.close()


@pytest.mark.parametrize("layout", ["level", "boxes"])
def test_self_partial_tiles_advance(ctx, layout):
    """Boxes narrower than a tile and not multiples of it, SSPRK3 steps."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    n = (40, 28, 12)
    cls = P.closures_reference()
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.7, 0.3), (1, 1, 1))
    boxes = [((0, 0, 0), (23, 11, 11)), ((24, 0, 0), (39, 11, 11)), ((0, 12, 0), (7, 27, 11)),
             ((8, 12, 0), (39, 27, 5)), ((8, 12, 6), (39, 27, 11))]
    u = P.smooth_random_state(cls, n, 17)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout=layout)

    def run(form):
        print(f"case self partial tiles {layout} form {form}", flush=True)
        S.upload(u)
        for _ in range(2):
            lev.advance(S, 0.0, 1e-6)
        return S.download()

    a, b = _both_forms(lev, (16, 16, 5), run)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    for _ in range(2):
        ol.advance(cls, sch, 1e-6)
    assert np.array_equal(a, b)
    assert rel_err(b, ol.get_state()) <= TOL
    lev.close()


def test_self_periodic_images(ctx):
    """A periodic level held as one array with its ghost cells left zero:
    halo cells come from their periodic images."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    cls, g, boxes, u = _setup(23)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state()
    S.upload(u)

    def run(form):
        print(f"case self periodic form {form}", flush=True)
        D = lev.new_state(ng=0)
        lev.compute_rhs(S, D)
        return D.download()

    a, b = _both_forms(lev, (16, 16, 16), run)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    assert np.array_equal(a, b)
    assert rel_err(b, ol.rhs(cls, sch)) <= TOL
    lev.close()


@pytest.mark.parametrize("kind", ["source", "body_force"])
def test_self_source_terms(ctx, kind):
    """A registered user_source array, or the channel's x body force
    (rho fx on x-momentum, mx fx on energy), on the fused path."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    cls = P.closures_reference()
    n = (32, 16, 12)
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.5, 0.4), (1, 1, 1))
    boxes = P.decompose(n, (16, 16, 12))
    u = P.smooth_random_state(cls, n, 31)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, source=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    src = None
    fx = 0.0
    if kind == "source":
        src = np.random.default_rng(4).standard_normal(u.shape) * np.max(np.abs(u), axis=(1, 2, 3))[:, None, None, None]
        Q = lev.new_state(ng=0)
        Q.upload(src)
        lev.set_source(Q)
    else:
        fx = 0.37
        lev.set_problem(body_force_x=fx)
    S = lev.new_state()

    def run(form):
        print(f"case self {kind} form {form}", flush=True)
        S.upload(u)
        for _ in range(2):
            lev.advance(S, 0.0, 1e-6)
        return S.download()

    a, b = _both_forms(lev, (16, 16, 16), run)
    O.set_problem(cls, 0.0, fx)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    for _ in range(2):
        ol.advance(cls, sch, 1e-6, source=src)
    O.set_problem(None)
    assert np.array_equal(a, b)
    assert rel_err(b, ol.get_state()) <= TOL
    lev.close()


def test_self_uncovered_ghosts_from_un(ctx):
    """A refined TGV level (uneven boxes, coarse-fine ghost cells that no box
    covers, read from U^n by the stages after the first) with level tiles."""
    from amrsolver_amd import level as L
    from oracle import oracle as O
    g, boxes = P.tgv_amr_hierarchy(n_base=32, nlev=2)[2]
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    n = [g.dom_hi[d] + 1 for d in range(3)]
    full = P.tgv_state_box(cls, g, (0, 0, 0), tuple(x - 1 for x in n), 3)
    S = lev.new_state(layout="level")
    dt = 5e-5

    def run(form):
        print(f"case self uncovered form {form}", flush=True)
        S.upload_ghosted(full)
        lev.fill_patch(S)
        plan = lev.step_plan(S)
        assert plan["tiles"] == "level" and plan["form"] == form and plan["uncovered_from_un"], plan
        for _ in range(2):
            lev.advance(S, 0.0, dt)
        out = np.zeros_like(full[:, 3:-3, 3:-3, 3:-3])
        got = S.download()
        for lo, hi in boxes:
            sl = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
            out[sl] = got[sl]
        return out

    a, b = _both_forms(lev, (16, 16, 16), run)
    ol = O.OracleLevel(g, boxes)
    for bi, (lo, hi) in enumerate(boxes):
        ol.fabs[bi][...] = P.tgv_state_box(cls, g, lo, hi, 3)
    for _ in range(2):
        ol.advance(cls, sch, dt, nthreads=8)
    assert np.array_equal(a, b)
    assert rel_err(b, ol.get_state()) <= TOL
    lev.close()


def test_self_graph_replays(ctx):
    """Captured steps of the self-loading form equal eager ones on every
    replay."""
    from amrsolver_amd import level as L
    cls, g, boxes, u = _setup(3)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    lev.set_tiling(16, 16, 10, form=1)
    A, B = lev.new_state(), lev.new_state()
    A.upload(u)
    B.upload(u)
    for rep in range(3):
        print(f"case self graph replay {rep}", flush=True)
        lev.advance_graph(A, 1e-6, 2)
        for _ in range(2):
            lev.advance(B, 0.0, 1e-6)
        assert np.array_equal(A.download(), B.download()), f"replay {rep}"
    lev.set_tiling(0)
    lev.close()


def test_self_form_rejections(ctx):
    from amrsolver_amd import level as L
    from amrsolver_amd.capi import AmrError
    cls, g, boxes, u = _setup(1)
    lev = L.Level(g, boxes, cls, P.scheme(flux_euler=1, order_keep=4, visc=False), ctx=ctx)
    with pytest.raises(AmrError, match="self-loading"):
        lev.set_tiling(16, 16, 16, form=1)
    lev2 = L.Level(g, boxes, cls, P.scheme(flux_euler=1, order_keep=4, visc=True), ctx=ctx)
    with pytest.raises(AmrError, match="self-loading"):
        lev2.set_tiling(8, 8, 16, form=1)
    with pytest.raises(AmrError, match="form must be"):
        lev2.set_tiling(16, 16, 16, form=2)
    lev.close()
    lev2.close()
