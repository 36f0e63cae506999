"""Reverse mode against exact directional derivatives: the oracle's
dual-number instantiation (forward mode, tests/test_oracle_tangent.py pins
it) gives J v to round-off, so <lam, J v> and <J^T lam, v> from the GPU
reverse mode agree to round-off, not to a finite-difference truncation
error.  Bars: 1e-12 of |lam| |J v| (Cauchy-Schwarz scale) for the RHS
transpose, 1e-11 of |grad| |v| for whole checkpointed runs."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SCHEMES = [
    ("keep2", dict(flux_euler=1, order_keep=2, visc=False)),
    ("keep4", dict(flux_euler=1, order_keep=4, visc=False)),
    ("keep6", dict(flux_euler=1, order_keep=6, visc=False)),
    ("keep2v", dict(flux_euler=1, order_keep=2, visc=True)),
    ("keep4v", dict(flux_euler=1, order_keep=4, visc=True)),
    ("keep6v", dict(flux_euler=1, order_keep=6, visc=True)),
]


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _scaled_dir(rng, u):
    v = rng.standard_normal(u.shape)
    for n in range(u.shape[0]):
        v[n] *= max(np.max(np.abs(u[n])), 1e-3)
    return v


def _geom(n, bc):
    if bc == "periodic":
        return P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.8, 0.7), (1, 1, 1))
    # foextrap in x, slip wall (lo) / no-slip wall (hi) in y, periodic z
    return P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.8, 0.7), (0, 0, 1), lo_bc=(2, 4, 0), hi_bc=(2, 5, 0))


@pytest.mark.parametrize("bc", ["periodic", "walls"])
@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_vjp_rhs_is_the_exact_transpose(ctx, name, kw, layout, bc):
    from amrsolver_amd import level as L
    print(f"case exact vjp_rhs {name} {layout} {bc}", flush=True)
    cls = P.closures_scaled_sutherland(0.02, T0=300.0)
    n = (16, 12, 10)
    g = _geom(n, bc)
    boxes = P.decompose(n, (8, 6, 10))
    sch = P.scheme(**kw)
    u = P.smooth_random_state(cls, n, 7)
    rng = np.random.default_rng(11)
    lam = rng.standard_normal(u.shape)
    v = _scaled_dir(rng, u)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ol.set_tangent(v)
    _, jv = ol.rhs_tangent(cls, sch)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout=layout)
    Lm, G = lev.new_state(layout=layout, ng=0), lev.new_state(layout=layout, ng=0)
    S.upload(u)
    lev.fill_patch(S)
    Lm.upload(lam)
    lev.vjp_rhs(S, Lm, G)
    got = float(np.sum(G.download() * v))
    want = float(np.sum(lam * jv))
    scale = np.linalg.norm(lam) * np.linalg.norm(jv)
    assert abs(got - want) <= 1e-12 * scale, f"{name}: vjp {got:.17e} tangent {want:.17e}"


def _kinetic_energy_tangent(u, du, g):
    """d/de of sum 0.5 |m|^2 / rho dV along du (the objective of kind 0)."""
    dV = g.dx[0] * g.dx[1] * g.dx[2]
    rho, m = u[0], u[1:4]
    m2 = np.sum(m * m, axis=0)
    return float(np.sum((-0.5 * m2 / rho ** 2 * du[0] + np.sum(m * du[1:4], axis=0) / rho) * dV))


@pytest.mark.parametrize("nsteps,kcheck,layout,rk", [(6, 2, "level", (3, 3)), (5, 3, "boxes", (3, 4)),
                                                     (4, 1, "level", (2, 3))],
                         ids=["6x2-level-ssprk33", "5x3-boxes-ssprk43", "4x1-level-rk2m3"])
def test_adjoint_run_is_the_exact_gradient(ctx, nsteps, kcheck, layout, rk):
    """Channel (walls in y), KEEP4 + viscous: grad . v of the checkpointed
    run against the dual-number derivative of nsteps oracle steps."""
    from amrsolver_amd import level as L
    print(f"case exact adjoint_run {nsteps}x{kcheck} {layout} {rk}", flush=True)
    cls, p0, _ = P.channel_closures()
    n = (32, 16, 16)
    g = P.channel_geom(n)
    boxes = P.decompose(n, (16, 16, 16))
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=rk[0], stages_rk=rk[1])
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    u = P.channel_state(cls, g, n, p0)
    S0 = lev.new_state(layout=layout)
    S0.upload(u)
    lev.fill_patch(S0)
    dt = lev.est_time_step(S0, 0.3)
    grad = lev.new_state(layout=layout, ng=0)
    lev.adjoint_run(S0, dt, nsteps, kcheck, grad)
    gr = grad.download()
    rng = np.random.default_rng(31)
    for t in range(2):
        v = _scaled_dir(rng, u)
        ol = O.OracleLevel(g, boxes)
        ol.set_state(u)
        ol.set_tangent(v)
        for _ in range(nsteps):
            ol.advance_tangent(cls, sch, dt, nthreads=8)
        want = _kinetic_energy_tangent(ol.get_state(), ol.get_tangent(), g)
        got = float(np.sum(gr * v))
        scale = np.linalg.norm(gr) * np.linalg.norm(v)
        assert abs(got - want) <= 1e-11 * scale, f"direction {t}: adjoint {got:.17e} tangent {want:.17e}"


def test_adjoint_run_on_a_refined_level(ctx):
    """Refined TGV level (uneven boxes, coarse-fine ghost cells holding the
    analytic field as constants), level layout, linear objective."""
    from amrsolver_amd import level as L
    print("case exact adjoint_run amr level", flush=True)
    g, boxes = P.tgv_amr_hierarchy(n_base=16, nlev=1)[1]
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    nd = [g.dom_hi[d] + 1 for d in range(3)]
    full = P.tgv_state_box(cls, g, (0, 0, 0), tuple(x - 1 for x in nd), 3)
    S0 = lev.new_state(layout="level")
    S0.upload_ghosted(full)
    rng = np.random.default_rng(4)
    w = rng.standard_normal((5, nd[2], nd[1], nd[0]))
    covered = np.zeros(w.shape[1:], dtype=bool)
    for lo, hi in boxes:
        covered[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = True
    w[:, ~covered] = 0.0
    W, grad = lev.new_state(layout="level", ng=0), lev.new_state(layout="level", ng=0)
    W.upload(w)
    dt, nsteps = 2e-4, 4
    lev.adjoint_run(S0, dt, nsteps, 2, grad, objective=1, weights=W)
    gr = grad.download()
    gr[:, ~covered] = 0.0
    v = _scaled_dir(rng, full[:, 3:-3, 3:-3, 3:-3])
    v[:, ~covered] = 0.0
    ol = O.OracleLevel(g, boxes)
    for b, (lo, hi) in enumerate(boxes):
        ol.fabs[b][...] = full[:, lo[2]:hi[2] + 7, lo[1]:hi[1] + 7, lo[0]:hi[0] + 7]
    ol.set_tangent(v)
    for _ in range(nsteps):
        ol.advance_tangent(cls, sch, dt, nthreads=8)
    want = float(np.sum(w * ol.get_tangent()))
    got = float(np.sum(gr * v))
    scale = np.linalg.norm(gr) * np.linalg.norm(v)
    assert abs(got - want) <= 1e-11 * scale, f"adjoint {got:.17e} tangent {want:.17e}"


@pytest.mark.parametrize("layout", ["boxes", "level"])
def test_adjoint_runs_repeat_bit_for_bit(ctx, layout):
    """The transposed ghost fill adds in a fixed order (no atomics): two
    checkpointed runs with outflow and reflecting walls, on 24 boxes whose
    ghost regions overlap at edges and corners, give the same bits, and so
    do two RHS transposes."""
    from amrsolver_amd import level as L
    print(f"case adjoint repeat {layout}", flush=True)
    cls = P.closures_scaled_sutherland(0.02, T0=300.0)
    n = (24, 18, 20)
    g = _geom(n, "walls")
    boxes = P.decompose(n, (8, 6, 10))
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    u = P.smooth_random_state(cls, n, 5)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S0 = lev.new_state(layout=layout)
    S0.upload(u)
    lev.fill_patch(S0)
    dt = 0.5 * lev.est_time_step(S0, 0.3)
    runs, vjps = [], []
    lam = np.random.default_rng(3).standard_normal(u.shape)
    for _ in range(2):
        grad = lev.new_state(layout=layout, ng=0)
        J = lev.adjoint_run(S0, dt, 4, 2, grad)
        runs.append((J, grad.download()))
        Lm, G = lev.new_state(layout=layout, ng=0), lev.new_state(layout=layout, ng=0)
        Lm.upload(lam)
        lev.vjp_rhs(S0, Lm, G)
        vjps.append(G.download())
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(vjps[0], vjps[1])
    lev.close()
