"""GPU checks of the hand-written reverse mode (C ABI amr_vjp_rhs,
amr_vjp_stage, amr_adjoint_run).

The reference has no adjoint and its autodiff submodule is empty, so the
reverse mode is checked against the forward path itself: dot products
<lam, dL v> from fourth-order central differences of amr_rhs / amr_rk_stage /
amr_advance (the forward path is parity-checked against the oracle elsewhere)
against <VJP(lam), v>.  The bar: relative difference <= 1e-7 for single RHS
and stage products, <= 1e-6 for whole checkpointed runs (directional
derivatives of the objective on seeded unit directions).
"""
import numpy as np
import pytest

from amrsolver_amd import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _dir(rng, u, scaled=True):
    """Random direction; scaled per component by the state's magnitude."""
    v = rng.standard_normal(u.shape)
    if scaled:
        for n in range(u.shape[0]):
            v[n] *= max(np.max(np.abs(u[n])), 1e-3)
    return v


def _fd4(f, h):
    """Fourth-order central difference of a scalar function at 0."""
    return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


SCHEMES = [
    ("keep2", dict(flux_euler=1, order_keep=2, visc=False)),
    ("keep4", dict(flux_euler=1, order_keep=4, visc=False)),
    ("keep6", dict(flux_euler=1, order_keep=6, visc=False)),
    ("keep2v", dict(flux_euler=1, order_keep=2, visc=True)),
    ("keep4v", dict(flux_euler=1, order_keep=4, visc=True)),
    ("keep6v", dict(flux_euler=1, order_keep=6, visc=True)),
]


def _level(ctx, n, mg, cls, sch, g=None):
    from amrsolver_amd import level as L
    if g is None:
        g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.8, 0.7), (1, 1, 1))
    return L.Level(g, P.decompose(n, mg), cls, sch, ctx=ctx), g


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_vjp_rhs_matches_directional_derivative(ctx, name, kw, layout):
    print(f"case vjp_rhs {name} {layout}", flush=True)
    cls = P.closures_reference()
    n, mg = (16, 12, 10), (8, 6, 10)
    sch = P.scheme(**kw)
    lev, g = _level(ctx, n, mg, cls, sch)
    u = P.smooth_random_state(cls, n, 7)
    rng = np.random.default_rng(11)
    lam = rng.standard_normal(u.shape)
    v = _dir(rng, u)
    S = lev.new_state(layout=layout)
    D = lev.new_state(layout=layout, ng=0)
    Lm = lev.new_state(layout=layout, ng=0)
    G = lev.new_state(layout=layout, ng=0)

    def rhs_dot(h):
        S.upload(u + h * v)
        lev.fill_patch(S)
        lev.compute_rhs(S, D)
        return float(np.sum(lam * D.download()))

    fd = _fd4(rhs_dot, 1e-5)
    S.upload(u)
    lev.fill_patch(S)
    Lm.upload(lam)
    lev.vjp_rhs(S, Lm, G)
    got = float(np.sum(G.download() * v))
    assert _rel(got, fd) <= 1e-7, f"{name} {layout}: vjp {got:.15e} fd {fd:.15e}"


def _closures(kind):
    from amrsolver_amd.capi import Closures
    g = P.perfect_gas()
    if kind == "suth":  # Sutherland viscosity and conductivity, 1e4 x the air values
        return Closures(visc_law=1, visc_ref=1.458e-2, Tvisc_ref=110.4, cond_law=1, cond_ref=24.95,
                        Tcond_ref=194.0, Pr=0.71, **g)
    if kind == "const":
        return Closures(visc_law=0, visc_ref=0.2, Tvisc_ref=110.4, cond_law=0, cond_ref=150.0,
                        Tcond_ref=194.0, Pr=0.71, **g)
    return P.closures_scaled_sutherland(0.2, T0=300.0)  # k = mu cp / Pr


@pytest.mark.parametrize("kind", ["suth", "const", "prandtl"])
def test_vjp_rhs_viscous_dominated(ctx, kind):
    """Closures with viscosity large enough that the viscous terms dominate
    the RHS, so the directional-derivative check resolves them."""
    print(f"case vjp_rhs viscous {kind}", flush=True)
    cls = _closures(kind)
    n, mg = (12, 10, 8), (6, 10, 8)
    sch = P.scheme(flux_euler=1, order_keep=2, visc=True)
    lev, g = _level(ctx, n, mg, cls, sch)
    u = P.smooth_random_state(cls, n, 21)
    rng = np.random.default_rng(23)
    lam = rng.standard_normal(u.shape)
    v = _dir(rng, u)
    S = lev.new_state(layout="boxes")
    D, Lm, G = (lev.new_state(layout="boxes", ng=0) for _ in range(3))

    def rhs_dot(h):
        S.upload(u + h * v)
        lev.fill_patch(S)
        lev.compute_rhs(S, D)
        return float(np.sum(lam * D.download()))

    fd = _fd4(rhs_dot, 1e-5)
    S.upload(u)
    lev.fill_patch(S)
    Lm.upload(lam)
    lev.vjp_rhs(S, Lm, G)
    got = float(np.sum(G.download() * v))
    assert _rel(got, fd) <= 1e-7, f"{kind}: vjp {got:.15e} fd {fd:.15e}"


@pytest.mark.parametrize("bc", [(2, 2), (3, 4), (5, 2)], ids=["foextrap", "slipwall-noslip", "symmetry"])
def test_vjp_rhs_physical_boundaries(ctx, bc):
    """Transposed foextrap / reflect fills (x and y non-periodic)."""
    print(f"case vjp_rhs bc {bc}", flush=True)
    cls = P.closures_reference()
    n, mg = (16, 12, 8), (8, 12, 8)
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.8, 0.7), (0, 0, 1), lo_bc=(bc[0], bc[1], 0), hi_bc=(bc[1], bc[0], 0))
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev, _ = _level(ctx, n, mg, cls, sch, g=g)
    u = P.smooth_random_state(cls, n, 3)
    rng = np.random.default_rng(5)
    lam = rng.standard_normal(u.shape)
    v = _dir(rng, u)
    S = lev.new_state(layout="boxes")
    D = lev.new_state(layout="boxes", ng=0)
    Lm = lev.new_state(layout="boxes", ng=0)
    G = lev.new_state(layout="boxes", ng=0)

    def rhs_dot(h):
        S.upload(u + h * v)
        lev.fill_patch(S)
        lev.compute_rhs(S, D)
        return float(np.sum(lam * D.download()))

    fd = _fd4(rhs_dot, 1e-5)
    S.upload(u)
    lev.fill_patch(S)
    Lm.upload(lam)
    lev.vjp_rhs(S, Lm, G)
    got = float(np.sum(G.download() * v))
    assert _rel(got, fd) <= 1e-7, f"bc {bc}: vjp {got:.15e} fd {fd:.15e}"


@pytest.mark.parametrize("mode,coef", [(0, (0.75, 0.25, 1.0, 2e-6)), (1, (0.5, 0.0, 0.7, 3e-6)),
                                       (2, (0.0, 0.0, 0.0, 0.0)), (0, (0.0, 1.0, 1.0, 1e-6))],
                         ids=["mode0", "mode1", "mode2", "mode0-a0"])
def test_vjp_stage(ctx, mode, coef):
    print(f"case vjp_stage mode {mode} {coef}", flush=True)
    cls = P.closures_reference()
    n, mg = (16, 12, 10), (8, 12, 10)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    lev, g = _level(ctx, n, mg, cls, sch)
    u = P.smooth_random_state(cls, n, 9)
    u1 = P.smooth_random_state(cls, n, 10)
    rng = np.random.default_rng(13)
    lam = rng.standard_normal(u.shape)
    v, v1 = _dir(rng, u), _dir(rng, u1)
    a, b, c, d = coef
    S, S1 = lev.new_state(), lev.new_state()
    O = lev.new_state()
    Lo, Ls1, Lsin = (lev.new_state(ng=0) for _ in range(3))

    def stage_dot(h):
        S.upload(u + h * v)
        S1.upload(u1 + h * v1)
        lev.fill_patch(S)
        lev.rk_stage(mode, a, b, c, d, S1, S, O)
        return float(np.sum(lam * O.download()))

    fd = _fd4(stage_dot, 1e-5)
    S.upload(u)
    S1.upload(u1)
    lev.fill_patch(S)
    Lo.upload(lam)
    seed = rng.standard_normal(u.shape)
    Ls1.upload(seed)
    Lsin.upload(seed)
    lev.vjp_stage(mode, a, b, c, d, S, Lo, Ls1, Lsin, accumulate=True)
    got = float(np.sum((Lsin.download() - seed) * v) + np.sum((Ls1.download() - seed) * v1))
    assert _rel(got, fd) <= 1e-7, f"mode {mode}: vjp {got:.15e} fd {fd:.15e}"


def _channel(ctx, n, mg, layout, order_rk=3, stages_rk=3):
    from amrsolver_amd import level as L
    cls, p0, _ = P.channel_closures()
    g = P.channel_geom(n)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=order_rk, stages_rk=stages_rk)
    lev = L.Level(g, P.decompose(n, mg), cls, sch, ctx=ctx)
    u = P.channel_state(cls, g, n, p0)
    S = lev.new_state(layout=layout)
    S.upload(u)
    lev.fill_patch(S)
    dt = lev.est_time_step(S, 0.3)
    return lev, g, u, dt


@pytest.mark.parametrize("nsteps,kcheck,layout,rk", [(6, 2, "level", (3, 3)), (7, 3, "boxes", (3, 3)),
                                                     (4, 4, "level", (3, 4)), (3, 1, "boxes", (2, 3))],
                         ids=["6x2-level", "7x3-boxes", "4x4-ssprk43", "3x1-rk2m3"])
def test_adjoint_run_kinetic_energy(ctx, nsteps, kcheck, layout, rk):
    print(f"case adjoint_run {nsteps} {kcheck} {layout} {rk}", flush=True)
    n, mg = (32, 16, 16), (16, 16, 16)
    lev, g, u, dt = _channel(ctx, n, mg, layout, *rk)
    S0 = lev.new_state(layout=layout)
    S = lev.new_state(layout=layout)
    grad = lev.new_state(layout=layout, ng=0)
    S0.upload(u)
    J = lev.adjoint_run(S0, dt, nsteps, kcheck, grad)
    gr = grad.download()

    def J_of(h, v):
        S.upload(u + h * v)
        for _ in range(nsteps):
            lev.advance(S, 0.0, dt)
        return P.kinetic_energy(S.download(), g)

    assert _rel(J, J_of(0.0, np.zeros_like(u))) <= 1e-12
    rng = np.random.default_rng(20261202)
    for t in range(3):
        v = rng.standard_normal(u.shape)
        v /= np.linalg.norm(v)
        fd = _fd4(lambda h: J_of(h, v), 0.02)
        got = float(np.sum(gr * v))
        assert _rel(got, fd) <= 1e-6, f"direction {t}: adjoint {got:.15e} fd {fd:.15e}"


def test_adjoint_run_linear_objective(ctx):
    """Objective 1 (weights . U): also checks the state is left unchanged."""
    print("case adjoint_run linear", flush=True)
    n, mg = (32, 16, 16), (16, 16, 16)
    lev, g, u, dt = _channel(ctx, n, mg, "level")
    S0, S = lev.new_state(), lev.new_state()
    W = lev.new_state(ng=0)
    grad = lev.new_state(ng=0)
    rng = np.random.default_rng(3)
    w = rng.standard_normal(u.shape)
    W.upload(w)
    S0.upload(u)
    J = lev.adjoint_run(S0, dt, 5, 2, grad, objective=1, weights=W)
    assert np.array_equal(S0.download(), u)
    v = _dir(rng, u)

    def J_of(h):
        S.upload(u + h * v)
        for _ in range(5):
            lev.advance(S, 0.0, dt)
        return float(np.sum(w * S.download()))

    assert _rel(J, J_of(0.0)) <= 1e-12
    fd = _fd4(J_of, 1e-4)
    got = float(np.sum(grad.download() * v))
    assert _rel(got, fd) <= 1e-6, f"adjoint {got:.15e} fd {fd:.15e}"


def test_adjoint_rejects_unsupported(ctx):
    from amrsolver_amd import level as L
    from amrsolver_amd.capi import AmrError
    cls = P.closures_reference()
    n = (8, 8, 8)
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1))
    lev = L.Level(g, [((0, 0, 0), (7, 7, 7))], cls, P.scheme(flux_euler=2, visc=False), ctx=ctx)
    S, Lm, G = lev.new_state(), lev.new_state(ng=0), lev.new_state(ng=0)
    with pytest.raises(AmrError, match="KEEP"):
        lev.vjp_rhs(S, Lm, G)
