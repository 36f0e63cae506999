"""Pins of the CPU oracle (oracle/cns_oracle.cpp) by checks that do not
restate its formulas: closed forms, invariants, orders of accuracy,
an independent exact Riemann solver and the reference notebook's exact Sod
state.  CPU only."""
import math

import numpy as np
import pytest

from amrsolver_amd import problems as P
from oracle import oracle as O
import exact_riemann as ER  # noqa: E402

CLS = P.closures_reference()
G = CLS.gamma


# ------------------------------------------------------------------ closures
def test_sutherland_air_values():
    # published air properties at 300 K: mu ~ 1.846e-5 Pa s, k ~ 0.0262 W/(m K)
    assert abs(O.visc(CLS, 300.0) - 1.846e-5) / 1.846e-5 < 2e-3
    assert abs(O.cond(CLS, 300.0) - 0.0262) / 0.0262 < 1e-2
    # at T = Tref the Sutherland law reduces to visc_ref * sqrt(Tref) / 2
    assert O.visc(CLS, 110.4) == pytest.approx(1.458e-6 * math.sqrt(110.4) / 2.0, rel=1e-15)


def test_constant_prandtl_conductivity():
    c = P.tgv_closures()
    for T in (100.0, 300.0, 1000.0):
        assert O.cond(c, T) == pytest.approx(O.visc(c, T) * c.cp / c.Pr, rel=1e-15)


def test_gas_constants():
    # cp - cv = R, cp/cv = gamma
    assert CLS.cp - CLS.cv == pytest.approx(CLS.Rspec, rel=1e-14)
    assert CLS.cp / CLS.cv == pytest.approx(G, rel=1e-14)
    assert CLS.Rspec == pytest.approx(287.1, rel=1e-3)


# ------------------------------------------------------------------ cons2prim
def test_cons2prim_closed_form_and_round_trip():
    rng = np.random.default_rng(0)
    for _ in range(50):
        rho, u, v, w = rng.uniform(0.1, 2), *rng.uniform(-3, 3, 3)
        p = rng.uniform(0.1, 5)
        E = p / (G - 1) + 0.5 * rho * (u * u + v * v + w * w)
        q = O.cons2prim(CLS, [rho, rho * u, rho * v, rho * w, E])
        assert q[0] == rho
        np.testing.assert_allclose(q[1:4], [u, v, w], rtol=1e-14, atol=1e-15)
        assert q[5] == pytest.approx(p, rel=1e-12)
        assert q[4] == pytest.approx(p / (rho * CLS.Rspec), rel=1e-12)


def test_cons2prim_density_clamp():
    q = O.cons2prim(CLS, [0.0, 0.0, 0.0, 0.0, 1.0])
    assert q[0] == 1e-40


# ------------------------------------------------------------------ helpers
def line_state(n, fields, ny=4, nz=4, axis=0):
    """State varying along `axis` only; fields(x) -> rho,u,v,w,p at cell
    centres x in [0,1)."""
    x = (np.arange(n) + 0.5) / n
    rho, un, ut1, ut2, p = fields(x)
    # the field's first velocity is normal to the variation direction
    vel = [[un, ut1, ut2], [ut1, un, ut2], [ut1, ut2, un]][axis]
    u, v, w = vel
    E = p / (G - 1) + 0.5 * rho * (u * u + v * v + w * w)
    U = np.stack([rho, rho * u, rho * v, rho * w, E])  # (5, n)
    if axis == 0:
        full = np.broadcast_to(U[:, :, None, None], (5, n, ny, nz))
        dims = (n, ny, nz)
    elif axis == 1:
        full = np.broadcast_to(U[:, None, :, None], (5, ny, n, nz))
        dims = (ny, n, nz)
    else:
        full = np.broadcast_to(U[:, None, None, :], (5, ny, nz, n))
        dims = (ny, nz, n)
    return P.to_fab_order(full), dims


def periodic_rhs(u, dims, sch, length=(1.0, 1.0, 1.0), cls=CLS, mg=None):
    g = P.geom(dims, (0, 0, 0), length, (1, 1, 1))
    boxes = P.decompose(dims, mg or max(dims))
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    return ol.rhs(cls, sch), g


# ------------------------------------------------------------------ KEEP
@pytest.mark.parametrize("order", [2, 4, 6])
def test_keep_constant_state_is_steady(order):
    u, dims = line_state(16, lambda x: (np.ones_like(x) * 1.2, 0.3 + 0 * x, -0.2 + 0 * x, 0.1 + 0 * x, 0.9 + 0 * x))
    r, _ = periodic_rhs(u, dims, P.scheme(flux_euler=1, order_keep=order, visc=False))
    assert np.max(np.abs(r)) == 0.0


@pytest.mark.parametrize("order", [2, 4, 6])
def test_keep_periodic_sum_vanishes(order):
    u = P.smooth_random_state(CLS, (12, 10, 8), 3)
    r, _ = periodic_rhs(u, (12, 10, 8), P.scheme(flux_euler=1, order_keep=order, visc=True), mg=6)
    for c in range(5):
        assert abs(r[c].sum()) <= 1e-12 * np.abs(r[c]).sum()


def euler_flux_divergence(x):
    """Analytic -dF/dx of the 1D Euler flux for the fields of
    test_keep_order_of_accuracy."""
    tp = 2 * np.pi
    rho = 1 + 0.2 * np.sin(tp * x)
    drho = 0.2 * tp * np.cos(tp * x)
    u = 0.5 + 0.1 * np.cos(tp * x)
    du = -0.1 * tp * np.sin(tp * x)
    p = 1 + 0.1 * np.sin(tp * x + 1)
    dp = 0.1 * tp * np.cos(tp * x + 1)
    E = p / (G - 1) + 0.5 * rho * u * u
    dE = dp / (G - 1) + 0.5 * drho * u * u + rho * u * du
    d_mass = drho * u + rho * du
    d_mom = drho * u * u + 2 * rho * u * du + dp
    d_en = (dE + dp) * u + (E + p) * du
    return np.stack([-d_mass, -d_mom, np.zeros_like(x), np.zeros_like(x), -d_en])


def keep_fields(x):
    tp = 2 * np.pi
    return (1 + 0.2 * np.sin(tp * x), 0.5 + 0.1 * np.cos(tp * x), 0 * x, 0 * x, 1 + 0.1 * np.sin(tp * x + 1))


@pytest.mark.parametrize("order", [2, 4, 6])
def test_keep_order_of_accuracy(order):
    errs = []
    for n in (16, 32, 64):
        u, dims = line_state(n, keep_fields)
        r, _ = periodic_rhs(u, dims, P.scheme(flux_euler=1, order_keep=order, visc=False))
        x = (np.arange(n) + 0.5) / n
        ex = euler_flux_divergence(x)
        errs.append(max(np.max(np.abs(r[c][0, 0, :] - ex[c])) for c in (0, 1, 4)))
    rate = math.log2(errs[1] / errs[2])
    assert abs(rate - order) < 0.35, (errs, rate)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_keep_momentum_consistency(order):
    """Uniform velocity and pressure: the momentum RHS is u0 times the mass
    RHS (the split convective forms reduce to u0 * mass flux)."""
    u0 = 0.7
    u, dims = line_state(24, lambda x: (1 + 0.3 * np.sin(2 * np.pi * x), u0 + 0 * x, 0 * x, 0 * x, 1 + 0 * x))
    r, _ = periodic_rhs(u, dims, P.scheme(flux_euler=1, order_keep=order, visc=False))
    np.testing.assert_allclose(r[1], u0 * r[0], rtol=0, atol=1e-12 * np.max(np.abs(r[0])))


def test_keep_direction_symmetry():
    """The same 1D field along x, y or z gives the same RHS (KEEP treats the
    three directions alike)."""
    outs = []
    for ax in range(3):
        u, dims = line_state(16, keep_fields, axis=ax)
        r, _ = periodic_rhs(u, dims, P.scheme(flux_euler=1, order_keep=4, visc=False))
        line = [r[0].transpose(2, 1, 0)[:, 0, 0], r[0].transpose(2, 1, 0)[0, :, 0], r[0].transpose(2, 1, 0)[0, 0, :]][ax]
        outs.append(line)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


# ------------------------------------------------------------------ WENO5
def cell_avgs(poly, centres):
    """Exact cell averages of a polynomial over unit cells."""
    P_ = np.polyint(poly)
    return np.array([np.polyval(P_, c + 0.5) - np.polyval(P_, c - 0.5) for c in centres])


@pytest.mark.parametrize("poly", [[3.0], [0.5, -1.0], [0.2, -0.7, 1.3], [-1.1, 0.0, 2.0]])
def test_weno_candidates_exact_for_quadratics(poly):
    s = cell_avgs(poly, [0, 1, 2, 3, 4])
    assert O.weno5js(s) == pytest.approx(np.polyval(poly, 2.5), rel=1e-13, abs=1e-13)


def test_weno_eno_selection():
    # a jump between cells 2 and 3: only the left stencil is smooth
    assert abs(O.weno5js([0, 0, 0, 1, 1])) < 1e-14
    # a jump between cells 1 and 2: only the right stencil is smooth
    assert O.weno5js([0, 0, 1, 1, 1]) == pytest.approx(1.0, abs=1e-14)


def test_weno_beta1_is_not_scale_invariant():
    """With beta1 = 13/12 d2^2 + 1/4 (s1-s3)^4 (High_resolution.h:137-140)
    the weights depend on the data scale; a textbook smoothness indicator
    would give weno(c*s) = c*weno(s).  This demonstrates the quartic term;
    its exact form stays unpinned (no reference vectors exist)."""
    s = np.array([0.0, 0.3, 1.0, 1.2, 2.5])
    a = O.weno5js(s)
    b = O.weno5js(10 * s) / 10
    assert abs(a - b) > 1e-6


# ------------------------------------------------------------------ Riemann
def test_limiter_is_monotonized_central():
    rng = np.random.default_rng(4)
    for _ in range(200):
        a, b = rng.normal(size=2)
        mc = 0.0 if a * b <= 0 else math.copysign(min(2 * abs(a), 2 * abs(b), abs(a + b) / 2), a)
        assert O.limiter(a, b) == pytest.approx(mc, rel=1e-15, abs=0)


def test_riemann_equal_states_give_euler_flux():
    rho, u, p, t1, t2 = 0.8, 0.35, 0.7, -0.2, 0.45
    f = O.riemann_prob(G, rho, u, p, t1, t2, rho, u, p, t1, t2)
    E = p / (G - 1) + 0.5 * rho * (u * u + t1 * t1 + t2 * t2)
    np.testing.assert_allclose(f, [rho * u, rho * u * u + p, rho * u * t1, rho * u * t2, u * (E + p)],
                               rtol=1e-14, atol=1e-15)


def test_riemann_mirror_symmetry():
    L = (1.0, 0.2, 1.0, 0.1, -0.3)
    R = (0.3, -0.4, 0.2, 0.5, 0.2)
    f = O.riemann_prob(G, *L, *R)
    m = O.riemann_prob(G, R[0], -R[1], R[2], R[3], R[4], L[0], -L[1], L[2], L[3], L[4])
    np.testing.assert_allclose([m[0], m[1], m[4]], [-f[0], f[1], -f[4]], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("seed", range(6))
def test_riemann_weak_waves_match_exact_solver(seed):
    """Weak waves: the two-shock approximation and its 3 iterations are
    exact to O(jump^3); compare against the exact solver."""
    rng = np.random.default_rng(seed)
    eps = 1e-3
    rl, ul, pl = 1.0, 0.2, 1.0
    rr, ur, pr = rl * (1 + eps * rng.uniform(-1, 1)), ul + eps * rng.uniform(-1, 1), pl * (1 + eps * rng.uniform(-1, 1))
    f = O.riemann_prob(G, rl, ul, pl, 0.0, 0.0, rr, ur, pr, 0.0, 0.0)
    e = ER.godunov_flux(rl, ul, pl, rr, ur, pr, G)
    for k in (0, 1, 4):
        assert abs(f[k] - e[k]) <= 1e-8 * max(1.0, abs(e[k]))


def test_exact_solver_reproduces_reference_sod_values():
    """Known answers the reference holds: the exact Sod solution printed by
    its notebook (tests/golden/sod_exact_notebook.json, extracted by
    tests/golden/extract_sod_notebook.py): star state, region densities and
    the four wave positions at t = 0.2."""
    import json
    import os
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sod_exact_notebook.json")))
    t, g, x0 = ref["case"]["t"], ref["case"]["gamma"], ref["case"]["x0"]
    pl, rl, ul = ref["case"]["left_p_rho_u"]
    pr, rr, ur = ref["case"]["right_p_rho_u"]
    p3, r3, u3 = ref["regions_p_rho_u"]["3"]
    p4, r4, u4 = ref["regions_p_rho_u"]["4"]
    p, u = ER.star_state(rl, ul, pl, rr, ur, pr, g)
    assert p == pytest.approx(p3, rel=1e-12) and p == pytest.approx(p4, rel=1e-12)
    assert u == pytest.approx(u3, rel=1e-12) and u == pytest.approx(u4, rel=1e-12)
    # region 3 (left of the contact) and region 4 (between contact and shock)
    xc = ref["positions"]["Contact Discontinuity"]
    assert ER.sample((xc - 0.01 - x0) / t, rl, ul, pl, rr, ur, pr, g)[0] == pytest.approx(r3, rel=1e-12)
    assert ER.sample((xc + 0.01 - x0) / t, rl, ul, pl, rr, ur, pr, g)[0] == pytest.approx(r4, rel=1e-12)
    # wave positions x0 + speed t (Toro 4.55-4.56, 4.59-4.60)
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    c3 = cl * (p / pl) ** ((g - 1.0) / (2.0 * g))
    ss = ur + cr * math.sqrt((g + 1.0) / (2.0 * g) * p / pr + (g - 1.0) / (2.0 * g))
    pos = ref["positions"]
    assert x0 + (ul - cl) * t == pytest.approx(pos["Head of Rarefaction"], rel=1e-12)
    assert x0 + (u - c3) * t == pytest.approx(pos["Foot of Rarefaction"], rel=1e-12)
    assert x0 + u * t == pytest.approx(xc, rel=1e-12)
    assert x0 + ss * t == pytest.approx(pos["Shock"], rel=1e-12)
    assert ss == pytest.approx(ref["us"], rel=1e-12)


def test_riemann_sod_interface_flux_near_exact():
    f = O.riemann_prob(G, 1.0, 0.0, 1.0, 0.0, 0.0, 0.125, 0.0, 0.1, 0.0, 0.0)
    e = ER.godunov_flux(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, G)
    # the Godunov state lies inside the left rarefaction's star region; the
    # two-shock approximation is within a few per cent there
    for k in (0, 1, 4):
        assert abs(f[k] - e[k]) <= 0.05 * abs(e[k])


def test_riemann_z_direction_quirk_is_present():
    """MUSCL+Riemann: the same 1D problem along x and y agree exactly; along
    z the left-state sound speed comes from T(k+1) (Riemann.h:316)."""
    def fields(x):  # smooth, so that the limited slopes are not zero
        tp = 2 * np.pi
        return (1 + 0.3 * np.sin(tp * x), 0.1 + 0 * x, 0 * x, 0 * x, 1 + 0.4 * np.cos(tp * x))
    lines = []
    for ax in range(3):
        u, dims = line_state(16, fields, axis=ax)
        r, _ = periodic_rhs(u, dims, P.scheme(flux_euler=0, visc=False))
        t = r[0].transpose(2, 1, 0)
        lines.append([t[:, 0, 0], t[0, :, 0], t[0, 0, :]][ax])
    assert np.array_equal(lines[0], lines[1])
    assert not np.allclose(lines[0], lines[2], rtol=1e-6, atol=0)


# ------------------------------------------------------------------ viscous
def test_viscous_face_flux_linear_fields():
    """Linear velocity and temperature with constant mu and k: the central
    differences are exact, so the viscous face flux equals the analytic
    Newtonian stress and Fourier flux."""
    cls = P.closures_reference()
    cls.visc_law, cls.visc_ref = 0, 2e-3
    cls.cond_law, cls.cond_ref = 0, 5e-2
    n = (8, 8, 8)
    lo, hi, ng = (0, 0, 0), (7, 7, 7), 3
    I, J, K = np.meshgrid(*(np.arange(-ng, 8 + ng) + 0.5 for _ in range(3)), indexing="ij")
    dx = 0.1
    X, Y, Z = I * dx, J * dx, K * dx
    a = dict(ux=0.3, uy=0.1, vy=0.2, vz=-0.1, wx=0.05, Tx=20.0)
    uu = a["ux"] * X + a["uy"] * Y
    vv = a["vy"] * Y + a["vz"] * Z
    ww = a["wx"] * X
    T = 300.0 + a["Tx"] * X
    rho = np.ones_like(T)
    p = rho * cls.Rspec * T
    U = np.stack([rho, rho * uu, rho * vv, rho * ww, p / (cls.gamma - 1) + 0.5 * rho * (uu * uu + vv * vv + ww * ww)])
    state = P.to_fab_order(U)
    dxinv = [1 / dx] * 3
    fv = O.box_numflux(cls, P.scheme(flux_euler=1, order_keep=2, visc=True), lo, hi, ng, state, dxinv)
    fe = O.box_numflux(cls, P.scheme(flux_euler=1, order_keep=2, visc=False), lo, hi, ng, state, dxinv)
    mu, k = 2e-3, 5e-2
    div = a["ux"] + a["vy"]
    txx = 2 * mu * (a["ux"] - div / 3)
    txy = mu * (a["uy"] + 0.0)
    txz = mu * (0.0 + a["wx"])
    # x face at i = 4 (between cells 3 and 4), j = k = 2
    i, j, kk = 4, 2, 2
    xf, yc, zc = 4 * dx, 2.5 * dx, 2.5 * dx
    uf, vf, wf = a["ux"] * xf + a["uy"] * yc, a["vy"] * yc + a["vz"] * zc, a["wx"] * xf
    vis = [f[:, kk + ng, j + ng, i + ng] for f in (fv[0] - fe[0],)][0]
    expect = [0.0, -txx, -txy, -txz, -k * a["Tx"] - txx * uf - txy * vf - txz * wf]
    # the viscous part is the difference of two fluxes that include the
    # (much larger) Euler flux: compare to that scale
    scale = np.max(np.abs(fe[0][:, kk + ng, j + ng, i + ng]))
    np.testing.assert_allclose(vis, expect, rtol=1e-9, atol=1e-14 * scale)


# ------------------------------------------------------------------ SSPRK
@pytest.mark.parametrize("order_rk,stages_rk,R", [
    (3, 3, lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6),
    (3, 4, lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 48),
    (2, 2, lambda z: 0.5 + 0.5 * (1 + z) ** 2),
    (2, 3, lambda z: 1 / 3 + 2 / 3 * (1 + z / 2) ** 3),
    (2, 5, lambda z: 1 / 5 + 4 / 5 * (1 + z / 4) ** 5),
    (1, 1, lambda z: 1 + z),
    (-2, 2, lambda z: 1 + z + z ** 2 / 2),
])
def test_rk_stability_polynomials(order_rk, stages_rk, R):
    for z in (-0.3, -1.1, 0.4, -2.0):
        dt = 0.1
        out = O.rk_linear(order_rk, stages_rk, z / dt, dt, [1.0, 2.0])
        assert out[0] == pytest.approx(R(z), rel=1e-13)
        assert out[1] == pytest.approx(2 * R(z), rel=1e-13)


def test_rk_rhs_only_returns_rhs():
    out = O.rk_linear(0, 0, -3.0, 0.1, [2.0])
    assert out[0] == -6.0


def test_ssprk_m2_is_second_order():
    for m in (2, 3, 4, 5):
        R = lambda z: O.rk_linear(2, m, z, 1.0, [1.0])[0]
        e1 = abs(R(0.02) - math.exp(0.02))
        e2 = abs(R(0.01) - math.exp(0.01))
        assert 7.0 < e1 / e2 < 9.0


# ------------------------------------------------------------------ ghost fill, estdt, conservation
def test_fill_patch_closed_forms():
    n = (8, 6, 4)
    g = P.geom(n, (0, 0, 0), (1, 1, 1), (0, 0, 1), lo_bc=(2, 4, 0), hi_bc=(2, 5, 0))
    boxes = P.decompose(n, (4, 6, 4))
    rng = np.random.default_rng(1)
    u = rng.uniform(0.5, 1.5, size=(5, 4, 6, 8))
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ol.fill_patch()
    f0 = ol.fabs[0]  # box 0 covers i = 0..3
    ng = 3
    A = lambda c, i, j, k: f0[c, k + ng, j + ng, i + ng]
    V = lambda c, i, j, k: u[c, k % 4, j, i]
    # x-lo outflow: foextrap copies cell 0
    assert A(0, -2, 2, 1) == V(0, 0, 2, 1)
    # interior neighbour box
    assert A(2, 5, 3, 2) == V(2, 5, 3, 2)
    # z periodic image
    assert A(1, 1, 1, -1) == V(1, 1, 1, 3)
    # y-lo slip wall: reflect_even for rho, reflect_odd for v (normal momentum)
    assert A(0, 1, -1, 0) == V(0, 1, 0, 0)
    assert A(2, 1, -2, 0) == -V(2, 1, 1, 0)
    assert A(1, 1, -2, 0) == V(1, 1, 1, 0)
    # y-hi no-slip wall: tangential momentum reflect_odd
    assert A(1, 1, 6, 0) == -V(1, 1, 5, 0)


def test_estdt_closed_form_with_umy_quirk():
    n = (4, 4, 4)
    g = P.geom(n, (0, 0, 0), (0.4, 0.8, 1.2), (1, 1, 1))
    rho, u, v, w, p = 1.3, 0.1, 0.2, 5.0, 2.0
    E = p / (G - 1) + 0.5 * rho * (u * u + v * v + w * w)
    U = np.zeros((5, 4, 4, 4))
    U[:] = np.array([rho, rho * u, rho * v, rho * w, E])[:, None, None, None]
    ol = O.OracleLevel(g, P.decompose(n, 4))
    ol.set_state(U)
    # the z velocity is taken from the y momentum (CNS_K.h:29)
    ke = 0.5 * rho * (u * u + v * v + v * v)
    ps = (G - 1) * (E - ke)
    c = math.sqrt(G * ps / rho)
    expect = min(0.1 / (u + c), 0.2 / (v + c), 0.3 / (v + c))
    assert ol.est_dt(CLS, 0.5) == pytest.approx(0.5 * expect, rel=1e-14)


def test_totals_conserved_over_steps():
    n = (16, 16, 16)
    cls = P.tgv_closures()
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    ol = O.OracleLevel(g, P.decompose(n, 8))
    ol.set_state(P.tgv_state(cls, g, n))
    t0 = ol.totals()
    for _ in range(3):
        ol.advance(cls, P.scheme(flux_euler=1, order_keep=4, visc=True), 4e-4)
    t1 = ol.totals()
    for c in (0, 4):
        assert abs(t1[c] - t0[c]) <= 1e-13 * abs(t0[c])


@pytest.mark.parametrize("flux", [2, 0], ids=["weno", "riemann"])
def test_sod_against_exact_solution(flux):
    """The full path (fill, RHS, SSPRK3, CFL dt with the 1.1x cap) on the
    Sod tube, 128 cells to t = 0.2, against the exact solution whose star
    state the reference notebook lists."""
    from amrsolver_amd.level import compute_new_dt
    n = (128, 4, 4)
    g = P.geom(n, (0, 0, 0), (1.0, 4 / 128, 4 / 128), (0, 1, 1), lo_bc=(2, 0, 0), hi_bc=(2, 0, 0))
    ol = O.OracleLevel(g, P.decompose(n, (64, 4, 4)))
    ol.set_state(P.sod_state(CLS, g, n))
    sch = P.scheme(flux_euler=flux, visc=False)
    t, dt, step = 0.0, None, 0
    while t < 0.2 - 1e-14 and step < 400:
        dt = compute_new_dt(ol.est_dt(CLS, 0.3), dt, t, 0.2, step == 0)
        ol.advance(CLS, sch, dt)
        t += dt
        step += 1
    assert t == pytest.approx(0.2, rel=1e-12)
    rho = ol.get_state()[0, 0, 0, :]
    x = (np.arange(128) + 0.5) / 128
    exact = np.array([ER.sample((xi - 0.5) / 0.2, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1)[0] for xi in x])
    l1 = np.mean(np.abs(rho - exact))
    assert l1 < 0.012, l1
    # plateau between contact (0.685) and shock (0.850) at rho_4 = 0.26557
    plateau = rho[(x > 0.72) & (x < 0.82)]
    assert np.all(np.abs(plateau - 0.2655737117053072) < 0.02)
