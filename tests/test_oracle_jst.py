"""Pins of the oracle's JST artificial dissipation (art_diss = 1;
CNS_advance.cpp:235-265, CNS_hydro_K.h:38-104).  CPU only.

- a constant state gets no dissipation (bit for bit, for values the damping
  stencil sums exactly);
- the face fluxes equal the plain scheme's minus an independent numpy
  statement of JSTflux, bit for bit (same operations in the same order);
- the reference's quirks, each of which a textbook JST would not show:
  the z damping stencil reads cons(k-1) three times (a linear profile in z
  is damped; in x and y it is not), every direction scales by the x spectral
  radius |u| + c (a y velocity does not enter the y term), and the sensor
  and spectral radius are zero outside the box (one box and two boxes differ
  at the shared face).
"""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from oracle import oracle as O

NG = 3


def _closures():
    return P.closures_scaled_sutherland(0.01, T0=300.0)


def _sensor(pp, pl, pr):
    pjst = pr + 2.0 * pp + pl
    ptvd = np.abs(pr - pp) + np.abs(pp - pl)
    return np.abs(2.0 * (pr - 2.0 * pp + pl) / (pjst + ptvd + 1.0e-40))


def _jst_faces(c, u, q, n):
    """JST term of every face of the all-nodal box, (3, 5, nz+1, ny+1, nx+1),
    from the ghosted state u and primitives q of a box with n cells."""
    nx, ny, nz = n
    g = NG
    # sensor / spectral radius on the valid box, zero on grow(1)
    sen = np.zeros((3, nz + 2, ny + 2, nx + 2))
    lam = np.zeros((3, nz + 2, ny + 2, nx + 2))
    P_ = q[5]
    V = (slice(g, g + nz), slice(g, g + ny), slice(g, g + nx))
    pp = P_[V]
    sh = lambda a, dk, dj, di: a[g + dk:g + dk + nz, g + dj:g + dj + ny, g + di:g + di + nx]
    sen[0, 1:-1, 1:-1, 1:-1] = _sensor(pp, sh(P_, 0, 0, -1), sh(P_, 0, 0, 1))
    sen[1, 1:-1, 1:-1, 1:-1] = _sensor(pp, sh(P_, 0, -1, 0), sh(P_, 0, 1, 0))
    sen[2, 1:-1, 1:-1, 1:-1] = _sensor(pp, sh(P_, -1, 0, 0), sh(P_, 1, 0, 0))
    cs = np.sqrt(c.gamma * pp / q[0][V])
    lam[0, 1:-1, 1:-1, 1:-1] = np.abs(q[1][V]) + cs
    lam[1, 1:-1, 1:-1, 1:-1] = np.abs(q[2][V]) + cs
    lam[2, 1:-1, 1:-1, 1:-1] = np.abs(q[3][V]) + cs
    out = np.zeros((3, 5, nz + 1, ny + 1, nx + 1))
    N = (nz + 1, ny + 1, nx + 1)

    def U(comp, dk, dj, di):
        return u[comp, g + dk:g + dk + N[0], g + dj:g + dj + N[1], g + di:g + di + N[2]]

    def S(a, dk, dj, di):  # grow(1) arrays at nodal points
        return a[1 + dk:1 + dk + N[0], 1 + dj:1 + dj + N[1], 1 + di:1 + di + N[2]]

    for m in range(5):
        uijk = U(m, 0, 0, 0)
        # x
        dw = uijk - U(m, 0, 0, -1)
        sn = c.Cshock * np.maximum(S(sen[0], 0, 0, -1), S(sen[0], 0, 0, 0))
        rr = np.maximum(S(lam[0], 0, 0, -1), S(lam[0], 0, 0, 0))
        fd = U(m, 0, 0, 1) - 3.0 * U(m, 0, 0, 0) + 3.0 * U(m, 0, 0, -1) - U(m, 0, 0, -2)
        out[0, m] = (sn * dw - np.maximum(0.0, c.Cdamp - sn) * fd) * rr
        # y (component 0 of sensor and spectral radius)
        dw = uijk - U(m, 0, -1, 0)
        sn = c.Cshock * np.maximum(S(sen[0], 0, -1, 0), S(sen[0], 0, 0, 0))
        rr = np.maximum(S(lam[0], 0, -1, 0), S(lam[0], 0, 0, 0))
        fd = U(m, 0, 1, 0) - 3.0 * U(m, 0, 0, 0) + 3.0 * U(m, 0, -1, 0) - U(m, 0, -2, 0)
        out[1, m] = (sn * dw - np.maximum(0.0, c.Cdamp - sn) * fd) * rr
        # z (component 0; cons(k-1) in place of cons(k+1) and cons(k-2))
        dw = uijk - U(m, -1, 0, 0)
        sn = c.Cshock * np.maximum(S(sen[0], -1, 0, 0), S(sen[0], 0, 0, 0))
        rr = np.maximum(S(lam[0], -1, 0, 0), S(lam[0], 0, 0, 0))
        fd = U(m, -1, 0, 0) - 3.0 * U(m, 0, 0, 0) + 3.0 * U(m, -1, 0, 0) - U(m, -1, 0, 0)
        out[2, m] = (sn * dw - np.maximum(0.0, c.Cdamp - sn) * fd) * rr
    return out


def _nodal(f):
    """The all-nodal box [lo, hi+1] of a numflux FAB (ghosts NG)."""
    return f[:, NG:-NG, NG:-NG, NG:-NG]


def _state(cls, n, seed, shock=True):
    rng = np.random.default_rng(seed)
    u = P.smooth_random_state(cls, tuple(x + 2 * NG for x in n), seed)
    if shock:  # a pressure jump half-way in x activates the sensor
        u[4, :, :, (n[0] + 2 * NG) // 2:] *= 3.0
    u *= 1.0 + 1e-3 * rng.standard_normal(u.shape)
    return u


@pytest.mark.parametrize("kw", [dict(flux_euler=1, order_keep=4, visc=False), dict(flux_euler=2, visc=False),
                                dict(flux_euler=0, visc=False), dict(flux_euler=1, order_keep=2, visc=True)],
                         ids=["keep4", "weno", "muscl", "keep2v"])
def test_faces_are_plain_minus_jst(kw):
    cls = _closures()
    n = (10, 8, 6)
    lo, hi = (0, 0, 0), tuple(x - 1 for x in n)
    u = _state(cls, n, 3)
    dx = [8.0, 9.0, 7.0]
    plain = O.box_numflux(cls, P.scheme(**kw), lo, hi, NG, u, dx)
    jst = O.box_numflux(cls, P.scheme(jst=True, **kw), lo, hi, NG, u, dx)
    q = O.box_prims(cls, lo, hi, NG, u)
    term = _jst_faces(cls, u, q, n)
    if kw.get("visc"):
        # viscous face terms are added after JST: compare the Euler part only
        e_plain = O.box_numflux(cls, P.scheme(**dict(kw, visc=False)), lo, hi, NG, u, dx)
        e_jst = O.box_numflux(cls, P.scheme(jst=True, **dict(kw, visc=False)), lo, hi, NG, u, dx)
        for d in range(3):
            assert np.array_equal(_nodal(e_jst[d]), _nodal(e_plain[d]) - term[d])
            # and the viscous part is added on top in both cases
            vis = _nodal(plain[d]) - _nodal(e_plain[d])
            assert np.allclose(_nodal(jst[d]) - _nodal(e_jst[d]), vis, rtol=1e-12, atol=1e-12 * np.abs(vis).max())
        return
    for d in range(3):
        assert np.array_equal(_nodal(jst[d]), _nodal(plain[d]) - term[d])
    assert np.abs(term).max() > 0.0


def test_constant_state_gets_no_dissipation():
    cls = _closures()
    n = (8, 6, 5)
    lo, hi = (0, 0, 0), tuple(x - 1 for x in n)
    u = np.empty((5,) + tuple(x + 2 * NG for x in reversed(n)))
    # values whose damping stencil sums exactly (the reference's
    # u(i+1) - 3u(i) + 3u(i-1) - u(i-2) rounds to a few ulps otherwise)
    for m, v in enumerate((1.0, 0.5, -0.25, 0.125, 2.5e5)):
        u[m] = v
    dx = [1.0, 1.0, 1.0]
    for fe in (0, 1, 2):
        a = O.box_numflux(cls, P.scheme(flux_euler=fe, visc=False), lo, hi, NG, u, dx)
        b = O.box_numflux(cls, P.scheme(flux_euler=fe, visc=False, jst=True), lo, hi, NG, u, dx)
        for d in range(3):
            assert np.array_equal(a[d], b[d])


def _profile_state(cls, n, axis, slope, vel=(0.0, 0.0, 0.0)):
    """rho linear along `axis`, constant pressure and velocity."""
    shape = tuple(x + 2 * NG for x in reversed(n))
    idx = np.indices(shape)[2 - axis].astype(float)
    rho = 1.0 + slope * idx
    p = 1.0e5
    u = np.empty((5,) + shape)
    u[0] = rho
    for d in range(3):
        u[1 + d] = rho * vel[d]
    u[4] = p / (cls.gamma - 1.0) + 0.5 * rho * sum(v * v for v in vel)
    return u


def test_z_damping_reads_k_minus_1():
    """Density linear in z, pressure constant (sensor 0): the textbook third
    difference vanishes, the reference's z stencil gives 3 (u(k-1) - u(k))."""
    cls = _closures()
    n = (6, 6, 6)
    lo, hi = (0, 0, 0), tuple(x - 1 for x in n)
    for axis in range(3):
        u = _profile_state(cls, n, axis, 0.01)
        a = O.box_numflux(cls, P.scheme(flux_euler=1, visc=False), lo, hi, NG, u, [1.0] * 3)
        b = O.box_numflux(cls, P.scheme(flux_euler=1, visc=False, jst=True), lo, hi, NG, u, [1.0] * 3)
        diff = _nodal(a[axis])[0, 1:-1, 1:-1, 1:-1] - _nodal(b[axis])[0, 1:-1, 1:-1, 1:-1]  # interior faces, rho
        if axis < 2:  # the third difference of a linear profile: round-off only
            assert np.max(np.abs(diff)) <= 1e-12
        else:
            # faces k = 1..5: term = -(Cdamp * fd) * rr with fd = 3 (rho(k-1) - rho(k)) = -0.03 and
            # rr = max(c(k-1), c(k)) = c(k-1) (density grows with k), cell k-1 at array index NG + k - 1
            c_km1 = np.sqrt(cls.gamma * 1.0e5 / (1.0 + 0.01 * np.arange(NG, NG + 5)))
            want = -(cls.Cdamp * (-0.03)) * c_km1
            assert np.allclose(diff[:, 0, 0], want, rtol=1e-9)


def test_every_direction_scales_by_the_x_spectral_radius():
    """Density cubic in y, pressure constant (sensor 0), a y velocity: the y
    term is Cdamp * (third difference) * (|u_x| + c), not (|u_y| + c)."""
    cls = _closures()
    n = (6, 6, 6)
    lo, hi = (0, 0, 0), tuple(x - 1 for x in n)
    shape = tuple(x + 2 * NG for x in reversed(n))
    j = np.indices(shape)[1].astype(float)
    rho = 1.0 + 1e-3 * j ** 3
    vy = 80.0
    u = np.zeros((5,) + shape)
    u[0] = rho
    u[2] = rho * vy
    u[4] = 1.0e5 / (cls.gamma - 1.0) + 0.5 * rho * vy * vy
    a = O.box_numflux(cls, P.scheme(flux_euler=1, visc=False), lo, hi, NG, u, [1.0] * 3)
    b = O.box_numflux(cls, P.scheme(flux_euler=1, visc=False, jst=True), lo, hi, NG, u, [1.0] * 3)
    term = _nodal(a[1])[0, 1:-1, 1:-1, 1:-1] - _nodal(b[1])[0, 1:-1, 1:-1, 1:-1]  # rho, faces j = 1..n-1
    jj = np.arange(NG + 1, NG + n[1]).astype(float)  # array index of cell j
    fd = 1e-3 * 6.0  # third difference of j^3
    c_j = np.sqrt(cls.gamma * 1.0e5 / (1.0 + 1e-3 * jj ** 3))
    c_m = np.sqrt(cls.gamma * 1.0e5 / (1.0 + 1e-3 * (jj - 1.0) ** 3))
    want = -(cls.Cdamp * fd) * np.maximum(c_j, c_m)
    assert np.allclose(term[0, :, 0], want, rtol=1e-9)
    with_vy = -(cls.Cdamp * fd) * (np.maximum(c_j, c_m) + vy)
    assert np.all(np.abs(term[0, :, 0] - with_vy) > 0.5 * np.abs(with_vy - want))


def test_sensor_is_zero_outside_the_box():
    """One box and the same cells as two boxes differ at the shared x face
    (the sensor of the neighbour cell reads 0), as in the reference."""
    cls = _closures()
    n = (16, 4, 4)
    g = P.geom(n, (0, 0, 0), (1.0, 0.25, 0.25), (1, 1, 1))
    rng = np.random.default_rng(2)
    u = P.smooth_random_state(cls, n, 5)
    u[4, :, :, 8:12] *= 2.0  # pressure step at the shared face i = 8: sensors of cells 7 and 8 > 0
    u *= 1.0 + 1e-3 * rng.standard_normal(u.shape)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=False, jst=True)
    r = []
    for boxes in ([((0, 0, 0), (15, 3, 3))], [((0, 0, 0), (7, 3, 3)), ((8, 0, 0), (15, 3, 3))]):
        ol = O.OracleLevel(g, boxes)
        ol.set_state(u)
        ol.fill_patch()
        r.append(ol.rhs(cls, sch))
    d = np.abs(r[0] - r[1]).max(axis=(0, 1, 2))
    assert d[7] > 0 and d[8] > 0  # the cells on either side of the shared face
    assert np.all(d[:7] == 0) and np.all(d[9:] == 0)
