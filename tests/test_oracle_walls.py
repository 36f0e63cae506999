"""Pins of the oracle's bc 6 wall path (exm/Channel): the ghost cells its
bcnormal writes satisfy the physics they encode, and the wall terms stay
local to the rows next to the walls.  CPU only."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from oracle import oracle as O

N = (12, 16, 8)


@pytest.fixture
def channel():
    cls, g, sch, parm = P.wall_channel(N)
    u = P.wall_channel_state(cls, g, N, parm)
    O.set_problem(cls, parm["Tw"], parm["fx"])
    yield cls, g, sch, parm, u
    O.set_problem(None)


def _prims(f, cls):
    rho = f[0]
    v = f[1:4] / rho
    p = (cls.gamma - 1.0) * (f[4] - 0.5 * rho * np.sum(v * v, axis=0))
    return rho, v, p, p / (rho * cls.Rspec)


def test_wall_ghosts_are_no_slip_isothermal_and_keep_the_pressure(channel):
    """Every ghost layer mirrors the first interior cell: velocity reversed
    (no slip), same pressure (dp/dn = 0), and the first ghost's temperature
    is the linear extrapolation through the wall temperature (mean with the
    first interior cell = Tw)."""
    cls, g, sch, parm, u = channel
    ol = O.OracleLevel(g, P.decompose(N, (6, 8, 8)))
    ol.set_state(u)
    ol.fill_patch()
    ng = 3
    for (lo, hi), f in zip(ol.boxes, ol.fabs):
        for side in (0, 1):
            if (side == 0 and lo[1] != 0) or (side == 1 and hi[1] != N[1] - 1):
                continue
            jint = ng if side == 0 else f.shape[2] - ng - 1
            _, vi, pi_, Ti = _prims(f[:, :, jint:jint + 1, :], cls)
            for layer in range(1, ng + 1):
                jg = jint - layer if side == 0 else jint + layer
                _, vg, pg, Tg = _prims(f[:, :, jg:jg + 1, :], cls)
                scale = np.max(np.abs(vi)) + 1e-300
                assert np.max(np.abs(vg + vi)) <= 1e-12 * scale
                assert np.max(np.abs(pg - pi_) / pi_) <= 1e-12
                if layer == 1:
                    T_mean = 0.5 * (Tg + Ti)
                    ok = Tg > 50.0 * (1 + 1e-12)
                    assert np.max(np.abs(T_mean - parm["Tw"])[ok]) <= 1e-9 * parm["Tw"]


def test_wall_terms_are_local(channel):
    """dSdt with bc 6 walls equals dSdt with foextrap walls (bc 2) bit for bit
    in every cell three or more rows from the walls, and differs next to
    them (ghost values, order reduction, one-sided viscous derivatives)."""
    cls, g, sch, parm, u = channel
    sch0 = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=0, stages_rk=3, source=True)
    out = []
    for code in (6, 2):
        gg = P.geom(N, (0.0, 0.0, 0.0), (g.prob_hi[0], g.prob_hi[1], g.prob_hi[2]), (1, 0, 1),
                    lo_bc=(0, code, 0), hi_bc=(0, code, 0))
        ol = O.OracleLevel(gg, P.decompose(N, (6, 8, 8)))
        ol.set_state(u)
        ol.advance(cls, sch0, 1e-7)  # order_rk 0: the new state is dSdt
        out.append(ol.get_state())
    inner = slice(3, N[1] - 3)
    assert np.array_equal(out[0][:, :, inner, :], out[1][:, :, inner, :])
    assert not np.array_equal(out[0][:, :, 0, :], out[1][:, :, 0, :])
    assert not np.array_equal(out[0][:, :, -1, :], out[1][:, :, -1, :])


def test_body_force_adds_rho_fx_and_mx_fx(channel):
    """With rhs_source = 1 and no source array, dSdt gains rho*fx on rho*u and
    (rho*u)*fx on rho*E exactly, and nothing elsewhere."""
    cls, g, sch, parm, u = channel
    out = []
    for src in (True, False):
        s0 = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=0, stages_rk=3, source=src)
        ol = O.OracleLevel(g, P.decompose(N, (6, 8, 8)))
        ol.set_state(u)
        ol.advance(cls, s0, 1e-7)
        out.append(ol.get_state())
    d = out[0] - out[1]
    for n in (0, 2, 3):
        assert np.array_equal(out[0][n], out[1][n])
    assert np.allclose(d[1], u[0] * parm["fx"], rtol=1e-12, atol=1e-12 * np.max(np.abs(out[1][1])))
    assert np.allclose(d[4], u[1] * parm["fx"], rtol=1e-12, atol=1e-12 * np.max(np.abs(out[1][4])))


def test_walls_need_the_problem_parameters():
    cls, g, sch, parm = P.wall_channel(N)
    O.set_problem(None)
    ol = O.OracleLevel(g, P.decompose(N, (6, 8, 8)))
    ol.set_state(P.wall_channel_state(cls, g, N, parm))
    with pytest.raises(RuntimeError, match="wall"):
        ol.fill_patch()
