"""Pins of the oracle's forward-mode (dual-number) instantiation, which the
reverse-mode GPU tests use as their exact directional derivative.  CPU only.

- the value part of every dual evaluation equals the plain double path bit
  for bit (same operations in the same order);
- the derivative part agrees with fourth-order central differences of the
  plain path (bar 1e-7 relative, the differences' own truncation/rounding);
- the derivative is linear in the direction (to round-off);
- the ghost fill of a tangent is the fill itself (it is linear).

The same templates instantiated on an operation-counting scalar give the
reference arithmetic's FP64 operation count; its values equal the plain path.
"""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from oracle import oracle as O

SCHEMES = [("keep2", dict(flux_euler=1, order_keep=2, visc=False)),
           ("keep4", dict(flux_euler=1, order_keep=4, visc=False)),
           ("keep6", dict(flux_euler=1, order_keep=6, visc=False)),
           ("keep4+visc", dict(flux_euler=1, order_keep=4, visc=True))]


def _box_case(seed=5):
    cls = P.closures_scaled_sutherland(0.05, T0=300.0)
    n = (10, 8, 6)
    lo, hi, ng = (0, 0, 0), tuple(x - 1 for x in n), 3
    rng = np.random.default_rng(seed)
    u = P.smooth_random_state(cls, tuple(x + 2 * ng for x in n), seed)
    scale = np.abs(u).reshape(5, -1).max(axis=1)[:, None, None, None]
    v = rng.standard_normal(u.shape) * scale
    dx = [1.0 / 0.1, 1.0 / 0.12, 1.0 / 0.09]
    return cls, lo, hi, ng, u, v, dx


def _fd4(f, h):
    return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_box_rhs_tangent_value_is_plain_path(name, kw):
    cls, lo, hi, ng, u, v, dx = _box_case()
    sch = P.scheme(**kw)
    d, _ = O.box_rhs_tangent(cls, sch, lo, hi, ng, u, v, dx)
    assert np.array_equal(d, O.box_rhs(cls, sch, lo, hi, ng, u, dx))


@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_box_rhs_tangent_matches_central_differences(name, kw):
    cls, lo, hi, ng, u, v, dx = _box_case()
    sch = P.scheme(**kw)
    _, dd = O.box_rhs_tangent(cls, sch, lo, hi, ng, u, v, dx)
    rng = np.random.default_rng(9)
    lam = rng.standard_normal(dd.shape)
    got = float(np.sum(lam * dd))
    best = min(abs(_fd4(lambda h: float(np.sum(lam * O.box_rhs(cls, sch, lo, hi, ng, u + h * v, dx))), h) - got)
               for h in (1e-4, 3e-4, 1e-3, 3e-3))
    assert best / abs(got) <= 1e-7, f"{name}: {best / abs(got):.2e}"


def test_box_rhs_tangent_is_linear_in_the_direction():
    cls, lo, hi, ng, u, v, dx = _box_case()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    w = np.random.default_rng(2).standard_normal(u.shape) * np.abs(u).reshape(5, -1).max(axis=1)[:, None, None, None]
    _, a = O.box_rhs_tangent(cls, sch, lo, hi, ng, u, v, dx)
    _, b = O.box_rhs_tangent(cls, sch, lo, hi, ng, u, w, dx)
    _, c = O.box_rhs_tangent(cls, sch, lo, hi, ng, u, 2.0 * v - 0.5 * w, dx)
    scale = np.abs(c).reshape(5, -1).max(axis=1)[:, None, None, None]
    assert np.max(np.abs(c - (2.0 * a - 0.5 * b)) / scale) <= 1e-13


def test_tangent_rejects_weno_and_riemann():
    cls, lo, hi, ng, u, v, dx = _box_case()
    for fe in (0, 2):
        with pytest.raises(RuntimeError, match="tangent"):
            O.box_rhs_tangent(cls, P.scheme(flux_euler=fe), lo, hi, ng, u, v, dx)


def _level_case(bc):
    cls = P.closures_scaled_sutherland(0.05, T0=300.0)
    n = (12, 10, 8)
    if bc == "periodic":
        g = P.geom(n, (0, 0, 0), (1.2, 1.0, 0.8), (1, 1, 1))
    else:  # slip wall / no-slip wall in y, foextrap in x, periodic z
        g = P.geom(n, (0, 0, 0), (1.2, 1.0, 0.8), (0, 0, 1), lo_bc=(2, 4, 0), hi_bc=(2, 5, 0))
    boxes = P.decompose(n, (6, 5, 8))
    u = P.smooth_random_state(cls, n, 4)
    rng = np.random.default_rng(8)
    v = rng.standard_normal(u.shape) * np.abs(u).reshape(5, -1).max(axis=1)[:, None, None, None]
    return cls, g, boxes, u, v


@pytest.mark.parametrize("bc", ["periodic", "walls"])
def test_fill_patch_tangent_is_the_fill(bc):
    cls, g, boxes, u, v = _level_case(bc)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ol.set_tangent(v)
    ol.fill_patch_tangent()
    got = [t.copy() for t in ol.tfabs]
    want = O.OracleLevel(g, boxes)
    want.set_state(v)
    want.fill_patch()
    for a, b in zip(got, want.fabs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("bc", ["periodic", "walls"])
@pytest.mark.parametrize("rk", [(3, 3), (3, 4), (2, 3)], ids=["ssprk33", "ssprk43", "rk2m3"])
def test_advance_tangent(bc, rk):
    cls, g, boxes, u, v = _level_case(bc)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=rk[0], stages_rk=rk[1])
    dt = 2e-5
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ol.set_tangent(v)
    for _ in range(2):
        ol.advance_tangent(cls, sch, dt)
    plain = O.OracleLevel(g, boxes)
    plain.set_state(u)
    for _ in range(2):
        plain.advance(cls, sch, dt)
    assert np.array_equal(ol.get_state(), plain.get_state())
    dv = ol.get_tangent()
    lam = np.random.default_rng(1).standard_normal(u.shape)
    got = float(np.sum(lam * dv))

    def f(h):
        o = O.OracleLevel(g, boxes)
        o.set_state(u + h * v)
        for _ in range(2):
            o.advance(cls, sch, dt)
        return float(np.sum(lam * o.get_state()))

    best = min(abs(_fd4(f, h) - got) for h in (1e-4, 3e-4, 1e-3, 3e-3))
    assert best / abs(got) <= 1e-7, f"{bc} {rk}: {best / abs(got):.2e}"


def test_operation_counting_scalar():
    """The counting instantiation (the reference arithmetic's flop count)
    returns the plain path's values bit for bit, and its counts depend on the
    scheme and the grid only, not on the data."""
    import math
    cls = P.tgv_closures()
    n = (16, 16, 16)
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=3)
    counts = []
    for seed in (1, 2):
        u = P.smooth_random_state(cls, n, seed)
        a = O.OracleLevel(g, P.decompose(n, 8))
        a.set_state(u)
        counts.append(a.advance_flops(cls, sch, 1e-5))
        b = O.OracleLevel(g, P.decompose(n, 8))
        b.set_state(u)
        b.advance(cls, sch, 1e-5)
        assert np.array_equal(a.get_state(), b.get_state())
    assert counts[0] == counts[1]
    assert counts[0]["div"] > 0 and counts[0]["sqrt"] > 0


@pytest.mark.parametrize("order,name", [(4, "FP64_OPS_PER_CELL_STAGE"), (6, "FP64_OPS_KEEP6V")],
                         ids=["keep4v", "keep6v"])
def test_bench_operation_count_constant(order, name):
    """bench.py's FP64 operations per cell-stage are the counted values for
    its workload (TGV, one 64^3 box, KEEP4 or KEEP6 + viscous, SSPRK(3,3))."""
    import math
    import bench
    cls = P.tgv_closures()
    n = (64, 64, 64)
    g = P.geom(n, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    sch = P.scheme(flux_euler=1, order_keep=order, visc=True, order_rk=3, stages_rk=3)
    ol = O.OracleLevel(g, P.decompose(n, 64))
    ol.set_state(P.tgv_state(cls, g, n))
    c = ol.advance_flops(cls, sch, 5e-5)
    ops = (c["add"] + c["mul"] + c["div"] + c["sqrt"]) / (n[0] * n[1] * n[2] * 3)
    assert abs(ops - getattr(bench, name)) <= 0.5, ops


@pytest.mark.parametrize("flux,name", [(2, "FP64_OPS_SOD_W"), (0, "FP64_OPS_SOD_R")], ids=["weno", "muscl"])
def test_counting_weno_and_riemann(flux, name):
    """The counting instantiation of the WENO and MUSCL + Riemann paths
    returns the plain path's values bit for bit, and bench.py's operations
    per cell-stage of the Sod workloads are its counts."""
    import bench
    cls = P.closures_reference()
    n = (256, 4, 4)
    g = P.geom(n, (0, 0, 0), (1.0, 4 / 256, 4 / 256), (0, 1, 1), lo_bc=(2, 0, 0), hi_bc=(2, 0, 0))
    boxes = P.decompose(n, (64, 4, 4))
    u = P.sod_state(cls, g, n)
    sch = P.scheme(flux_euler=flux, visc=False)
    a = O.OracleLevel(g, boxes)
    a.set_state(u)
    c = a.advance_flops(cls, sch, 1e-3)
    b = O.OracleLevel(g, boxes)
    b.set_state(u)
    b.advance(cls, sch, 1e-3)
    assert np.array_equal(a.get_state(), b.get_state())
    assert c["div"] > 0 and c["sqrt"] > 0 and c["cmp"] > 0
    ops = (c["add"] + c["mul"] + c["div"] + c["sqrt"]) / (n[0] * n[1] * n[2] * 3)
    assert abs(ops - getattr(bench, name)) <= 0.5, ops
