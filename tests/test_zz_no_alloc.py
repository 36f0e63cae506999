"""amr_est_dt_noalloc (the CFL reduction that allocates nothing) against
amr_est_dt through the C ABI: equal time steps on two calls in a row, the
second on a smaller input (fewer, smaller boxes), and the box-count limit
rejected before any device work."""
import ctypes
import math

import pytest

from amrsolver_amd import capi
from amrsolver_amd import problems as P

pytestmark = pytest.mark.gpu


def _est_both(lev, S, cfl):
    lib = capi.load()
    a = ctypes.c_double()
    b = ctypes.c_double()
    rc = lib.amr_est_dt(lev.h, ctypes.byref(lev.closures), S.fabs, cfl, ctypes.byref(a))
    assert rc == 0, capi.load().amr_last_error()
    rc = lib.amr_est_dt_noalloc(lev.h, ctypes.byref(lev.closures), S.fabs, cfl, ctypes.byref(b))
    assert rc == 0, capi.load().amr_last_error()
    return a.value, b.value


def test_est_dt_noalloc_matches_est_dt():
    from amrsolver_amd import level as L
    ctx = L.Context(0)
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
    # first call: TGV 64^3 in 27 uneven boxes, then a smaller Sod-like level
    n1 = (64, 64, 64)
    g1 = P.geom(n1, (0, 0, 0), (2 * math.pi,) * 3, (1, 1, 1))
    lev1 = L.Level(g1, P.decompose(n1, (24, 24, 24)), cls, sch, ctx=ctx)
    S1 = lev1.new_state(layout="boxes")
    S1.upload(P.tgv_state(cls, g1, n1))
    print(f"case noalloc est_dt: {len(lev1.boxes)} boxes of 64^3", flush=True)
    a1, b1 = _est_both(lev1, S1, 0.3)
    assert a1 > 0 and a1 == b1, (a1, b1)
    n2 = (32, 4, 4)
    g2 = P.geom(n2, (0, 0, 0), (1.0, 4 / 32, 4 / 32), (0, 1, 1), lo_bc=(2, 0, 0), hi_bc=(2, 0, 0))
    cls2 = P.closures_reference()
    lev2 = L.Level(g2, P.decompose(n2, (16, 4, 4)), cls2, P.scheme(flux_euler=2, visc=False), ctx=ctx)
    S2 = lev2.new_state()
    S2.upload(P.sod_state(cls2, g2, n2))
    print(f"case noalloc est_dt: {len(lev2.boxes)} boxes of 32x4x4", flush=True)
    a2, b2 = _est_both(lev2, S2, 0.5)
    assert a2 > 0 and a2 == b2, (a2, b2)
    # the first level again after the smaller call
    a3, b3 = _est_both(lev1, S1, 0.3)
    assert a3 == a1 and b3 == b1
    lev2.close()
    lev1.close()
    ctx.close()


def test_est_dt_noalloc_rejects_too_many_boxes():
    from amrsolver_amd import level as L
    ctx = L.Context(0)
    n = (8200, 1, 1)
    g = P.geom(n, (0, 0, 0), (1.0, 1.0 / 8200, 1.0 / 8200), (1, 0, 0), lo_bc=(0, 2, 2), hi_bc=(0, 2, 2))
    boxes = [((i, 0, 0), (i, 0, 0)) for i in range(8200)]
    lev = L.Level(g, boxes, P.closures_reference(), P.scheme(flux_euler=2, visc=False), ctx=ctx)
    fabs = (capi.Fab * len(boxes))()
    out = ctypes.c_double()
    rc = capi.load().amr_est_dt_noalloc(lev.h, ctypes.byref(lev.closures), fabs, 0.3, ctypes.byref(out))
    assert rc == -1 and b"at most 8192" in capi.load().amr_last_error()
    lev.close()
    ctx.close()
