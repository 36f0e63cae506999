"""Runs the oracle's entry points on small cases for the sanitizer check
(tests/test_oracle_asan.py starts it as `python -m oracle.asan_check` from
the repository root, with the AddressSanitizer runtime
preloaded and ORACLE_LIB pointing at oracle/_build/asan/libcns_oracle.so).
Test infrastructure only: every scheme's RHS and steps, the tangent and
operation-counting instantiations, threads inside a box, the fills, the
coarse-fine coupling and the point functions."""
import os

import numpy as np

from oracle import oracle as O  # noqa: E402

O.LIBPATH = os.environ["ORACLE_LIB"]
O._lib = None

from amrsolver_amd import problems as P  # noqa: E402


def main():
    cls = P.closures_scaled_sutherland(0.02, T0=300.0)
    n = (12, 10, 8)
    for per, lo_bc, hi_bc in (((1, 1, 1), (0, 0, 0), (0, 0, 0)), ((0, 0, 1), (2, 4, 0), (2, 5, 0))):
        g = P.geom(n, (0, 0, 0), (1.0, 0.8, 0.7), per, lo_bc=lo_bc, hi_bc=hi_bc)
        boxes = P.decompose(n, (6, 5, 8))
        u = P.smooth_random_state(cls, n, 3)
        for kw in (dict(flux_euler=1, order_keep=4, visc=True), dict(flux_euler=1, order_keep=6, visc=False),
                   dict(flux_euler=2, visc=True), dict(flux_euler=0, visc=True), dict(flux_euler=2, jst=True)):
            ol = O.OracleLevel(g, boxes)
            ol.set_state(u)
            sch = P.scheme(**kw)
            ol.rhs(cls, sch)
            ol.advance(cls, sch, 1e-6, nthreads=2)
            ol.advance(cls, sch, 1e-6, nthreads=16)  # more threads than boxes: threads inside each box
            ol.est_dt(cls, 0.3)
            ol.totals()
        sch = P.scheme(flux_euler=1, order_keep=4, visc=True)
        ol = O.OracleLevel(g, boxes)
        ol.set_state(u)
        ol.set_tangent(u)
        ol.rhs_tangent(cls, sch)
        ol.advance_tangent(cls, sch, 1e-6, nthreads=16)
        ol.advance_flops(cls, sch, 1e-6)
    q = O.box_prims(cls, (0, 0, 0), tuple(x - 1 for x in n), 0, np.ascontiguousarray(u))
    assert q.shape == (6,) + u.shape[1:]
    O.weno5js([0.1, 0.4, 0.3, 0.9, 0.2])
    O.riemann_prob(1.4, 1.0, 0.5, 1.0, 0.1, 0.0, 0.5, -0.2, 0.4, 0.0, 0.3)
    O.riemann_face(cls, 2, np.linspace(0.5, 2.0, 18), np.linspace(-0.01, 0.01, 10))
    print("oracle sanitizer check: ok")


if __name__ == "__main__":
    main()
