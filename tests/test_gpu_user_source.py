"""amr_set_user_source: a problem's user_source (CNS::compute_rhs calls it
after the flux assembly, src/tim/CNS_advance.cpp:397-407) given as HIP text
and compiled at run time.  The channel's body force written as the
reference's user_source (exm/Channel/prob.h:189-193) equals the library's
built-in body force bit for bit over whole SSPRK steps (both layouts, also
captured in a graph); a nonlinear state-dependent source gives amr_rhs equal
to the oracle's dSdt plus the same source evaluated in numpy, bit for bit;
compile errors are reported."""
import math

import numpy as np
import pytest

from amrsolver_amd import capi
from amrsolver_amd import problems as P

pytestmark = pytest.mark.gpu

CHANNEL_SOURCE = r"""
// exm/Channel/prob.h:189-193 with ProbParm::fx as pparm[0]
__device__ void user_source(int i, int j, int k, const StateView& state, const RhsView& rhs, const double* pparm,
                            const amr_closures_t& cls, const double* dx) {
  rhs(i, j, k, UMX) += state(i, j, k, URHO) * pparm[0];
  rhs(i, j, k, UET) += state(i, j, k, UMX) * pparm[0];
}
"""

NONLINEAR_SOURCE = r"""
__device__ void user_source(int i, int j, int k, const StateView& state, const RhsView& rhs, const double* pparm,
                            const amr_closures_t& cls, const double* dx) {
  const double rho = state(i, j, k, URHO);
  const double ke = 0.5 * (state(i, j, k, UMX) * state(i, j, k, UMX) + state(i, j, k, UMY) * state(i, j, k, UMY) +
                           state(i, j, k, UMZ) * state(i, j, k, UMZ)) / rho;
  const double p = (cls.gamma - 1.0) * (state(i, j, k, UET) - ke);
  rhs(i, j, k, UET) = rhs(i, j, k, UET) - pparm[0] * (p - pparm[1]);
  rhs(i, j, k, UMX) = rhs(i, j, k, UMX) + pparm[2] * rho * (j + 0.5) * dx[1];
}
"""


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _setup(ctx, sch):
    from amrsolver_amd import level as L
    cls = P.closures_reference()
    n = (32, 24, 16)
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.75, 0.5), (1, 1, 1))
    boxes = P.decompose(n, (16, 12, 16))
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    return lev, cls, g, boxes, P.smooth_random_state(cls, n, 29)


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("name,kw", [("keep4v", dict(flux_euler=1, order_keep=4, visc=True)),
                                     ("weno", dict(flux_euler=2, visc=False))], ids=["keep4v", "weno"])
def test_channel_source_as_code_equals_the_builtin_body_force(ctx, name, kw, layout):
    print(f"case user source body force {name} {layout}", flush=True)
    sch = P.scheme(source=True, **kw)
    lev, cls, g, boxes, u = _setup(ctx, sch)
    fx = 0.37
    lev.set_problem(body_force_x=fx)
    A = lev.new_state(layout=layout)
    A.upload(u)
    for _ in range(3):
        lev.advance(A, 0.0, 1e-5)
    want = A.download()
    lev.set_user_source(CHANNEL_SOURCE, [fx])
    B = lev.new_state(layout=layout)
    B.upload(u)
    for _ in range(3):
        lev.advance(B, 0.0, 1e-5)
    assert np.array_equal(B.download(), want)
    # captured steps replay equal to eager ones
    C = lev.new_state(layout=layout)
    C.upload(u)
    for rep in range(2):
        lev.advance_graph(C, 1e-5, 3)
        if rep == 0:
            assert np.array_equal(C.download(), want)
    D = lev.new_state(layout=layout)
    D.upload(u)
    for _ in range(6):
        lev.advance(D, 0.0, 1e-5)
    assert np.array_equal(C.download(), D.download())
    # removed: back to the built-in body force
    lev.set_user_source(None)
    E = lev.new_state(layout=layout)
    E.upload(u)
    for _ in range(3):
        lev.advance(E, 0.0, 1e-5)
    assert np.array_equal(E.download(), want)
    lev.close()


def test_nonlinear_source_rhs_equals_oracle_plus_source(ctx):
    from oracle import oracle as O
    print("case user source nonlinear rhs", flush=True)
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, source=True)
    lev, cls, g, boxes, u = _setup(ctx, sch)
    params = [0.05, 2.0e5, 3.0]
    lev.set_user_source(NONLINEAR_SOURCE, params)
    S = lev.new_state()
    S.upload(u)
    lev.fill_patch(S)
    D = lev.new_state(ng=0)
    lev.compute_rhs(S, D)
    got = D.download()
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    want = ol.rhs(cls, P.scheme(flux_euler=1, order_keep=4, visc=True, source=False))
    rho, mx, my, mz, E = u
    ke = 0.5 * (mx * mx + my * my + mz * mz) / rho
    p = (cls.gamma - 1.0) * (E - ke)
    want[4] = want[4] - params[0] * (p - params[1])
    jj = np.arange(u.shape[2], dtype=np.float64)[None, :, None]
    want[1] = want[1] + params[2] * rho * (jj + 0.5) * g.dx[1]
    assert np.array_equal(got, want)
    lev.close()


def test_compile_errors_are_reported(ctx):
    print("case user source compile error", flush=True)
    lev, cls, g, boxes, u = _setup(ctx, P.scheme(source=True))
    with pytest.raises(capi.AmrError, match="does not compile"):
        lev.set_user_source("__device__ void user_source(int i) { undefined_name = 1; }")
    lev.close()
