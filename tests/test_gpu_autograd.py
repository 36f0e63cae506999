"""The harness's torch.autograd bindings (amrsolver_amd/autograd.py): torch
losses built on a level's RHS or time steps differentiate through the
hand-written reverse mode; checked against the dual-number oracle's exact
directional derivatives (round-off bars, as in test_gpu_adjoint_exact)."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _case(ctx, rk=(3, 3)):
    from amrsolver_amd import level as L
    cls = P.closures_scaled_sutherland(0.02, T0=300.0)
    n = (16, 12, 10)
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.8, 0.7), (1, 1, 1))
    boxes = P.decompose(n, (8, 6, 10))
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=rk[0], stages_rk=rk[1])
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    u = P.smooth_random_state(cls, n, 7)
    rng = np.random.default_rng(17)
    v = rng.standard_normal(u.shape)
    for c in range(5):
        v[c] *= np.max(np.abs(u[c]))
    return lev, cls, g, boxes, sch, u, v


@pytest.mark.parametrize("layout", ["level", "boxes"])
def test_rhs_loss_gradient(ctx, layout):
    import torch
    from amrsolver_amd.autograd import level_rhs
    print(f"case autograd rhs {layout}", flush=True)
    lev, cls, g, boxes, sch, u, v = _case(ctx)
    ut = torch.tensor(u, device="cuda", requires_grad=True)
    y = level_rhs(ut, lev, layout)
    loss = (y * y).sum()  # d loss = 2 <y, J v>
    loss.backward()
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ol.set_tangent(v)
    rhs, jv = ol.rhs_tangent(cls, sch)
    assert np.max(np.abs(y.detach().cpu().numpy() - rhs)) <= 1e-12 * np.max(np.abs(rhs))
    want = 2.0 * float(np.sum(rhs * jv))
    got = float(np.sum(ut.grad.cpu().numpy() * v))
    scale = 2.0 * np.linalg.norm(rhs) * np.linalg.norm(jv)
    assert abs(got - want) <= 1e-12 * scale, f"{got:.17e} {want:.17e}"


@pytest.mark.parametrize("layout,rk,kcheck", [("level", (3, 3), 2), ("boxes", (3, 4), 3)])
def test_steps_loss_gradient(ctx, layout, rk, kcheck):
    import torch
    from amrsolver_amd.autograd import level_steps
    print(f"case autograd steps {layout} {rk} k={kcheck}", flush=True)
    lev, cls, g, boxes, sch, u, v = _case(ctx, rk)
    dt, nsteps = 2e-5, 4
    w = np.random.default_rng(5).standard_normal(u.shape)
    ut = torch.tensor(u, device="cuda", requires_grad=True)
    y = level_steps(ut, lev, dt, nsteps, kcheck, layout)
    loss = (y * torch.tensor(w, device="cuda")).sum()
    loss.backward()
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ol.set_tangent(v)
    for _ in range(nsteps):
        ol.advance_tangent(cls, sch, dt, nthreads=8)
    assert np.max(np.abs(y.detach().cpu().numpy() - ol.get_state()) /
                  np.abs(ol.get_state()).reshape(5, -1).max(axis=1)[:, None, None, None]) <= 1e-12
    want = float(np.sum(w * ol.get_tangent()))
    got = float(np.sum(ut.grad.cpu().numpy() * v))
    scale = np.linalg.norm(ut.grad.cpu().numpy()) * np.linalg.norm(v)
    assert abs(got - want) <= 1e-11 * scale, f"{got:.17e} {want:.17e}"
