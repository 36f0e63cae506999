"""Method of manufactured solutions (exm/MMS_euler) on the HIP path: with the
exact flux divergence registered as the rhs_source term, amr_rhs returns the
truncation error; its L2 norm must fall at the KEEP order (2, 4, 6) under
grid refinement, as exm/MMS_euler/convergence_order.py measures.  Also the
source path itself against the oracle (parity 1e-12)."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err
import mms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


@pytest.mark.parametrize("order", [2, 4, 6])
def test_mms_keep_order(ctx, order):
    from amrsolver_amd import level as L
    cls = P.closures_reference()
    sch = P.scheme(flux_euler=1, order_keep=order, visc=False, order_rk=0, stages_rk=0, source=True)
    ns = [16, 32, 64]
    l2 = []
    for n in ns:
        print(f"case mms keep{order} n={n}", flush=True)
        g, u = mms.mms_state(cls, n)
        lev = L.Level(g, P.decompose((n, n, n), 32), cls, sch, ctx=ctx)
        S, D, Q = lev.new_state(), lev.new_state(ng=0), lev.new_state(ng=0)
        S.upload(u)
        Q.upload(mms.mms_source(cls, n))
        lev.set_source(Q)
        lev.fill_patch(S)
        lev.compute_rhs(S, D)
        l2.append(mms.norms(D.download())[0])
        lev.close()
    r = mms.rates(l2, ns)
    print("L2 errors", [e.tolist() for e in l2], "rates", [x.tolist() for x in r], flush=True)
    for c in range(5):
        assert r[-1][c] > order - 0.3, f"keep{order} comp {c}: rate {r[-1][c]:.3f}"


@pytest.mark.parametrize("tiling", [(16, 16, 16), (-1, 0, 0)], ids=["fused", "passes"])
def test_source_parity(ctx, tiling):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    cls = P.closures_reference()
    n = (32, 16, 12)
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.5, 0.4), (1, 1, 1))
    boxes = P.decompose(n, (16, 16, 12))
    u = P.smooth_random_state(cls, n, 31)
    src = np.random.default_rng(4).standard_normal(u.shape) * np.max(np.abs(u), axis=(1, 2, 3))[:, None, None, None]
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, source=True)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    lev.set_tiling(*tiling)
    S, Q = lev.new_state(), lev.new_state(ng=0)
    S.upload(u)
    Q.upload(src)
    lev.set_source(Q)
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    for _ in range(2):
        lev.advance(S, 0.0, 1e-6)
        ol.advance(cls, sch, 1e-6, source=src)
    lev.set_tiling(0)
    assert rel_err(S.download(), ol.get_state()) <= 1e-12
