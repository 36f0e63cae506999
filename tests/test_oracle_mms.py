"""Oracle pin: the method of manufactured solutions of exm/MMS_euler (source =
exact Euler flux divergence, prob.h:219-283) gives L2 truncation-error rates
at the KEEP orders 2, 4, 6 (convergence_order.py:161-186 computes the same
log-ratio rates)."""
import pytest

from amrsolver_amd import problems as P
import mms


@pytest.mark.parametrize("order", [2, 4, 6])
def test_oracle_mms_rates(order):
    from oracle import oracle as O
    cls = P.closures_reference()
    sch = P.scheme(flux_euler=1, order_keep=order, visc=False, order_rk=0, stages_rk=0, source=True)
    ns = [8, 16, 32]
    l2 = []
    for n in ns:
        g, u = mms.mms_state(cls, n)
        ol = O.OracleLevel(g, [((0, 0, 0), (n - 1, n - 1, n - 1))])
        ol.set_state(u)
        l2.append(mms.norms(ol.rhs(cls, sch, source=mms.mms_source(cls, n)))[0])
    r = mms.rates(l2, ns)
    for c in range(5):
        assert abs(r[-1][c] - order) < 0.2, f"keep{order} comp {c}: rate {r[-1][c]:.3f}"


def test_oracle_source_is_added_after_assembly():
    """rhs with a source = rhs without it + the source, exactly."""
    import numpy as np
    from oracle import oracle as O
    cls = P.closures_reference()
    n = (12, 10, 8)
    g = P.geom(n, (0, 0, 0), (1.0, 0.8, 0.6), (1, 1, 1))
    u = P.smooth_random_state(cls, n, 3)
    src = np.random.default_rng(1).standard_normal(u.shape)
    ol = O.OracleLevel(g, P.decompose(n, 6))
    ol.set_state(u)
    a = ol.rhs(cls, P.scheme(flux_euler=1, order_keep=4, visc=True))
    b = ol.rhs(cls, P.scheme(flux_euler=1, order_keep=4, visc=True, source=True), source=src)
    assert np.array_equal(b, a + src)
