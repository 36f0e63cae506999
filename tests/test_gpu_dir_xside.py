"""Directional passes on levels whose boxes span their periodic y and z
extents: the x pass reads the cells beyond each box's x ends from the
neighbour box (or its periodic image) or by the physical boundary's rule
(foextrap, reflect even/odd), and a step fills no ghost cell.  The steps
must not depend on what the ghost cells hold (NaN there), and must equal,
bit for bit, the same problem decomposed so that the passes read filled
ghost cells, and the oracle to 1e-12."""
import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import rel_err

pytestmark = pytest.mark.gpu

CASES = [
    # name, n, lo_bc, hi_bc, periodic, boxes with x ends read directly, boxes that need fills
    ("sod_foextrap", (128, 8, 8), (2, 0, 0), (2, 0, 0), (0, 1, 1), (32, 8, 8), (32, 4, 8)),
    ("slipwall_reflect", (96, 8, 8), (4, 0, 0), (3, 0, 0), (0, 1, 1), (24, 8, 8), (24, 8, 4)),
    ("periodic_x", (64, 8, 8), (0, 0, 0), (0, 0, 0), (1, 1, 1), (16, 8, 8), (16, 4, 8)),
    # a 4-deep periodic y (repeated pencil rows in the y pass) with z 8 deep, and the reverse
    ("rows4_y", (64, 4, 8), (2, 0, 0), (3, 0, 0), (0, 1, 1), (16, 4, 8), (16, 4, 4)),
    ("rows4_z", (64, 8, 4), (4, 0, 0), (2, 0, 0), (0, 1, 1), (16, 8, 4), (16, 4, 4)),
]


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _state(cls, g, n, seed):
    u = P.smooth_random_state(cls, n, seed)
    if n[0] >= 64:  # a Sod-like jump across the middle on top
        u[:, :, :, n[0] // 2:] *= 0.5
    return u


def _nan_ghosts(lev, S, u):
    for b, (lo, hi) in enumerate(lev.boxes):
        a = np.full(S.tensors[b].shape, np.nan)
        a[:, 3:-3, 3:-3, 3:-3] = u[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        S.upload_box(b, a)


@pytest.mark.parametrize("flux", [2, 0], ids=["weno", "muscl"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_x_ends_read_directly(ctx, case, flux):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    name, n, lo_bc, hi_bc, per, mg_direct, mg_fill = case
    print(f"case dir xside {name} flux {flux}", flush=True)
    cls = P.closures_reference()
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, n[1] / n[0], n[2] / n[0]), per, lo_bc=lo_bc, hi_bc=hi_bc)
    u = _state(cls, g, n, 90 + flux)
    sch = P.scheme(flux_euler=flux, visc=False, order_rk=3, stages_rk=3)
    out = {}
    dt = None
    # the fill-based decomposition first: its CFL estimate (0.3) sets the step
    for mg, nan in ((mg_fill, False), (mg_direct, True)):
        boxes = P.decompose(n, mg)
        lev = L.Level(g, boxes, cls, sch, ctx=ctx)
        S = lev.new_state(layout="boxes")
        if nan:
            _nan_ghosts(lev, S, u)
        else:
            S.upload(u)
            lev.fill_patch(S)
            dt = lev.est_time_step(S, 0.3)
        for _ in range(3):
            lev.advance(S, 0.0, dt)
        out[nan] = S.download()
        lev.close()
    assert np.isfinite(out[False]).all()
    assert np.isfinite(out[True]).all(), "a ghost cell was read"
    assert np.array_equal(out[True], out[False])
    ol = O.OracleLevel(g, P.decompose(n, mg_direct))
    ol.set_state(u)
    for _ in range(3):
        ol.advance(cls, sch, dt)
    err = rel_err(out[True], ol.get_state())
    print(f"  max normalised error {err:.3e}", flush=True)
    assert err <= 1e-12
