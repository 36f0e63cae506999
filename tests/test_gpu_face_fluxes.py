"""amr_face_fluxes, the numflxmf of CNS::compute_rhs
(src/tim/CNS_advance.cpp:218-346): bit for bit against the oracle's face
fluxes for every scheme (KEEP 2/4/6 +- viscous, WENO, MUSCL + Riemann, JST)
in both layouts, and their assembly dxinv*(F(i) - F(i+1)) summed over the
directions (numpy, in the reference's order) equal to amr_rhs bit for bit,
also next to the channel's bc 6 walls (order-reduced KEEP and wall viscous
fluxes)."""
import numpy as np
import pytest

from amrsolver_amd import capi
from amrsolver_amd import problems as P

pytestmark = pytest.mark.gpu

SCHEMES = [
    ("keep4v", dict(flux_euler=1, order_keep=4, visc=True)),
    ("keep2", dict(flux_euler=1, order_keep=2, visc=False)),
    ("keep6v", dict(flux_euler=1, order_keep=6, visc=True)),
    ("weno", dict(flux_euler=2, visc=False)),
    ("wenov", dict(flux_euler=2, visc=True)),
    ("muscl", dict(flux_euler=0, visc=False)),
    ("keep4_jst", dict(flux_euler=1, order_keep=4, visc=False, jst=True)),
]


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _assemble(F, lev, b, dxi):
    fx, fy, fz = (F.download(d, b) for d in range(3))
    return dxi[0] * (fx[:, :, :, :-1] - fx[:, :, :, 1:]) + dxi[1] * (fy[:, :, :-1, :] - fy[:, :, 1:, :]) + \
        dxi[2] * (fz[:, :-1, :, :] - fz[:, 1:, :, :])


@pytest.mark.parametrize("layout", ["level", "boxes"])
@pytest.mark.parametrize("name,kw", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_face_fluxes_match_the_oracle(ctx, name, kw, layout):
    from amrsolver_amd import level as L
    from oracle import oracle as O
    print(f"case face fluxes {name} {layout}", flush=True)
    cls = P.closures_reference()
    n = (40, 24, 16)
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 0.6, 0.4), (1, 1, 1))
    boxes = P.decompose(n, (20, 12, 16))
    u = P.smooth_random_state(cls, n, 41)
    sch = P.scheme(**kw)
    lev = L.Level(g, boxes, cls, sch, ctx=ctx)
    S = lev.new_state(layout=layout)
    S.upload(u)
    lev.fill_patch(S)
    F = L.FaceArrays(lev)
    lev.face_fluxes(S, F)
    D = lev.new_state(ng=0)
    lev.compute_rhs(S, D)
    rhs = D.download()
    ol = O.OracleLevel(g, boxes)
    ol.set_state(u)
    ol.fill_patch()
    dxi = [1.0 / g.dx[d] for d in range(3)]
    ng = ol.ng
    for b, (lo, hi) in enumerate(boxes):
        nf = O.box_numflux(cls, sch, lo, hi, ng, ol.fabs[b], dxi)
        m = [hi[d] - lo[d] + 1 for d in range(3)]
        for d in range(3):
            e = [1 if x == d else 0 for x in range(3)]
            want = nf[d][:, ng:ng + m[2] + e[2], ng:ng + m[1] + e[1], ng:ng + m[0] + e[0]]
            assert np.array_equal(F.download(d, b), want), f"box {b} dir {d}"
        got = _assemble(F, lev, b, dxi)
        assert np.array_equal(got, rhs[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]), f"box {b}"
    lev.close()


@pytest.mark.parametrize("order,visc", [(4, True), (6, True), (2, False)], ids=["keep4v", "keep6v", "keep2"])
def test_face_fluxes_next_to_walls_assemble_to_the_rhs(ctx, order, visc):
    from amrsolver_amd import level as L
    print(f"case face fluxes walls keep{order} visc {visc}", flush=True)
    N, MG = (24, 32, 16), (12, 16, 16)
    cls, g, _, parm = P.wall_channel(N)
    sch = P.scheme(flux_euler=1, order_keep=order, visc=visc)
    lev = L.Level(g, P.decompose(N, MG), cls, sch, ctx=ctx)
    lev.set_problem(wall_T=parm["Tw"], body_force_x=parm["fx"])
    S = lev.new_state(layout="boxes")
    S.upload(P.wall_channel_state(cls, g, N, parm))
    lev.fill_patch(S)
    F = L.FaceArrays(lev)
    lev.face_fluxes(S, F)
    D = lev.new_state(ng=0)
    lev.compute_rhs(S, D)
    rhs = D.download()
    dxi = [1.0 / g.dx[d] for d in range(3)]
    for b, (lo, hi) in enumerate(lev.boxes):
        got = _assemble(F, lev, b, dxi)
        assert np.array_equal(got, rhs[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]), f"box {b}"
    lev.close()


def test_face_fluxes_reject_cell_boxes(ctx):
    from amrsolver_amd import level as L
    cls = P.closures_reference()
    n = (16, 16, 16)
    g = P.geom(n, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1))
    lev = L.Level(g, P.decompose(n, 8), cls, P.scheme(flux_euler=1, order_keep=4, visc=True), ctx=ctx)
    S = lev.new_state()
    F = L.FaceArrays(lev)
    F.fabs[1][2].hi[1] -= 1  # a cell box where the y face box is required
    with pytest.raises(capi.AmrError, match="face box"):
        lev.face_fluxes(S, F)
    lev.close()
