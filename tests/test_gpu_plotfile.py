"""Plotfiles (CNS::writePlotFile, src/CNS.cpp:644-874) written from GPU
states of a two-level hierarchy, read back by an independent parser of the
AMReX plotfile layout:
- Header: format tag, names in CNS's order (state, then derived variables in
  registration order whatever order they are asked in), time, finest level,
  extent, ratio, domains, steps, cell sizes, boxes in physical coordinates;
- Cell_H: the boxes, FAB offsets that point at each FAB header, per-FAB
  minima and maxima equal to the data;
- the FABs: state components bit for bit; pressure, temperature and the
  velocities equal the primitive fields the state was built from (closed
  form, 1e-13 relative)."""
import math

import numpy as np
import pytest

from amrsolver_amd import problems as P
from plotfile_io import read_plotfile

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _levels():
    L = 2 * math.pi
    g0 = P.geom((16, 16, 16), (0, 0, 0), (L, L, L), (1, 1, 1))
    g1 = P.geom((32, 32, 32), (0, 0, 0), (L, L, L), (1, 1, 1))
    return [(g0, P.decompose((16, 16, 16), 8)), (g1, [((8, 8, 4), (23, 15, 27)), ((8, 16, 4), (23, 23, 27))])]


def _prims(cls, g, seed):
    n = [g.dom_hi[d] + 1 for d in range(3)]
    rng = np.random.default_rng(seed)
    shape = (n[2], n[1], n[0])
    rho = 1.0 + 0.2 * rng.random(shape)
    vel = [rng.standard_normal(shape) for _ in range(3)]
    p = 1.0e5 * (1.0 + 0.1 * rng.random(shape))
    u = np.empty((5,) + shape)
    u[0] = rho
    for d in range(3):
        u[1 + d] = rho * vel[d]
    u[4] = p / (cls.gamma - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return u, rho, vel, p


@pytest.mark.parametrize("layout", ["level", "boxes"])
def test_plotfile_two_levels(ctx, tmp_path, layout):
    from amrsolver_amd import level as LV
    print(f"case plotfile {layout}", flush=True)
    cls = P.tgv_closures()
    sch = P.scheme()
    levels = _levels()
    Ls = [LV.Level(g, b, cls, sch, ctx=ctx) for g, b in levels]
    states, prims = [], []
    for l, (g, _) in enumerate(levels):
        u, rho, vel, p = _prims(cls, g, 3 + l)
        S = Ls[l].new_state(layout=layout)
        S.upload(u)
        states.append(S)
        prims.append((u, rho, vel, p))
    path = tmp_path / "plt00010"
    # asked out of order: written in registration order
    names = LV.write_plotfile(path, Ls, states, 0.125, [10, 20], 2,
                              derive=("z_velocity", "temperature", "x_velocity", "pressure", "y_velocity"))
    hd = read_plotfile(str(path))
    want = ["density", "xmom", "ymom", "zmom", "energy", "pressure", "temperature", "x_velocity", "y_velocity",
            "z_velocity"]
    assert names == want and hd["names"] == want
    assert hd["tag"] == "HyperCLaw-V1.1" and hd["dim"] == 3 and hd["time"] == 0.125 and hd["finest"] == 1
    assert hd["ratio"] == [2] and hd["steps"] == [10, 20] and hd["coord"] == 0 and hd["bndry"] == 0
    assert np.allclose(hd["prob_lo"], [0, 0, 0]) and np.allclose(hd["prob_hi"], [2 * math.pi] * 3, rtol=1e-14)
    assert hd["domains"] == ["((0,0,0)", "(15,15,15)", "(0,0,0))", "((0,0,0)", "(31,31,31)", "(0,0,0))"]
    for l, (g, boxes) in enumerate(levels):
        L = hd["levels"][l]
        assert L["level"] == l and L["step"] == [10, 20][l] and L["path"] == f"Level_{l}/Cell"
        assert np.allclose(hd["dx"][l], [g.dx[d] for d in range(3)], rtol=1e-14)
        assert L["ncomp"] == 10 and L["ngrow"] == 0 and len(L["fabs"]) == len(boxes)
        u, rho, vel, p = prims[l]
        for b, (lo, hi) in enumerate(boxes):
            rb = L["rboxes"][b]
            for d in range(3):
                assert abs(rb[d][0] - lo[d] * g.dx[d]) < 1e-12 and abs(rb[d][1] - (hi[d] + 1) * g.dx[d]) < 1e-12
            flo, fhi, data = L["fabs"][b]
            assert (flo, fhi) == (tuple(lo), tuple(hi))
            assert L["boxes"][b] == f"(({lo[0]},{lo[1]},{lo[2]}) ({hi[0]},{hi[1]},{hi[2]}) (0,0,0))"
            sl = (slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
            assert np.array_equal(data[:5], u[(slice(None),) + sl])
            T = p[sl] / (rho[sl] * cls.Rspec)
            for c, ref in zip(range(5, 10), [p[sl], T] + [v[sl] for v in vel]):
                assert np.allclose(data[c], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
            assert L["min"][b] == [float(data[c].min()) for c in range(10)]
            assert L["max"][b] == [float(data[c].max()) for c in range(10)]
