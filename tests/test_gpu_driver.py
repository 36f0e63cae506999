"""The C++ host driver (amrsolver_amd/lib/cns_driver, the C ABI alone, no
Python or torch in the process) against the Python host mirror: the same
three-level subcycled TGV hierarchy from the same initial data, two coarse
steps of SSPRK(3,3) and (3,4), read back from the driver's plotfile; every
level equal to level.Hierarchy's result to 1e-12 per cell (the same kernels
in the same order: bit for bit in practice).  Also a default single-level
run that reports conserved totals."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from amrsolver_amd import problems as P
from conftest import ROOT, rel_err
from plotfile_io import read_plotfile

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "amrsolver_amd", "lib", "cns_driver")


@pytest.fixture(scope="module")
def ctx():
    from amrsolver_amd import level
    return level.Context(0)


def _levels():
    L = 2 * math.pi
    g0 = P.geom((16, 16, 16), (0, 0, 0), (L, L, L), (1, 1, 1))
    g1 = P.geom((32, 32, 32), (0, 0, 0), (L, L, L), (1, 1, 1))
    g2 = P.geom((64, 64, 64), (0, 0, 0), (L, L, L), (1, 1, 1))
    return [(g0, P.decompose((16, 16, 16), 8)),
            (g1, [((8, 8, 4), (23, 15, 27)), ((8, 16, 4), (23, 23, 27))]),
            (g2, [((24, 24, 16), (39, 39, 47))])]


def _run(args, timeout=120):
    print("driver:", " ".join(args), flush=True)
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("stages", [3, 4])
def test_driver_hierarchy_matches_python(ctx, tmp_path, stages):
    from amrsolver_amd import level as LV
    assert os.path.exists(DRIVER), "build the driver first (python -m amrsolver_amd.build)"
    cls = P.tgv_closures()
    sch = P.scheme(flux_euler=1, order_keep=4, visc=True, order_rk=3, stages_rk=stages)
    levels = _levels()
    us = [P.tgv_state(cls, g, tuple(g.dom_hi[d] + 1 for d in range(3))) for g, _ in levels]
    init = tmp_path / "init"
    init.mkdir()
    with open(tmp_path / "grids.txt", "w") as f:
        for l, (g, boxes) in enumerate(levels):
            np.ascontiguousarray(us[l]).tofile(str(init / f"level_{l}.bin"))
            for lo, hi in boxes:
                f.write(f"{l} {lo[0]} {lo[1]} {lo[2]} {hi[0]} {hi[1]} {hi[2]}\n")
    dt = 2e-3
    out = _run(["--n", "16", "--steps", "2", "--dt", repr(dt), "--stages", str(stages), "--grids",
                str(tmp_path / "grids.txt"), "--init", str(init), "--plot", str(tmp_path / "plt")])
    assert out["levels"] == 3 and out["steps"] == 2
    H = LV.Hierarchy([LV.Level(g, b, cls, sch, ctx=ctx) for g, b in levels], layout="boxes")
    H.upload(us)
    t = 0.0
    for _ in range(2):
        H.advance(t, dt)
        t += dt
    hd = read_plotfile(str(tmp_path / "plt"))
    assert hd["steps"] == [2, 4, 8] and abs(hd["time"] - t) <= 1e-15
    for l, (g, boxes) in enumerate(levels):
        want = H.download(l)
        for b, (lo, hi) in enumerate(boxes):
            flo, fhi, data = hd["levels"][l]["fabs"][b]
            assert (flo, fhi) == (tuple(lo), tuple(hi))
            ref = want[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
            err = rel_err(data[:5], ref)
            assert err <= 1e-12, f"level {l} box {b}: {err:.3e}"


def test_driver_single_level_default(tmp_path):
    assert os.path.exists(DRIVER)
    out = _run(["--n", "32", "--steps", "3"])
    assert out["levels"] == 1 and out["steps"] == 3
    tot = out["totals_level0"]
    assert all(math.isfinite(x) for x in tot) and tot[0] > 0
