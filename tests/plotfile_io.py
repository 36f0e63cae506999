"""Reader of the AMReX plotfile layout (Header, Level_l/Cell_H, FAB data
files) for the tests: parses every field the writer emits and returns the
FABs as numpy arrays (component, z, y, x)."""
import os

import numpy as np


def read_plotfile(path):
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    it = iter(lines)
    hd = {"tag": next(it), "nvar": int(next(it))}
    hd["names"] = [next(it) for _ in range(hd["nvar"])]
    hd["dim"] = int(next(it))
    hd["time"] = float(next(it))
    hd["finest"] = int(next(it))
    hd["prob_lo"] = [float(x) for x in next(it).split()]
    hd["prob_hi"] = [float(x) for x in next(it).split()]
    hd["ratio"] = [int(x) for x in next(it).split()]
    hd["domains"] = next(it).split()
    hd["steps"] = [int(x) for x in next(it).split()]
    hd["dx"] = [[float(x) for x in next(it).split()] for _ in range(hd["finest"] + 1)]
    hd["coord"] = int(next(it))
    hd["bndry"] = int(next(it))
    hd["levels"] = []
    for _ in range(hd["finest"] + 1):
        lev, nb, t = next(it).split()
        step = int(next(it))
        boxes = [[tuple(float(x) for x in next(it).split()) for _ in range(3)] for _ in range(int(nb))]
        rel = next(it)
        hd["levels"].append({"level": int(lev), "time": float(t), "step": step, "rboxes": boxes, "path": rel})
    for L in hd["levels"]:
        with open(os.path.join(path, L["path"] + "_H")) as f:
            cl = f.read().split("\n")
        c = iter(cl)
        L["vers"], L["how"], L["ncomp"], L["ngrow"] = (int(next(c)) for _ in range(4))
        nb = int(next(c)[1:].split()[0])
        L["boxes"] = [next(c) for _ in range(nb)]
        assert next(c) == ")"
        assert int(next(c)) == nb
        L["fod"] = []
        for _ in range(nb):
            tag, fname, off = next(c).split()
            assert tag == "FabOnDisk:"
            L["fod"].append((fname, int(off)))
        assert next(c) == ""
        assert next(c) == f"{nb},{L['ncomp']}"
        L["min"] = [[float(x) for x in next(c).rstrip(",").split(",")] for _ in range(nb)]
        assert next(c) == ""
        assert next(c) == f"{nb},{L['ncomp']}"
        L["max"] = [[float(x) for x in next(c).rstrip(",").split(",")] for _ in range(nb)]
        L["fabs"] = []
        ldir = os.path.dirname(os.path.join(path, L["path"]))
        for fname, off in L["fod"]:
            with open(os.path.join(ldir, fname), "rb") as f:
                f.seek(off)
                head = f.readline().decode()
                assert head.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))")
                box = head[head.index(")))") + 3:head.rindex(" ")]
                ncomp = int(head.split()[-1])
                lo, hi = [tuple(int(v) for v in part.strip("() ").split(",")) for part in box.split(" ")[:2]]
                n = [hi[d] - lo[d] + 1 for d in range(3)]
                data = np.frombuffer(f.read(8 * ncomp * n[0] * n[1] * n[2]), dtype="<f8")
                L["fabs"].append((lo, hi, data.reshape(ncomp, n[2], n[1], n[0])))
    return hd
