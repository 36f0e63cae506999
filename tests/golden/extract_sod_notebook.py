"""Writes sod_exact_notebook.json from the output the reference's own
notebook holds (exm/Sods/KEEP_AD_validation/KEEP_AD_validation.ipynb, cell 5:
exact Sod solution, test 1, t = 0.2, gamma = 1.4, diaphragm at 0.5 on [0, 1],
left (p, rho, u) = (1, 1, 0), right = (0.1, 0.125, 0)).

Reads the notebook as JSON text (no code from it runs); run in a checkout
that has the reference next to it:

    python tests/golden/extract_sod_notebook.py /path/to/reference
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
nb = json.load(open(os.path.join(ref, "exm/Sods/KEEP_AD_validation/KEEP_AD_validation.ipynb")))
text = "".join("".join(o.get("text", [])) for o in nb["cells"][5].get("outputs", []))
num = r"([-+0-9.eE]+)"
out = {"source": "exm/Sods/KEEP_AD_validation/KEEP_AD_validation.ipynb cell 5 output",
       "case": {"t": 0.2, "gamma": 1.4, "x0": 0.5, "left_p_rho_u": [1.0, 1.0, 0.0],
                "right_p_rho_u": [0.1, 0.125, 0.0]},
       "positions": {}, "regions_p_rho_u": {}}
for name in ("Head of Rarefaction", "Foot of Rarefaction", "Contact Discontinuity", "Shock"):
    out["positions"][name] = float(re.search(re.escape(name) + r"\s*:\s*" + num, text).group(1))
for r in (3, 4):
    m = re.search(rf"Region {r}\s*:\s*\({num}, {num}, {num}\)", text)
    out["regions_p_rho_u"][str(r)] = [float(m.group(i)) for i in (1, 2, 3)]
for key in ("c5", "us", "Ms"):
    out[key] = float(re.search(rf"^{key} =\s*{num}", text, re.M).group(1))
here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "sod_exact_notebook.json"), "w") as f:
    json.dump(out, f, indent=1)
print(json.dumps(out, indent=1))
