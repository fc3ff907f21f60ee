"""Extracts the reference's stored numeric outputs into fixtures.

The only numeric output the reference holds is the list of deformation
radii (km) printed by set_background_profiles.ipynb (cell 9, notebook lines
180-185) for its 32-layer demonstration profile.  This script reads the
notebook as JSON text (nothing in it is executed) and writes
notebook_radii.json.  Run it in a checkout where the reference is present:
    python tests/golden/make_golden.py /path/to/reference
"""
import json
import os
import re
import sys


def main(ref):
    nb = json.load(open(os.path.join(ref, "set_background_profiles.ipynb")))
    for cell in nb["cells"]:
        if cell["cell_type"] != "code":
            continue
        src = "".join(cell["source"])
        if "np.sort(1/np.sqrt(-vals))" not in src:
            continue
        for out in cell.get("outputs", []):
            text = "".join(out.get("text", []))
            if text.strip().startswith("["):
                body = text[text.index("[") + 1:text.index("]")]
                vals = [float(v) if v != "nan" else None for v in re.split(r"\s+", body.strip())]
                path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "notebook_radii.json")
                json.dump({"source": "set_background_profiles.ipynb cell 9 output (lines 180-185)",
                           "units": "km", "H_tot": 4800.0, "nz": 32, "alpha": 1.0,
                           "radii_sorted": vals}, open(path, "w"), indent=1)
                print("wrote", path, len(vals), "values")
                return
    raise SystemExit("radii output not found")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
