"""Extract the DDA golden vectors from the reference notebook outputs.

Reads dda_benchmarking.ipynb as JSON text (no reference code is imported or
run) and writes tests/golden/dda_goldens.json:

  ray A -- inputs dda_benchmarking.ipynb:11-15 (cell 0): direction
           (1,2,3)/sqrt(14), start (50,50,50), 101^3 grid, unit voxels;
           outputs: the 102 t-values printed by cell 3 (:105-206), identical
           to the first 102 values printed by the Cython variant (cell 10).
  ray C -- inputs :651-654 (cell 11): direction (0.997859, 0, -0.065403),
           start (14,14,4), 41^3 grid, unit voxels; outputs: the 29 voxels
           of cell 16 (:855-883) and the per-step t-states of cell 14
           (:786-814) at full precision (the appended t is the minimum of
           each state).

Run in the build container only:  python tests/golden/make_dda_goldens.py
"""
import ast
import json
import os
import re
import sys

NB = "/root/reference/dda_benchmarking.ipynb"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "dda_goldens.json")


def _text(cell):
    out = []
    for o in cell.get("outputs", []):
        t = "".join(o.get("text", [])) or \
            "".join(o.get("data", {}).get("text/plain", []))
        out.append(t)
    return "\n".join(out)


def main():
    nb = json.load(open(NB))
    cells = nb["cells"]
    # ray A: cell 3 prints the Python list of t-values
    ray_a_t = ast.literal_eval(_text(cells[3]).strip())
    fast = _text(cells[10]).strip()
    fast_vals = [float(x) for x in re.findall(r"[-+0-9.e]+|nan", fast)
                 if x not in ("-", "+")]
    assert fast_vals[:len(ray_a_t)] == ray_a_t, "cells 3 and 10 disagree"
    # ray C: cell 14 prints the t-state before each step
    states = [ast.literal_eval(line) for line in
              _text(cells[14]).strip().splitlines() if line.startswith("[")]
    vox_txt = _text(cells[16])
    vox = [list(map(int, m)) for m in
           re.findall(r"\[\s*(\d+),\s*(\d+),\s*(\d+)\]", vox_txt)]
    assert len(states) == len(vox) == 29
    data = {
        "source": "dda_benchmarking.ipynb (reference notebook outputs)",
        "ray_A": {
            "direction_expr": ["1.0/sqrt(14.0)", "2.0/sqrt(14.0)",
                               "3.0/sqrt(14.0)"],
            "grid_shape": [101, 101, 101],
            "first_voxel": [50, 50, 50],
            "voxel_size": [1.0, 1.0, 1.0],
            "t": [repr(float(x)) for x in ray_a_t],
        },
        "ray_C": {
            "direction": [0.997859, 0.0, -0.065403],
            "grid_shape": [41, 41, 41],
            "first_voxel": [14, 14, 4],
            "voxel_size": [1.0, 1.0, 1.0],
            "voxels": vox,
            "t_states": [[repr(float(v)) for v in s] for s in states],
        },
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT, len(ray_a_t), len(vox))


if __name__ == "__main__":
    sys.exit(main())
