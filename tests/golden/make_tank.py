"""Writes tank_10x10.json: the measured 6 MV 10x10 cm water-tank data that
the reference embeds in its plotting code (conehead/conehead.py:955-1061,
plot_10x10_profiles and plot_10x10_pdd) -- a depth-dose curve (depths in
mm, scaled by 0.1 to cm as the reference does) and cross-profiles at 5, 10,
20 and 30 cm depth (positions in mm, scaled by 0.1).

The numbers are read from the reference source as text (nothing is
imported or run) and stored as data.  Run in a tree where the reference is
mounted at /root/reference:
    python tests/golden/make_tank.py
"""
import json
import os
import re

SRC = "/root/reference/conehead/conehead.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "tank_10x10.json")


def arrays(text):
    """name -> list of floats for every `name = [0.1 *] np.array([...])`."""
    out = {}
    for m in re.finditer(r"(\w+)\s*=\s*(0\.1\s*\*\s*)?np\.array\(\[([^\]]*)\]\)",
                         text):
        vals = [float(v) for v in m.group(3).split(",") if v.strip()]
        scale = 0.1 if m.group(2) else 1.0
        out[m.group(1)] = [v * scale for v in vals]
    return out


def section(text, name):
    start = text.index("def %s(" % name)
    nxt = text.find("\n        def ", start + 1)
    return text[start:nxt if nxt > 0 else len(text)]


def main():
    text = open(SRC).read()
    prof = arrays(section(text, "plot_10x10_profiles"))
    pdd = arrays(section(text, "plot_10x10_pdd"))
    depth = [v * 0.1 for v in pdd["xpdd"]]   # xpdd *= 0.1 (:1058)
    dose = pdd["dpdd"]
    top = max(dose)
    data = {
        "source": "conehead/conehead.py:955-1061 (measured 6 MV, 10x10 cm)",
        "pdd": {"depth_cm": depth, "dose_pct": [100.0 * d / top
                                                 for d in dose]},
        "profiles": {},
    }
    for d in (5, 10, 20, 30):
        x, v = prof["x%d" % d], prof["d%d" % d]
        assert len(x) == len(v)
        c = v[len(v) // 2]                    # d / d[100] * 100 (:906)
        data["profiles"][str(d)] = {"x_cm": x,
                                    "dose_pct": [100.0 * a / c for a in v]}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=0)
    print("wrote", OUT, len(depth), "depths")


if __name__ == "__main__":
    main()
