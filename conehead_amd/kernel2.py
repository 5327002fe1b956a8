"""Polyenergetic energy-deposition kernel, the `kernel` argument of convolve_c.

Follows conehead/kernel2.py:19-49: differential values per EDKnrc cone
keyed by "%.2f" % (180 - angle) (so "176.25" is the forward 3.75-degree cone,
"0.00" backscatter), divided by f_total = 0.289622754141, summed cumulatively
over the 24 sphere radii and resampled with linear interpolation onto
r_i = 0.05 + 0.01 i cm, i = 0..5995 (numpy.interp).
"""
import os

import numpy as np

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                    "varian_6.csv")
F_TOTAL = 0.289622754141
RESAMPLE = np.linspace(0.05, 60.0, 5996)


def _unique_in_order(a):
    _, idx = np.unique(a, return_index=True)
    return a[np.sort(idx)]


class PolyenergeticKernel:
    """Cumulative kernel tables keyed by cone angle ("%.2f") plus "radii"."""

    def __init__(self, path=None):
        table = np.loadtxt(path or DATA, delimiter=",", skiprows=1,
                           dtype=np.float64)
        angle, radius, value = table[:, 0], table[:, 1], table[:, 2]
        differential = {"radii": _unique_in_order(radius)}
        for a in _unique_in_order(angle):
            key = "{:.2f}".format(180 - a)
            differential[key] = value[angle == a] / F_TOTAL
        self.differential = differential
        cumulative = {"radii": differential["radii"]}
        for key, d in differential.items():
            if key == "radii":
                continue
            cumulative[key] = np.interp(RESAMPLE, cumulative["radii"],
                                        np.cumsum(d))
        self.cumulative = cumulative

    def keys(self):
        """Cone keys in ascending angle order (excluding "radii")."""
        return sorted((k for k in self.cumulative if k != "radii"),
                      key=float)


class ConeKernel:
    """Any cumulative table set with the `.cumulative` interface convolve_c
    reads (conehead/convolve_c.pyx:269-273)."""

    def __init__(self, cumulative):
        self.cumulative = cumulative

    def keys(self):
        return sorted((k for k in self.cumulative if k != "radii"),
                      key=float)


def merged_kernel(kernel, n_phi):
    """Merge the 48 cones of 3.75 degrees into n_phi equal polar groups.

    Group g covers keys [g * w, (g + 1) * w) (w = 48 / n_phi bins); its table
    is the sum of the member tables and its key is the group's central angle,
    which is itself one of the existing keys (e.g. 8 groups -> 11.25, 33.75,
    ..., 168.75).  Summing tables keeps every bin's energy; the direction of
    the group is its central angle.
    """
    keys = kernel.keys()
    n = len(keys)
    if n % n_phi:
        raise ValueError("%d cones do not split into %d groups" % (n, n_phi))
    w = n // n_phi
    step = float(keys[1]) - float(keys[0])
    cum = {"radii": kernel.cumulative.get("radii")}
    for g in range(n_phi):
        members = keys[g * w:(g + 1) * w]
        centre = float(members[0]) + step * w / 2.0
        cum["{:.2f}".format(centre)] = np.sum(
            [np.asarray(kernel.cumulative[k], np.float64) for k in members],
            axis=0)
    return ConeKernel(cum)
