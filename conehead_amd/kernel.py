"""Monoenergetic EDKnrc kernels and their polyenergetic sum.

Follows conehead/kernel.py:19-72.  KernelMono parses an EDKnrc listing
(.egslst): the incident energy from the "KINETIC ENERGY OF THE INCIDENT
BEAM:" line, the (angle, radius, value) rows from two lines after the
"angle/deg.  radius/cm" header up to the line before "END OF RUN", reshaped
to 48 cones x 24 spheres, normalised to unit sum, summed cumulatively over
radius and resampled onto 5996 points from 0.05 to 60 cm.  KernelPoly does
the same for a given differential table.
"""
import numpy as np

RESAMPLE = np.linspace(0.05, 60.0, 5996)


def _cumulative_table(radii, kernel_diff):
    cum = kernel_diff.cumsum(axis=1)
    out = np.zeros((cum.shape[0], RESAMPLE.size))
    for i in range(cum.shape[0]):
        out[i, :] = np.interp(RESAMPLE, radii, cum[i, :])
    return out


class KernelMono:

    def __init__(self, egslst_path=None):
        if egslst_path is None:
            raise NotImplementedError
        self._from_egslst(egslst_path)

    def _from_egslst(self, path):
        with open(path) as f:
            lines = f.readlines()
        start = end = None
        for i, line in enumerate(lines):
            if "KINETIC ENERGY OF THE INCIDENT BEAM:" in line:
                self.energy = float(line.split()[-2])
            if "angle/deg.  radius/cm" in line:
                start = i + 2
            if "END OF RUN" in line:
                end = i - 1
        if start is None or end is None:
            raise ValueError("%s is not an EDKnrc listing" % path)
        rows = np.array([l.split()[:3] for l in lines[start:end]],
                        dtype=np.float32)
        self.angles = np.unique(rows[:, 0])
        self.radii = np.unique(rows[:, 1])
        diff = rows[:, 2].reshape((48, 24))
        self.kernel_diff = diff / diff.sum()
        self.kernel_cum = _cumulative_table(self.radii, self.kernel_diff)


class KernelPoly:

    def __init__(self, angles, radii, kernel_diff):
        self.angles = angles
        self.radii = radii
        # The reference normalises kernel_diff but accumulates the argument
        # as given (conehead/kernel.py:62-63); callers pass a normalised one.
        self.kernel_diff = kernel_diff / kernel_diff.sum()
        self.kernel_cum = _cumulative_table(radii, kernel_diff)
