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


# Incident energies (MeV) of the EDKnrc listings, as settings keys.
ENERGIES = ("0.5", "1.0", "1.5", "2.0", "2.5", "3.0", "3.5", "4.0", "4.5",
            "5.0", "5.5", "6.0")


def listing_files(root, reference_1mev=True):
    """Listing path per energy key under `root` (the kernels/ directory).
    With reference_1mev the "1.0" entry reads the 1.5 MeV listing, as
    conehead/conehead.py:688 does; False reads the 1.0 MeV one."""
    import os
    out = {}
    for e in ENERGIES:
        d = "1.5" if (e == "1.0" and reference_1mev) else e
        out[e] = os.path.join(root, "%sMeV" % d, "%sMeV.egslst" % d)
    return out


def polyenergetic(kernels, energy_weights):
    """Weighted sum of monoenergetic kernels, renormalised, as a KernelPoly
    (conehead/conehead.py:700-704).  kernels: energy key -> object with
    angles, radii, kernel_diff; energy_weights: settings['energy_weights']."""
    first = kernels[next(iter(kernels))]
    diff = np.zeros_like(first.kernel_diff)
    for e, w in energy_weights.items():
        diff += kernels[e].kernel_diff * w
    diff = diff / diff.sum()
    return KernelPoly(first.angles, first.radii, diff)


class EdkSet:
    """The twelve monoenergetic EDKnrc tables in one compact file
    (conehead_amd/data/edk_6mv.npz): per energy the normalised 48 x 24
    differential table of the listing, with the angle and radius grids.
    .mono(key) gives a KernelMono-like object; .poly(weights) the
    polyenergetic kernel of conehead/conehead.py:686-704."""

    def __init__(self, energies, angles, radii, diffs):
        self.energies = tuple(energies)
        self.angles = np.asarray(angles, np.float32)
        self.radii = np.asarray(radii, np.float32)
        self.diffs = np.asarray(diffs, np.float32)

    @classmethod
    def from_listings(cls, root):
        ks = [KernelMono(p) for p in listing_files(root, False).values()]
        return cls(ENERGIES, ks[0].angles, ks[0].radii,
                   np.stack([k.kernel_diff for k in ks]))

    @classmethod
    def load(cls, path=None):
        import os
        if path is None:
            path = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                "data", "edk_6mv.npz")
        z = np.load(path, allow_pickle=False)
        return cls([str(e) for e in z["energies"]], z["angles"],
                   z["radii"], z["diffs"])

    def save(self, path):
        np.savez_compressed(path, energies=np.array(self.energies),
                            angles=self.angles, radii=self.radii,
                            diffs=self.diffs)

    def mono(self, key):
        k = KernelMono.__new__(KernelMono)
        k.energy = float(key)
        k.angles = self.angles
        k.radii = self.radii
        k.kernel_diff = self.diffs[self.energies.index(key)]
        k.kernel_cum = _cumulative_table(self.radii, k.kernel_diff)
        return k

    def poly(self, energy_weights, reference_1mev=True):
        kernels = {}
        for e in energy_weights:
            src = "1.5" if (e == "1.0" and reference_1mev) else e
            kernels[e] = self.mono(src)
        return polyenergetic(kernels, energy_weights)


def merge_polar(angles, kernel_cum, n_phi):
    """Merges the EDKnrc polar bins (angles = upper bin edges, ascending,
    equal width) into n_phi equal groups for a coarser cone set: a group's
    table is the sum of its members' cumulative tables (so it keeps their
    energy) and its angle is the centre of its angular range.  Returns
    (angles[n_phi], cum[n_phi, row_len])."""
    angles = np.asarray(angles, dtype=np.float64)
    cum = np.asarray(kernel_cum, dtype=np.float64)
    n = len(angles)
    if n_phi < 1 or n % n_phi:
        raise ValueError("%d bins do not split into %d groups" % (n, n_phi))
    w = n // n_phi
    width = angles[1] - angles[0] if n > 1 else angles[0]
    out_a = np.empty(n_phi)
    out_c = np.empty((n_phi, cum.shape[1]))
    for g in range(n_phi):
        lo = angles[g * w] - width
        hi = angles[(g + 1) * w - 1]
        out_a[g] = 0.5 * (lo + hi)
        out_c[g] = cum[g * w:(g + 1) * w].sum(axis=0)
    return out_a, np.ascontiguousarray(out_c)

