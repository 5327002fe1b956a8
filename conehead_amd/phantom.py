"""Phantom container (conehead/phantom.py:5-16)."""
import numpy as np


class SimplePhantom:
    """40 cm water cube: 201^3 voxels of 2 mm, origin (-20.1, 0, -20.1)."""

    def __init__(self):
        self.size = [201, 201, 201]
        self.origin = np.array([-20.1, 0, -20.1], dtype=np.float32)
        self.spacing = np.array([.2, .2, .2], dtype=np.float32)
        self.densities = np.ones(self.size, dtype=np.float32)


class Phantom:
    """Arbitrary density grid with the same attributes as SimplePhantom."""

    def __init__(self, densities, origin, spacing):
        self.densities = np.ascontiguousarray(densities, dtype=np.float32)
        self.size = list(self.densities.shape)
        self.origin = np.asarray(origin, dtype=np.float32)
        self.spacing = np.asarray(spacing, dtype=np.float32)
