"""Dose grid container (conehead/dosegrid.py:5-11)."""
import numpy as np


class DoseGrid:

    def __init__(self, size, origin, spacing):
        self.size = size
        self.origin = origin
        self.spacing = spacing
        self.dose = np.zeros(self.size, dtype=np.float32)
