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


class CtPhantom:
    """A phantom given as CT numbers: the densities are the CT numbers
    mapped through a calibration curve on the GPU (ch_ct_to_density), with
    optional per-structure density overrides from a uint8 label map.
    `densities` (host float32) is what Phantom holds; calculate() takes the
    device densities directly and resamples them on the GPU."""

    def __init__(self, ct, origin, spacing, curve=None, labels=None,
                 overrides=None):
        ct = np.asarray(ct)
        if ct.dtype not in (np.int16, np.float32):
            ct = ct.astype(np.float32)
        self.ct = np.ascontiguousarray(ct)
        self.size = list(self.ct.shape)
        self.origin = np.asarray(origin, dtype=np.float32)
        self.spacing = np.asarray(spacing, dtype=np.float32)
        from .density import DEFAULT_CT_CURVE
        self.curve = DEFAULT_CT_CURVE if curve is None else curve
        self.labels = None if labels is None else \
            np.ascontiguousarray(labels, dtype=np.uint8)
        if self.labels is not None and self.labels.shape != self.ct.shape:
            raise ValueError("labels must have the CT grid's shape")
        self.overrides = dict(overrides or {})
        self._host = None

    def device_densities(self, stream=None):
        """Density grid on the device (float32, the CT grid's shape)."""
        from .density import ct_to_density
        from .device import DeviceArray
        d_ct = DeviceArray.from_host(self.ct, stream=stream)
        d_lab = None if self.labels is None else \
            DeviceArray.from_host(self.labels, stream=stream)
        return ct_to_density(d_ct, self.curve, d_lab, self.overrides,
                             stream=stream)

    @property
    def densities(self):
        if self._host is None:
            self._host = self.device_densities().download()
        return self._host
