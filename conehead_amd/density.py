"""Phantom densities on the dose grid, resampled on the GPU.

resample() runs ch_resample_linear: the reference's linear interpolation
of the phantom (scipy RegularGridInterpolator on the axes
linspace(origin, origin + n * spacing, n), fill 0 outside;
conehead/conehead.py:525-533) evaluated at the dose grid's axes with the
meshgrid 'xy' layout of :561-568 (swap_xy=True) or in [x][y][z] order.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray, _h


def _v3(v, t):
    dt = np.float64 if t is ctypes.c_double else np.int32
    return (t * 3)(*np.asarray(v, dtype=dt).reshape(3).tolist())


def resample(d_src, src_shape, src_origin, src_spacing, dst_shape,
             dst_origin, dst_spacing, swap_xy=True, d_dst=None, stream=None):
    """Device array of the densities on the dose grid (float32)."""
    dst_shape = tuple(int(v) for v in dst_shape)
    if d_dst is None:
        d_dst = DeviceArray(dst_shape, np.float32)
    _lib.check(_lib.lib().ch_resample_linear(
        d_src.ptr, _v3(src_shape, ctypes.c_int32),
        _v3(src_origin, ctypes.c_double), _v3(src_spacing, ctypes.c_double),
        d_dst.ptr, _v3(dst_shape, ctypes.c_int32),
        _v3(dst_origin, ctypes.c_double), _v3(dst_spacing, ctypes.c_double),
        1 if swap_xy else 0, _h(stream)))
    return d_dst


# A generic two-segment calibration (air, water, cortical bone) for CT
# numbers in Hounsfield units; a scanner's own curve replaces it.
DEFAULT_CT_CURVE = ((-1000.0, 0.0012), (0.0, 1.0), (3000.0, 2.9))

CT_TYPES = {np.dtype(np.int16): 0, np.dtype(np.float32): 1}


def _curve(curve):
    c = np.asarray(curve, dtype=np.float64).reshape(-1, 2)
    return (np.ascontiguousarray(c[:, 0]), np.ascontiguousarray(c[:, 1]))


def override_table(overrides):
    """256 float32 densities per label, NaN where the label keeps the
    curve's density; `overrides` maps label (1..255) -> density."""
    t = np.full(256, np.nan, dtype=np.float32)
    for label, rho in (overrides or {}).items():
        label = int(label)
        if not 0 <= label < 256:
            raise ValueError("label %d outside 0..255" % label)
        t[label] = np.float32(rho)
    return t


def ct_to_density(d_ct, curve=DEFAULT_CT_CURVE, d_labels=None,
                  overrides=None, d_rho=None, stream=None):
    """Device densities (float32, the CT grid's shape) from CT numbers
    (int16 or float32 DeviceArray) through the calibration curve
    [(ct, density), ...] as numpy.interp evaluates it; voxels of d_labels
    (uint8 DeviceArray) whose label is in `overrides` take that density."""
    dt = np.dtype(d_ct.dtype)
    if dt not in CT_TYPES:
        raise TypeError("CT numbers must be int16 or float32, not %s" % dt)
    hu, rho = _curve(curve)
    table = None
    lab = None
    if d_labels is not None:
        if np.dtype(d_labels.dtype) != np.uint8 or \
                tuple(d_labels.shape) != tuple(d_ct.shape):
            raise ValueError("labels must be uint8 of the CT grid's shape")
        table = override_table(overrides)
        lab = d_labels.ptr
    elif overrides:
        raise ValueError("density overrides need a label map")
    if d_rho is None:
        d_rho = DeviceArray(tuple(d_ct.shape), np.float32)
    n = int(np.prod(d_ct.shape))
    _lib.check(_lib.lib().ch_ct_to_density(
        d_ct.ptr, CT_TYPES[dt], n,
        hu.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        rho.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(hu), lab,
        None if table is None else
        table.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
        d_rho.ptr, _h(stream)))
    return d_rho
