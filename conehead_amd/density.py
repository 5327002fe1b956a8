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
