"""Python entry to the GPU 3D-DDA (the counterpart of conehead's dda_3d_c).

The reference's dda_3d_c is a C-level Cython function (conehead/dda_3d_c.pxd:8)
that appends to caller-owned `vector`s; here the traversal runs on the GPU in
FP64 (ch_dda_3d / ch_dda_trace_batch) and returns numpy arrays.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray, _h


def dda_3d_c(direction, grid_shape, first_voxel, voxel_size,
             mutate_direction=False):
    """Voxels and boundary-crossing t-values of one ray.

    Same traversal as conehead/dda_3d_c.pyx:33-99.  Returns
    (voxels int32[N, 3], t float64[N]).  With mutate_direction=True a numpy
    `direction` is replaced by its absolute value, as the reference does
    (dda_3d_c.pyx:38-43)."""
    L = _lib.lib()
    d = np.ascontiguousarray(direction, dtype=np.float64).reshape(3)
    shp = np.ascontiguousarray(grid_shape, dtype=np.int32).reshape(3)
    fv = np.ascontiguousarray(first_voxel, dtype=np.int32).reshape(3)
    vs = np.ascontiguousarray(voxel_size, dtype=np.float64).reshape(3)
    cap = int(np.clip(shp, 0, None).sum()) + 2
    vox = np.zeros((cap, 3), np.int32)
    tv = np.zeros(cap, np.float64)
    dabs = np.zeros(3, np.float64)
    n = ctypes.c_int32()
    _lib.check(L.ch_dda_3d(d.ctypes.data_as(_lib.c_f64p),
                           shp.ctypes.data_as(_lib.c_i32p),
                           fv.ctypes.data_as(_lib.c_i32p),
                           vs.ctypes.data_as(_lib.c_f64p),
                           vox.ctypes.data_as(_lib.c_i32p),
                           tv.ctypes.data_as(_lib.c_f64p), cap,
                           ctypes.byref(n), dabs.ctypes.data_as(_lib.c_f64p)))
    if mutate_direction and isinstance(direction, np.ndarray):
        direction[...] = dabs.reshape(direction.shape)
    return vox[:n.value].copy(), tv[:n.value].copy()


def dda_trace(directions, first_voxels, grid_shape, voxel_size, cap=None,
              stream=None):
    """Many traversals in one GPU launch.

    directions: float64[n, 3]; first_voxels: int32[n, 3].  Returns
    (counts int32[n], voxels int32[n, cap, 3], t float64[n, cap])."""
    L = _lib.lib()
    dirs = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 3)
    firsts = np.ascontiguousarray(first_voxels, dtype=np.int32).reshape(-1, 3)
    n = len(dirs)
    if len(firsts) != n:
        raise ValueError("one first voxel per direction")
    shp = (ctypes.c_int32 * 3)(*[int(s) for s in grid_shape])
    vs = (ctypes.c_double * 3)(*[float(v) for v in voxel_size])
    if cap is None:
        cap = int(sum(max(int(s), 0) for s in grid_shape)) + 2
    d_dirs = DeviceArray.from_host(dirs)
    d_first = DeviceArray.from_host(firsts)
    d_vox = DeviceArray((max(n, 1), cap, 3), np.int32)
    d_t = DeviceArray((max(n, 1), cap), np.float64)
    d_cnt = DeviceArray((max(n, 1),), np.int32)
    _lib.check(L.ch_dda_trace_batch(n, d_dirs.ptr, d_first.ptr, shp, vs, cap,
                                    d_vox.ptr, d_t.ptr, d_cnt.ptr,
                                    _h(stream)))
    counts = d_cnt.download(stream=stream)[:n]
    vox = d_vox.download(stream=stream)[:n]
    t = d_t.download(stream=stream)[:n]
    return counts, vox, t
