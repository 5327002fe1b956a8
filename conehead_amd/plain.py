"""The plain single-beam GPU path: the unbatched baseline.

One beam per call, every stage synchronised and its host transfers redone,
as the reference's calculate flow does (conehead/conehead.py:520-773):
densities and the 4000 x 4000 aperture raster go to the device per call
(:555, :568-573), TERMA runs over the full raster (no analytic aperture)
and comes back to the host, then TERMA and the cone tables go up again and
the scatter convolution (ch_convolve_scatter: one thread per voxel and
cone, FP64 DDA, FP32 atomics -- the structure of cuda_dose, :385-488) runs
and the dose comes back.  Nothing is cached between calls.

It computes the same dose as the gather engine (same taps, FP32 sums in a
different order), which tests/test_gpu_plain.py checks.
"""
import ctypes

import numpy as np

from . import _lib
from . import engine as E
from .device import DeviceArray, device_sync

SEMANTICS = {"convolve_c": 0, "cuda_dose": 1}


def scatter(d_terma, d_dose, shape, spacing, d_dirs, d_rows, d_cum, row_len,
            n_cones, semantics="convolve_c", stream=None):
    """dose += scatter convolution of terma (device arrays)."""
    shp = (ctypes.c_int32 * 3)(*[int(v) for v in shape])
    vs = (ctypes.c_double * 3)(*[float(v) for v in spacing])
    from .device import _h
    _lib.check(_lib.lib().ch_convolve_scatter(
        d_terma.ptr, d_dose.ptr, shp, vs, int(n_cones), d_dirs.ptr,
        d_rows.ptr, d_cum.ptr, int(row_len), SEMANTICS[semantics],
        _h(stream)))


def plain_beam(densities, origin, spacing, source, block, settings, cones):
    """Dose (float32 [X][Y][Z]) of one beam through the plain path; also
    returns the TERMA grid it computed."""
    densities = np.ascontiguousarray(densities, np.float32)
    shape = densities.shape
    # stage 1: TERMA
    d_rho = DeviceArray.from_host(densities)
    d_raster = DeviceArray.from_host(
        np.ascontiguousarray(block.block_values, np.float32))
    d_t = DeviceArray(shape)
    beam = E.beam_struct(source, settings, d_raster, d_rho, d_t, None)
    E.terma([beam], shape, origin, spacing)
    device_sync()
    terma = d_t.download()
    for a in (d_rho, d_raster, d_t):
        a.free()
    # stage 2: convolution
    d_t = DeviceArray.from_host(terma)
    d_dirs = DeviceArray.from_host(np.ascontiguousarray(cones.dirs,
                                                        np.float64))
    d_rows = DeviceArray.from_host(np.ascontiguousarray(cones.rows,
                                                        np.int32))
    cum = np.ascontiguousarray(cones.cum, np.float64)
    d_cum = DeviceArray.from_host(cum)
    d_dose = DeviceArray(shape, zero=True)
    scatter(d_t, d_dose, shape, spacing, d_dirs, d_rows, d_cum,
            cum.shape[1], cones.n_cones, cones.semantics)
    device_sync()
    dose = d_dose.download()
    for a in (d_t, d_dirs, d_rows, d_cum, d_dose):
        a.free()
    return dose, terma


def plain_case(case, beam=0):
    """plain_beam for beam `beam` of a workloads.Case."""
    src, blk, st = case.beams[beam]
    return plain_beam(case.densities, case.origin, case.spacing, src, blk,
                      st, case.cones)
