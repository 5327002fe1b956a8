"""Drop-in for conehead's convolve_c (conehead/convolve_c.pyx:17-197).

Same signature and semantics: `dose_grid_dose` (float64, any strides) is
accumulated in place with the collapsed-cone convolution of
`dose_grid_terma`, for the cones thetas x phis (degrees) and the cumulative
kernel tables kernel.cumulative["%.2f" % phi].  The work runs on the GPU as
the deterministic stencil gather (FP32 with compensated sums) and the result
is added to the FP64 dose on the device.
"""
import numpy as np

from .cones import ConeSet
from .device import DeviceArray
from . import engine as _engine

_ENGINE = None


def default_engine():
    global _ENGINE
    if _ENGINE is None:
        _ENGINE = _engine.Engine()
    return _ENGINE


def _check_grid(a, name, writable=False):
    if not isinstance(a, np.ndarray):
        a = np.asarray(a)
    if a.ndim != 3:
        raise ValueError("Buffer has wrong number of dimensions (expected 3,"
                         " got %d)" % a.ndim)
    if a.dtype != np.float64:
        raise ValueError("Buffer dtype mismatch, expected 'float64_t' but "
                         "got '%s' (%s)" % (a.dtype, name))
    if writable and not a.flags.writeable:
        raise ValueError("buffer source array is read-only (%s)" % name)
    return a


def convolve_c(dose_grid_terma, dose_grid_dose, dose_grid_dim, thetas, phis,
               kernel):
    """Accumulate the cone convolution of TERMA into `dose_grid_dose`."""
    terma = _check_grid(dose_grid_terma, "dose_grid_terma")
    dose = _check_grid(dose_grid_dose, "dose_grid_dose", writable=True)
    if dose.shape != terma.shape:
        # the reference indexes dose with the TERMA grid's voxel indices
        raise ValueError("dose grid %s does not match TERMA grid %s" %
                         (dose.shape, terma.shape))
    dims = np.asarray(dose_grid_dim, dtype=np.float64).reshape(3)
    cones = ConeSet(thetas, phis, kernel, semantics="convolve_c")
    if cones.n_cones == 0:
        return None
    eng = default_engine()
    plan = eng.plan(terma.shape, dims, cones)
    d_t = DeviceArray.from_host(np.ascontiguousarray(terma, np.float32))
    d_box = DeviceArray((6,), np.int32)
    _engine.support_box(d_t, terma.shape, d_box)
    d_dose = DeviceArray.from_host(np.ascontiguousarray(dose))
    _engine.convolve_f64(plan, d_t, d_dose, d_box)
    out = d_dose.download()
    if dose.flags.c_contiguous:
        dose[...] = out
    else:
        np.copyto(dose, out)
    return None
