"""Drop-in for conehead's convolve_c (conehead/convolve_c.pyx:17-197).

Same signature and semantics: `dose_grid_dose` (float64, any strides) is
accumulated in place with the collapsed-cone convolution of
`dose_grid_terma`, for the cones thetas x phis (degrees) and the cumulative
kernel tables kernel.cumulative["%.2f" % phi].  The work runs on the GPU:
the FP64 TERMA is converted to FP32 on the device, the convolution runs in
the FFT form (grids whose axes fit its frames) or as the deterministic
stencil gather, and the FP32 result is added into the FP64 dose on the
device.  Cone sets, plans and device buffers are kept between calls.
"""
import threading

import numpy as np

from . import _lib
from . import engine as _engine
from .cones import ConeSet, kernel_rows
from .device import DeviceArray

_ENGINE = None
_LOCK = threading.Lock()
_CALL_LOCK = threading.Lock()   # one call at a time owns the buffers
_CONES = {}     # (thetas, phis, id(kernel)) -> (kernel rows, ConeSet)
_BUFS = {}      # grid shape -> device buffers of one call


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


def _cones(thetas, phis, kernel):
    """The ConeSet of these angles and kernel, reused while the kernel's
    rows are unchanged (they are compared by value on every call)."""
    th = np.asarray(thetas, dtype=np.float64).ravel()
    ph = np.asarray(phis, dtype=np.float64).ravel()
    rows = kernel_rows(kernel, ph) if len(ph) else None
    key = (th.tobytes(), ph.tobytes(), id(kernel))
    with _LOCK:
        hit = _CONES.get(key)
    if hit is not None and (rows is None or np.array_equal(hit[0], rows)):
        return hit[1]
    cones = ConeSet(th, ph, kernel, semantics="convolve_c")
    with _LOCK:
        if len(_CONES) > 64:
            _CONES.clear()
        _CONES[key] = (None if rows is None else rows.copy(), cones)
    return cones


def _buffers(shape):
    with _LOCK:
        b = _BUFS.get(shape)
        if b is None:
            if len(_BUFS) > 4:
                _BUFS.clear()
            b = _BUFS[shape] = {
                "t64": DeviceArray(shape, np.float64),
                "t": DeviceArray(shape, np.float32),
                "part": DeviceArray(shape, np.float32),
                "dose": DeviceArray(shape, np.float64),
                "box": DeviceArray((6,), np.int32)}
        return b


def _convert(dst, src, accumulate=False):
    types = {np.dtype(np.float32): _lib.CH_F32,
             np.dtype(np.float64): _lib.CH_F64}
    _lib.check(_lib.lib().ch_convert(dst.ptr, types[dst.dtype], src.ptr,
                                     types[src.dtype], dst.size,
                                     1 if accumulate else 0, None))


def convolve_c(dose_grid_terma, dose_grid_dose, dose_grid_dim, thetas, phis,
               kernel, method="auto"):
    """Accumulate the cone convolution of TERMA into `dose_grid_dose`.

    method: "auto" (the FFT form where the grid fits its frames, else the
    gather), "fft" or "gather"."""
    terma = _check_grid(dose_grid_terma, "dose_grid_terma")
    dose = _check_grid(dose_grid_dose, "dose_grid_dose", writable=True)
    if dose.shape != terma.shape:
        # the reference indexes dose with the TERMA grid's voxel indices
        raise ValueError("dose grid %s does not match TERMA grid %s" %
                         (dose.shape, terma.shape))
    if method not in ("auto", "fft", "gather"):
        raise ValueError("method must be auto, fft or gather")
    dims = np.asarray(dose_grid_dim, dtype=np.float64).reshape(3)
    cones = _cones(thetas, phis, kernel)
    if cones.n_cones == 0:
        return None
    from .batch import spectral_ok
    shape = tuple(int(v) for v in terma.shape)
    fft = method == "fft" or (method == "auto" and spectral_ok(shape))
    eng = default_engine()
    plan = eng.plan(shape, dims, cones)
    with _CALL_LOCK:
        _run(plan, shape, terma, dose, fft)
    return None


def _run(plan, shape, terma, dose, fft):
    b = _buffers(shape)
    b["t64"].upload(np.ascontiguousarray(terma))
    _convert(b["t"], b["t64"])
    b["dose"].upload(np.ascontiguousarray(dose))
    if fft:
        plan.spectrum()
        _engine.convolve_fft(plan, b["t"], b["part"])
        _convert(b["dose"], b["part"], accumulate=True)
    else:
        _engine.support_box(b["t"], shape, b["box"])
        _engine.convolve_f64(plan, b["t"], b["dose"], b["box"])
    if dose.flags.c_contiguous:
        b["dose"].download(out=dose)
    else:
        np.copyto(dose, b["dose"].download())
