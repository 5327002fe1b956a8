"""ctypes wrapper around the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker and as the timed CPU baseline.
The product package (conehead_amd) never imports it.

Each wrapper follows the C function of the same name in oracle/ch_ref.c,
which cites the reference file:line it restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)


class RefBeam(ctypes.Structure):
    """Mirror of ch_ref_beam in oracle/ch_ref.h."""

    _fields_ = [
        ("source_position", ctypes.c_double * 3),
        ("v_y", ctypes.c_double * 3),
        ("transform", ctypes.c_double * 9),
        ("sad", ctypes.c_double),
        ("sPri", ctypes.c_double),
        ("sAnn", ctypes.c_double),
        ("zAnn", ctypes.c_double),
        ("rInner", ctypes.c_double),
        ("rOuter", ctypes.c_double),
        ("zExp", ctypes.c_double),
        ("sExp", ctypes.c_double),
        ("kExp", ctypes.c_double),
        ("softRatio", ctypes.c_double),
        ("softLimit", ctypes.c_double),
        ("hornRatio", ctypes.c_double),
        ("samples", ctypes.c_int32),
        ("n_energy", ctypes.c_int32),
        ("energy", ctypes.c_double * 32),
        ("weight", ctypes.c_double * 32),
        ("mu", ctypes.c_double * 32),
    ]


def build(force=False):
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-B" if force else "-s", "-C", _HERE],
                              stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ch_ref_dda_3d.restype = ctypes.c_int32
        L.ch_ref_dda_3d.argtypes = [_f64p, _i32p, _i32p, _f64p, _i32p, _f64p,
                                    ctypes.c_int32, _f64p]
        L.ch_ref_cone_direction.restype = None
        L.ch_ref_cone_direction.argtypes = [ctypes.c_double, ctypes.c_double,
                                            _f64p]
        L.ch_ref_kernel_index.restype = ctypes.c_int32
        L.ch_ref_kernel_index.argtypes = [ctypes.c_double]
        L.ch_ref_convolve_dirs.restype = ctypes.c_int64
        L.ch_ref_convolve_dirs.argtypes = [_f64p, _f64p, _i32p, _f64p,
                                           ctypes.c_int32, _f64p, _i32p,
                                           ctypes.c_int32, _f64p,
                                           ctypes.c_int32]
        L.ch_ref_convolve.restype = ctypes.c_int64
        L.ch_ref_convolve.argtypes = [_f64p, _f64p, _i32p, _f64p, _f64p,
                                      ctypes.c_int32, _f64p, ctypes.c_int32,
                                      ctypes.c_int32, _f64p, ctypes.c_int32]
        L.ch_ref_convolve_gather.restype = ctypes.c_int64
        L.ch_ref_convolve_gather.argtypes = [_f64p, _f64p, _i32p, _f64p,
                                             ctypes.c_int32, _f64p, _i32p,
                                             ctypes.c_int32, _f64p]
        L.ch_ref_cone_taps.restype = ctypes.c_int32
        L.ch_ref_cone_taps.argtypes = [_f64p, _i32p, _f64p, ctypes.c_int32,
                                       _f64p, _i32p, _f64p, ctypes.c_int32]
        L.ch_ref_kernel_index_cuda.restype = ctypes.c_int32
        L.ch_ref_kernel_index_cuda.argtypes = [ctypes.c_double]
        L.ch_ref_cone_path.restype = ctypes.c_int32
        L.ch_ref_cone_path.argtypes = [_f64p, _i32p, _f64p, _i32p, _f64p,
                                       ctypes.c_int32]
        L.ch_ref_convolve_rad.restype = ctypes.c_int64
        L.ch_ref_convolve_rad.argtypes = [_f64p, _f32p, _f64p, _i32p, _f64p,
                                          ctypes.c_int32, _f64p, _i32p,
                                          ctypes.c_int32, _f64p,
                                          ctypes.c_int32, ctypes.c_int32]
        L.ch_ref_terma.restype = ctypes.c_int32
        L.ch_ref_terma.argtypes = [ctypes.POINTER(RefBeam), _i32p, _f64p,
                                   _f64p, _f32p, _f32p, _f64p, _f64p, _f64p,
                                   _f64p, _f64p, ctypes.c_int32]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _c64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def _c32i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def dda_3d(direction, grid_shape, first_voxel, voxel_size):
    """Oracle 3D-DDA (conehead/dda_3d_c.pyx:33-99).

    Returns (voxels int32[N,3], t float64[N], |direction| float64[3])."""
    d = _c64(direction, (3,))
    shp = _c32i(grid_shape)
    fv = _c32i(first_voxel)
    vs = _c64(voxel_size, (3,))
    cap = int(max(shp.sum(), 0)) + 2
    vox = np.zeros((cap, 3), np.int32)
    tv = np.zeros(cap, np.float64)
    dabs = np.zeros(3, np.float64)
    n = lib().ch_ref_dda_3d(_p(d, _f64p), _p(shp, _i32p), _p(fv, _i32p),
                            _p(vs, _f64p), _p(vox, _i32p), _p(tv, _f64p),
                            cap, _p(dabs, _f64p))
    if n > cap:
        raise RuntimeError("oracle DDA overflow")
    return vox[:n].copy(), tv[:n].copy(), dabs


def cone_direction(theta_deg, phi_deg):
    """Discretised cone direction (conehead/convolve_c.pyx:95-117)."""
    out = np.zeros(3, np.float64)
    lib().ch_ref_cone_direction(float(theta_deg), float(phi_deg),
                                _p(out, _f64p))
    return out


def kernel_index(t):
    return int(lib().ch_ref_kernel_index(float(t)))


def cum_rows(kernel, phis):
    """Rows of kernel.cumulative keyed by "{:.2f}".format(phi)
    (conehead/convolve_c.pyx:257-274)."""
    rows = [np.asarray(kernel.cumulative["{:.2f}".format(float(p))],
                       dtype=np.float64) for p in phis]
    n = min(len(r) for r in rows)
    return np.ascontiguousarray(np.stack([r[:n] for r in rows]))


def convolve(terma, dose, dims, thetas, phis, kernel, n_threads=1):
    """Oracle convolve_c (conehead/convolve_c.pyx:17-197), scatter form.

    `dose` (float64, C-contiguous) is accumulated in place.  Returns the
    number of dose += T*k executions."""
    terma = _c64(terma)
    assert dose.dtype == np.float64 and dose.flags.c_contiguous
    shp = _c32i(terma.shape)
    thetas = _c64(thetas)
    phis = _c64(phis)
    cum = cum_rows(kernel, phis)
    r = lib().ch_ref_convolve(_p(terma, _f64p), _p(dose, _f64p),
                              _p(shp, _i32p), _p(_c64(dims, (3,)), _f64p),
                              _p(thetas, _f64p), len(thetas), _p(phis, _f64p),
                              len(phis), cum.shape[1], _p(cum, _f64p),
                              int(n_threads))
    if r < 0:
        raise RuntimeError("oracle convolve failed (%d)" % r)
    return int(r)


def convolve_dirs(terma, dose, dims, dirs, cone_row, cum, n_threads=1):
    """Scatter form with explicit cone directions and kernel rows."""
    terma = _c64(terma)
    assert dose.dtype == np.float64 and dose.flags.c_contiguous
    shp = _c32i(terma.shape)
    dirs = _c64(dirs).reshape(-1, 3)
    rows = _c32i(cone_row)
    cum = _c64(cum)
    r = lib().ch_ref_convolve_dirs(_p(terma, _f64p), _p(dose, _f64p),
                                   _p(shp, _i32p),
                                   _p(_c64(dims, (3,)), _f64p), len(dirs),
                                   _p(dirs, _f64p), _p(rows, _i32p),
                                   cum.shape[1], _p(cum, _f64p),
                                   int(n_threads))
    if r < 0:
        raise RuntimeError("oracle convolve failed (%d)" % r)
    return int(r)


def convolve_gather(terma, dose, dims, dirs, cone_row, cum):
    """Stencil (gather) form of the same convolution."""
    terma = _c64(terma)
    shp = _c32i(terma.shape)
    dirs = _c64(dirs).reshape(-1, 3)
    rows = _c32i(cone_row)
    cum = _c64(cum)
    r = lib().ch_ref_convolve_gather(_p(terma, _f64p), _p(dose, _f64p),
                                     _p(shp, _i32p),
                                     _p(_c64(dims, (3,)), _f64p), len(dirs),
                                     _p(dirs, _f64p), _p(rows, _i32p),
                                     cum.shape[1], _p(cum, _f64p))
    if r < 0:
        raise RuntimeError("oracle gather failed (%d)" % r)
    return int(r)


def cone_taps(direction, shape, dims, cum_row):
    """Tap table (offsets int32[M,3], increments float64[M]) of one cone."""
    d = _c64(direction, (3,))
    shp = _c32i(shape)
    cum_row = _c64(cum_row)
    cap = int(shp.sum()) + 2
    off = np.zeros((cap, 3), np.int32)
    k = np.zeros(cap, np.float64)
    n = lib().ch_ref_cone_taps(_p(d, _f64p), _p(shp, _i32p),
                               _p(_c64(dims, (3,)), _f64p), len(cum_row),
                               _p(cum_row, _f64p), _p(off, _i32p),
                               _p(k, _f64p), cap)
    if n < 0:
        raise RuntimeError("oracle tap overflow")
    return off[:n].copy(), k[:n].copy()


def kernel_index_cuda(r):
    """cuda_dose kernel index |floor(100 r - 5)| clamped at 5995."""
    return int(lib().ch_ref_kernel_index_cuda(float(r)))


def cone_path(direction, shape, dims):
    """DDA path of one cone from the origin of an unbounded grid while it
    can land inside `shape`: offsets int32[M,3] and exit t-values
    float64[M]."""
    d = _c64(direction, (3,))
    shp = _c32i(shape)
    cap = int(shp.sum()) + 2
    off = np.zeros((cap, 3), np.int32)
    t = np.zeros(cap, np.float64)
    n = lib().ch_ref_cone_path(_p(d, _f64p), _p(shp, _i32p),
                               _p(_c64(dims, (3,)), _f64p), _p(off, _i32p),
                               _p(t, _f64p), cap)
    if n < 0:
        raise RuntimeError("oracle path overflow")
    return off[:n].copy(), t[:n].copy()


def convolve_rad(terma, density, dose, dims, dirs, cone_row, cum,
                 radiological=True, n_threads=1):
    """Physics-mode gather (radiological kernel distance, cuda_dose index);
    radiological=False gives the cuda_dose convolution in gather form.
    `dose` (float64, C-contiguous) is accumulated in place; `density` is
    float32 on the same grid."""
    terma = _c64(terma)
    assert dose.dtype == np.float64 and dose.flags.c_contiguous
    assert dose.shape == terma.shape
    shp = _c32i(terma.shape)
    dens = np.ascontiguousarray(
        np.zeros(terma.shape, np.float32) if density is None else density,
        dtype=np.float32)
    assert dens.shape == terma.shape
    dirs = _c64(dirs).reshape(-1, 3)
    rows = _c32i(cone_row)
    cum = _c64(cum)
    r = lib().ch_ref_convolve_rad(_p(terma, _f64p), _p(dens, _f32p),
                                  _p(dose, _f64p), _p(shp, _i32p),
                                  _p(_c64(dims, (3,)), _f64p), len(dirs),
                                  _p(dirs, _f64p), _p(rows, _i32p),
                                  cum.shape[1], _p(cum, _f64p),
                                  int(bool(radiological)), int(n_threads))
    if r < 0:
        raise RuntimeError("oracle physics convolution failed (%d)" % r)
    return int(r)


def make_beam(source_position, v_y, transform, sad, settings, energy,
              weight, mu):
    b = RefBeam()
    for i in range(3):
        b.source_position[i] = float(source_position[i])
        b.v_y[i] = float(v_y[i])
    tr = np.asarray(transform, dtype=np.float64).reshape(9)
    for i in range(9):
        b.transform[i] = float(tr[i])
    b.sad = float(sad)
    for key in ("sPri", "sAnn", "zAnn", "rInner", "rOuter", "zExp", "sExp",
                "kExp", "softRatio", "softLimit", "hornRatio"):
        setattr(b, key, float(settings[key]))
    b.samples = int(settings["fluenceResampling"])
    b.n_energy = len(energy)
    for i in range(len(energy)):
        b.energy[i] = float(energy[i])
        b.weight[i] = float(weight[i])
        b.mu[i] = float(mu[i])
    return b


def terma(beam, shape, origin, spacing, density, raster, n_threads=1):
    """Oracle TERMA chain (conehead/conehead.py:71-377, :641-655).

    Returns dict of float64 grids: terma, blocked, d_eff, oad, fluence."""
    shp = _c32i(shape)
    density = np.ascontiguousarray(density, dtype=np.float32)
    raster = np.ascontiguousarray(raster, dtype=np.float32)
    assert raster.shape == (4000, 4000)
    out = {k: np.zeros(tuple(shp), np.float64)
           for k in ("terma", "blocked", "d_eff", "oad", "fluence")}
    r = lib().ch_ref_terma(ctypes.byref(beam), _p(shp, _i32p),
                           _p(_c64(origin, (3,)), _f64p),
                           _p(_c64(spacing, (3,)), _f64p),
                           _p(density, _f32p), _p(raster, _f32p),
                           _p(out["terma"], _f64p), _p(out["blocked"], _f64p),
                           _p(out["d_eff"], _f64p), _p(out["oad"], _f64p),
                           _p(out["fluence"], _f64p), int(n_threads))
    if r != 0:
        raise RuntimeError("no intersection or line is within plane")
    return out
