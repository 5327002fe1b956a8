"""ctypes binding of libconehead_amd.so (the C-ABI in include/conehead_amd.h).

The library is built in-tree by conehead_amd/build.py.  There is no
fallback: if the shared library is missing, or no GPU is present when a
compute entry point is called, an error is raised.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
# CONEHEAD_AMD_LIB selects another build of the library (same C-ABI).
LIB_PATH = os.environ.get("CONEHEAD_AMD_LIB",
                          os.path.join(HERE, "libconehead_amd.so"))

CH_OK = 0
CH_ERR_ARG = -1
CH_ERR_CAPACITY = -2
CH_ERR_HIP = -3
CH_ERR_NOMEM = -4
CH_ERR_GEOMETRY = -5

CH_SEM_CONVOLVE_C = 0
CH_SEM_CUDA_DOSE = 1
CH_SEM_RADIOLOGICAL = 2
CH_MAX_ENERGIES = 32

c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_f32p = ctypes.POINTER(ctypes.c_float)
c_vp = ctypes.c_void_p
c_vpp = ctypes.POINTER(ctypes.c_void_p)
c_size = ctypes.c_size_t


class CHError(RuntimeError):
    """A C-ABI call returned a non-zero ch_status."""

    def __init__(self, status, message):
        super().__init__("%s (status %d)" % (message, status))
        self.status = status


CH_APERTURE_RASTER = 0
CH_APERTURE_RECTS = 1
CH_APERTURE_MLC = 2
CH_F32 = 0          # ch_convert element types
CH_F64 = 1


class Aperture(ctypes.Structure):
    """Mirror of ch_aperture."""

    _fields_ = [
        ("kind", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("d_raster", c_vp),
        ("d_table", c_vp),
        ("bbox", ctypes.c_int32 * 4),
    ]


class Beam(ctypes.Structure):
    """Mirror of ch_beam."""

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
        ("energy", ctypes.c_double * CH_MAX_ENERGIES),
        ("weight", ctypes.c_double * CH_MAX_ENERGIES),
        ("mu", ctypes.c_double * CH_MAX_ENERGIES),
        ("aperture", Aperture),
        ("d_density", c_vp),
        ("d_terma", c_vp),
        ("d_box", c_vp),
        ("accumulate", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("d_status", c_vp),
    ]


class ConvJob(ctypes.Structure):
    """Mirror of ch_conv_job."""

    _fields_ = [
        ("plan", c_vp),
        ("d_terma", c_vp),
        ("d_dose", c_vp),
        ("d_box", c_vp),
        ("accumulate", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


CH_MAX_GROUP = 4


class ConvGroup(ctypes.Structure):
    """Mirror of ch_conv_group."""

    _fields_ = [
        ("plan", c_vp),
        ("n_beams", ctypes.c_int32),
        ("accumulate", ctypes.c_int32),
        ("d_terma", c_vp * CH_MAX_GROUP),
        ("d_dose", c_vp * CH_MAX_GROUP),
        ("d_box", c_vp * CH_MAX_GROUP),
    ]


# name -> (restype, argtypes); every function declared in the header.
SIGNATURES = {
    "ch_last_error": (ctypes.c_char_p, []),
    "ch_version": (ctypes.c_int32, []),
    "ch_init": (ctypes.c_int32, [ctypes.c_int32, c_i32p, c_i64p,
                                 ctypes.c_char_p, ctypes.c_int32]),
    "ch_malloc": (ctypes.c_int32, [c_vpp, c_size]),
    "ch_free": (ctypes.c_int32, [c_vp]),
    "ch_mem_info": (ctypes.c_int32, [ctypes.POINTER(c_size),
                                     ctypes.POINTER(c_size)]),
    "ch_memcpy_h2d": (ctypes.c_int32, [c_vp, c_vp, c_size, c_vp]),
    "ch_memcpy_d2h": (ctypes.c_int32, [c_vp, c_vp, c_size, c_vp]),
    "ch_memcpy_d2h_async": (ctypes.c_int32, [c_vp, c_vp, c_size, c_vp]),
    "ch_memcpy_d2d": (ctypes.c_int32, [c_vp, c_vp, c_size, c_vp]),
    "ch_zero": (ctypes.c_int32, [c_vp, c_size, c_vp]),
    "ch_host_alloc": (ctypes.c_int32, [c_vpp, c_size]),
    "ch_host_free": (ctypes.c_int32, [c_vp]),
    "ch_stream_create": (ctypes.c_int32, [c_vpp]),
    "ch_stream_destroy": (ctypes.c_int32, [c_vp]),
    "ch_stream_sync": (ctypes.c_int32, [c_vp]),
    "ch_device_sync": (ctypes.c_int32, []),
    "ch_event_create": (ctypes.c_int32, [c_vpp]),
    "ch_event_destroy": (ctypes.c_int32, [c_vp]),
    "ch_event_record": (ctypes.c_int32, [c_vp, c_vp]),
    "ch_stream_wait": (ctypes.c_int32, [c_vp, c_vp]),
    "ch_event_sync": (ctypes.c_int32, [c_vp]),
    "ch_event_query": (ctypes.c_int32, [c_vp, c_i32p]),
    "ch_event_elapsed": (ctypes.c_int32, [c_f32p, c_vp, c_vp]),
    "ch_dda_3d": (ctypes.c_int32, [c_f64p, c_i32p, c_i32p, c_f64p, c_i32p,
                                   c_f64p, ctypes.c_int32, c_i32p, c_f64p]),
    "ch_dda_trace_batch": (ctypes.c_int32, [ctypes.c_int32, c_vp, c_vp,
                                            c_i32p, c_f64p, ctypes.c_int32,
                                            c_vp, c_vp, c_vp, c_vp]),
    "ch_plan_create": (ctypes.c_int32, [c_i32p, c_f64p, ctypes.c_int32,
                                        c_f64p, c_i32p, ctypes.c_int32,
                                        ctypes.c_int32, c_f64p,
                                        ctypes.c_int32, c_vpp]),
    "ch_plan_destroy": (ctypes.c_int32, [c_vp]),
    "ch_plan_graphs": (ctypes.c_int32, [c_vp, c_i32p]),
    "ch_plan_info": (ctypes.c_int32, [c_vp, c_i32p, c_i64p, c_i32p]),
    "ch_plan_cone_taps": (ctypes.c_int32, [c_vp, ctypes.c_int32, c_i32p,
                                           c_f64p, ctypes.c_int32, c_i32p]),
    "ch_support_box": (ctypes.c_int32, [c_vp, c_i32p, c_vp, c_vp]),
    "ch_convolve": (ctypes.c_int32, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int32,
                                     c_vp]),
    "ch_convolve_f64": (ctypes.c_int32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "ch_plan_spectrum": (ctypes.c_int32, [c_vp, c_vp]),
    "ch_plan_spectrum_release": (ctypes.c_int32, [c_vp]),
    "ch_plan_spectrum_info": (ctypes.c_int32, [c_vp, c_i32p, c_i32p,
                                               c_i64p]),
    "ch_convolve_fft": (ctypes.c_int32, [c_vp, c_vp, c_vp, ctypes.c_int32,
                                         c_vp]),
    "ch_convolve_fft_batch": (ctypes.c_int32, [ctypes.c_int32,
                                               ctypes.POINTER(ConvJob),
                                               c_vp]),
    "ch_convolve_batch": (ctypes.c_int32, [ctypes.c_int32,
                                           ctypes.POINTER(ConvJob), c_vp]),
    "ch_set_launch_events": (ctypes.c_int32, [c_vp, c_vp]),
    "ch_convolve_rad": (ctypes.c_int32, [
        c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int32, c_vp]),
    "ch_resample_linear": (ctypes.c_int32, [
        c_vp, c_i32p, c_f64p, c_f64p, c_vp, c_i32p, c_f64p, c_f64p,
        ctypes.c_int32, c_vp]),
    "ch_ct_to_density": (ctypes.c_int32, [
        c_vp, ctypes.c_int32, ctypes.c_int64, c_f64p, c_f64p, ctypes.c_int32,
        c_vp, ctypes.POINTER(ctypes.c_float), c_vp, c_vp]),
    "ch_beamset_create": (ctypes.c_int32, [
        ctypes.c_int32, ctypes.POINTER(Beam), c_i32p, c_f64p, c_f64p,
        ctypes.POINTER(c_vp)]),
    "ch_beamset_update": (ctypes.c_int32, [c_vp, ctypes.c_int32,
                                           ctypes.POINTER(Beam), c_vp]),
    "ch_terma_beamset": (ctypes.c_int32, [c_vp, c_vp]),
    "ch_beamset_flags": (ctypes.c_int32, [c_vp, c_i32p, c_vp]),
    "ch_beamset_destroy": (ctypes.c_int32, [c_vp]),
    "ch_convset_create": (ctypes.c_int32, [
        ctypes.c_int32, ctypes.POINTER(ConvJob), ctypes.POINTER(c_vp)]),
    "ch_convolve_fft_set": (ctypes.c_int32, [c_vp, c_vp]),
    "ch_convset_destroy": (ctypes.c_int32, [c_vp]),
    "ch_convert": (ctypes.c_int32, [c_vp, ctypes.c_int32, c_vp,
                                    ctypes.c_int32, ctypes.c_size_t,
                                    ctypes.c_int32, c_vp]),
    "ch_graph_begin": (ctypes.c_int32, [c_vp]),
    "ch_graph_end": (ctypes.c_int32, [c_vp, ctypes.POINTER(c_vp)]),
    "ch_graph_launch": (ctypes.c_int32, [c_vp, c_vp]),
    "ch_graph_destroy": (ctypes.c_int32, [c_vp]),
    "ch_convolve_groups": (ctypes.c_int32, [ctypes.c_int32,
                                            ctypes.POINTER(ConvGroup), c_vp]),
    "ch_convolve_scatter": (ctypes.c_int32, [
        c_vp, c_vp, c_i32p, c_f64p, ctypes.c_int32, c_vp, c_vp, c_vp,
        ctypes.c_int32, ctypes.c_int32, c_vp]),
    "ch_plan_set_tile": (ctypes.c_int32, [c_vp, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_int32]),
    "ch_plan_get_tile": (ctypes.c_int32, [c_vp, c_i32p, c_i32p, c_i32p]),
    "ch_plan_set_fft_lines": (ctypes.c_int32, [c_vp, c_i32p]),
    "ch_plan_get_fft_lines": (ctypes.c_int32, [c_vp, c_i32p]),
    "ch_plan_set_accum": (ctypes.c_int32, [c_vp, ctypes.c_int32]),
    "ch_terma": (ctypes.c_int32, [ctypes.c_int32, ctypes.POINTER(Beam),
                                  c_i32p, c_f64p, c_f64p, c_vp, c_vp]),
    "ch_terma_check": (ctypes.c_int32, [c_vp]),
    "ch_terma_noalloc": (ctypes.c_int32, [ctypes.c_int32,
                                          ctypes.POINTER(Beam), c_i32p,
                                          c_f64p, c_f64p, c_vp, c_vp]),
    "ch_terma_rays": (ctypes.c_int32, [ctypes.POINTER(Beam), c_i32p, c_f64p,
                                       c_f64p, ctypes.c_int32, c_vp,
                                       ctypes.c_int32, c_vp, c_vp, c_vp, c_vp,
                                       c_vp]),
    "ch_raster_square": (ctypes.c_int32, [c_vp, ctypes.c_double, c_vp]),
    "ch_mlc_table_len": (ctypes.c_int32, [ctypes.c_int32]),
    "ch_mlc_table": (ctypes.c_int32, [c_vp, ctypes.c_int32, c_i32p, c_i32p,
                                      c_i32p, c_i32p, c_i32p, c_i32p]),
    "ch_raster_mlc": (ctypes.c_int32, [c_vp, c_vp, c_vp]),
    "ch_set_timing": (ctypes.c_int32, [ctypes.c_int32]),
    "ch_last_kernel_ms": (ctypes.c_int32, [c_f32p]),
}

_lib = None
_lock = threading.Lock()
_device_ready = False


def load():
    """Load the shared library (no device initialisation)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    "libconehead_amd.so is not built; run "
                    "`python -m conehead_amd.build` (hipcc, gfx950)")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def last_error():
    msg = load().ch_last_error()
    return msg.decode() if msg else ""


def check(status):
    if status != CH_OK:
        raise CHError(status, last_error())
    return status


def lib():
    """The library with the GPU initialised (raises without a GPU)."""
    global _device_ready
    L = load()
    if not _device_ready:
        init(0)
    return L


_device_info = {}


def init(device=0):
    global _device_ready
    L = load()
    n_cu = ctypes.c_int32()
    hbm = ctypes.c_int64()
    name = ctypes.create_string_buffer(128)
    check(L.ch_init(device, ctypes.byref(n_cu), ctypes.byref(hbm), name, 128))
    _device_info.update(device=device, n_cu=n_cu.value, hbm_bytes=hbm.value,
                        arch=name.value.decode())
    _device_ready = True
    return dict(_device_info)


def device_info():
    if not _device_ready:
        init(0)
    return dict(_device_info)
