"""Device-resident dose engine: plan store, aperture rasters, phantoms,
TERMA and the collapsed-cone convolution (the direct gather and the FFT
form), all through the C-ABI.

Everything that is reused across beams and requests stays in HBM: cone
plans (tap tables, and the kernel spectrum once the FFT form needs it)
keyed by (grid shape, voxel size, cone set), aperture rasters keyed by
their recipe, phantom densities keyed by the caller.
"""
import collections
import ctypes
import hashlib
import threading

import numpy as np

from . import _lib
from .cones import ConeSet
from .device import DeviceArray, Stream, _h
from .nist import mu_water

SETTINGS_KEYS = ("sPri", "sAnn", "zAnn", "rInner", "rOuter", "zExp", "sExp",
                 "kExp", "softRatio", "softLimit", "hornRatio")

SEMANTICS = {"convolve_c": _lib.CH_SEM_CONVOLVE_C,
             "cuda_dose": _lib.CH_SEM_CUDA_DOSE,
             "physics": _lib.CH_SEM_RADIOLOGICAL}


class Plan:
    """A ch_plan: tap tables of one cone set on one grid, in HBM."""

    def __init__(self, shape, voxel_size, cones):
        L = _lib.lib()
        self.shape = tuple(int(s) for s in shape)
        self.voxel_size = tuple(float(v) for v in voxel_size)
        self.cones = cones
        shp = (ctypes.c_int32 * 3)(*self.shape)
        vs = (ctypes.c_double * 3)(*self.voxel_size)
        dirs = np.ascontiguousarray(cones.dirs, dtype=np.float64)
        rows = np.ascontiguousarray(cones.rows, dtype=np.int32)
        cum = np.ascontiguousarray(cones.cum, dtype=np.float64)
        h = ctypes.c_void_p()
        _lib.check(L.ch_plan_create(
            shp, vs, len(rows), dirs.ctypes.data_as(_lib.c_f64p),
            rows.ctypes.data_as(_lib.c_i32p), cum.shape[0], cum.shape[1],
            cum.ctypes.data_as(_lib.c_f64p), SEMANTICS[cones.semantics],
            ctypes.byref(h)))
        self.handle = h.value
        nc = ctypes.c_int32()
        nt = ctypes.c_int64()
        mx = ctypes.c_int32()
        _lib.check(L.ch_plan_info(self.handle, ctypes.byref(nc),
                                  ctypes.byref(nt), ctypes.byref(mx)))
        self.n_cones, self.n_taps, self.max_taps = nc.value, nt.value, \
            mx.value

    def set_tile(self, tx, ty, tz):
        _lib.check(_lib.lib().ch_plan_set_tile(self.handle, tx, ty, tz))

    def tile(self):
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(_lib.lib().ch_plan_get_tile(self.handle, ctypes.byref(a),
                                               ctypes.byref(b),
                                               ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_fft_lines(self, lines):
        """Lines per workgroup of the FFT form's five passes (z forward,
        y forward, x convolution, y inverse, z inverse); 0 = default."""
        arr = (ctypes.c_int32 * 5)(*[int(v) for v in lines])
        _lib.check(_lib.lib().ch_plan_set_fft_lines(self.handle, arr))

    def fft_lines(self):
        """The lines per workgroup the five passes use."""
        arr = (ctypes.c_int32 * 5)()
        _lib.check(_lib.lib().ch_plan_get_fft_lines(self.handle, arr))
        return tuple(arr)

    def set_accum(self, mode):
        _lib.check(_lib.lib().ch_plan_set_accum(self.handle, int(mode)))

    def cone_taps(self, cone):
        """(offsets int32[M,3], FP64 increments[M]) of one cone."""
        cap = self.max_taps + 1
        off = np.zeros((cap, 3), np.int32)
        k = np.zeros(cap, np.float64)
        n = ctypes.c_int32()
        _lib.check(_lib.lib().ch_plan_cone_taps(
            self.handle, int(cone), off.ctypes.data_as(_lib.c_i32p),
            k.ctypes.data_as(_lib.c_f64p), cap, ctypes.byref(n)))
        return off[:n.value].copy(), k[:n.value].copy()

    def spectrum(self, stream=None):
        """Builds the kernel spectrum the FFT convolution needs (once);
        returns (frame lengths, half-spectrum pitch, work bytes per beam)."""
        _lib.check(_lib.lib().ch_plan_spectrum(self.handle, _h(stream)))
        f = (ctypes.c_int32 * 3)()
        pitch = ctypes.c_int32()
        work = ctypes.c_int64()
        _lib.check(_lib.lib().ch_plan_spectrum_info(
            self.handle, f, ctypes.byref(pitch), ctypes.byref(work)))
        self.spectrum_bytes = 8 * f[0] * f[1] * pitch.value
        return tuple(f), pitch.value, work.value

    def graphs(self):
        """Live captured graphs that read this plan (ch_plan_graphs)."""
        n = ctypes.c_int32()
        _lib.check(_lib.lib().ch_plan_graphs(self.handle, ctypes.byref(n)))
        return n.value

    def release_spectrum(self):
        """Frees the kernel spectrum (HBM); the next FFT use rebuilds it.
        Refused while a live graph reads the plan."""
        _lib.check(_lib.lib().ch_plan_spectrum_release(self.handle))
        self.spectrum_bytes = 0

    def close(self):
        if self.handle:
            _lib.lib().ch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_key(shape, voxel_size, cones):
    h = hashlib.sha1()
    h.update(np.asarray(shape, np.int64).tobytes())
    h.update(np.asarray(voxel_size, np.float64).tobytes())
    h.update(cones_digest(cones))
    return h.hexdigest()


def cones_digest(cones):
    """Content hash of a cone set, computed once per ConeSet (whose arrays
    are not changed after construction)."""
    d = getattr(cones, "_digest", None)
    if d is None:
        h = hashlib.sha1()
        for part in cones.key():
            h.update(part.encode() if isinstance(part, str) else part)
        d = cones._digest = h.digest()
    return d


def support_box(d_terma, shape, d_box, stream=None):
    shp = (ctypes.c_int32 * 3)(*[int(s) for s in shape])
    _lib.check(_lib.lib().ch_support_box(d_terma.ptr, shp, d_box.ptr,
                                         _h(stream)))


def convolve(plan, d_terma, d_dose, d_box=None, accumulate=False,
             stream=None):
    """dose (+)= cone convolution of terma (FP32 device arrays)."""
    _lib.check(_lib.lib().ch_convolve(plan.handle, d_terma.ptr, d_dose.ptr,
                                      d_box.ptr if d_box is not None else None,
                                      1 if accumulate else 0, _h(stream)))


def convolve_fft(plan, d_terma, d_dose, accumulate=False, stream=None):
    """dose (+)= the same convolution as convolve(), computed as one linear
    convolution of T with the plan's summed kernel through zero-padded FFTs
    (ch_convolve_fft).  Builds the plan's spectrum on first use."""
    plan.spectrum(stream)
    _lib.check(_lib.lib().ch_convolve_fft(plan.handle, d_terma.ptr,
                                          d_dose.ptr, 1 if accumulate else 0,
                                          _h(stream)))


def convolve_rad(plan, d_terma, d_density, d_dose, d_box=None,
                 accumulate=False, stream=None):
    """Physics mode: dose (+)= convolution of terma with the kernel looked
    up at the radiological distance through d_density (FP32, same grid).
    The plan must come from a "physics" cone set."""
    _lib.check(_lib.lib().ch_convolve_rad(
        plan.handle, d_terma.ptr, d_density.ptr, d_dose.ptr,
        d_box.ptr if d_box is not None else None, 1 if accumulate else 0,
        _h(stream)))


def convolve_f64(plan, d_terma, d_dose64, d_box=None, stream=None):
    """FP64 dose += cone convolution of FP32 terma."""
    _lib.check(_lib.lib().ch_convolve_f64(plan.handle, d_terma.ptr,
                                          d_dose64.ptr,
                                          d_box.ptr if d_box is not None
                                          else None, _h(stream)))


def convolve_batch(jobs, stream=None):
    """One launch over many (plan, terma, dose, box, accumulate) jobs."""
    n = len(jobs)
    if n == 0:
        return
    arr = (_lib.ConvJob * n)()
    for i, (plan, t, d, box, acc) in enumerate(jobs):
        arr[i].plan = plan.handle
        arr[i].d_terma = t.ptr
        arr[i].d_dose = d.ptr
        arr[i].d_box = box.ptr if box is not None else None
        arr[i].accumulate = 1 if acc else 0
    _lib.check(_lib.lib().ch_convolve_batch(n, arr, _h(stream)))


def convolve_fft_batch(jobs, stream=None):
    """convolve_fft over many (plan, terma, dose, box, accumulate) jobs: the
    jobs of one plan share each pass's launch (ch_convolve_fft_batch); the
    box is not used.  Builds missing plan spectra first."""
    n = len(jobs)
    if n == 0:
        return
    seen = set()
    arr = (_lib.ConvJob * n)()
    for i, (plan, t, d, box, acc) in enumerate(jobs):
        if plan.handle not in seen:
            plan.spectrum(stream)
            seen.add(plan.handle)
        arr[i].plan = plan.handle
        arr[i].d_terma = t.ptr
        arr[i].d_dose = d.ptr
        arr[i].d_box = None
        arr[i].accumulate = 1 if acc else 0
    _lib.check(_lib.lib().ch_convolve_fft_batch(n, arr, _h(stream)))


def convolve_groups(groups, stream=None):
    """Beam groups on shared plans: each item is (plan, [(terma, dose, box),
    ...] with 1..4 beams, accumulate)."""
    n = len(groups)
    if n == 0:
        return
    arr = (_lib.ConvGroup * n)()
    for i, (plan, beams, acc) in enumerate(groups):
        if not 1 <= len(beams) <= _lib.CH_MAX_GROUP:
            raise ValueError("a group holds 1..%d beams" % _lib.CH_MAX_GROUP)
        arr[i].plan = plan.handle
        arr[i].n_beams = len(beams)
        arr[i].accumulate = 1 if acc else 0
        for b, (t, d, box) in enumerate(beams):
            arr[i].d_terma[b] = t.ptr
            arr[i].d_dose[b] = d.ptr
            arr[i].d_box[b] = box.ptr if box is not None else None
    _lib.check(_lib.lib().ch_convolve_groups(n, arr, _h(stream)))


class Graph:
    """A captured pipeline (ch_graph_begin ... ch_graph_end), replayed with
    launch().  Use as:  with Graph.capture(stream) as g: <calls on stream>."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def capture(cls, stream):
        return _Capture(stream)

    def launch(self, stream):
        _lib.check(_lib.lib().ch_graph_launch(self.handle, _h(stream)))

    def close(self):
        if self.handle:
            _lib.lib().ch_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Capture:
    def __init__(self, stream):
        self.stream = stream
        self.graph = None

    def __enter__(self):
        _lib.check(_lib.lib().ch_graph_begin(_h(self.stream)))
        return self

    def __exit__(self, exc_type, exc, tb):
        h = ctypes.c_void_p()
        st = _lib.lib().ch_graph_end(_h(self.stream), ctypes.byref(h))
        if exc_type is None:
            _lib.check(st)
            self.graph = Graph(h.value)
        elif h.value:
            _lib.lib().ch_graph_destroy(h.value)
        return False


class ApertureRef:
    """A ch_aperture plus the device buffers it points into."""

    def __init__(self, kind, n=0, raster=None, table=None,
                 bbox=(0, 4000, 0, 4000)):
        self.raster = raster
        self.table = table
        self.struct = _lib.Aperture()
        self.struct.kind = kind
        self.struct.n = n
        self.struct.d_raster = raster.ptr if raster is not None else None
        self.struct.d_table = table.ptr if table is not None else None
        for i in range(4):
            self.struct.bbox[i] = int(bbox[i])


def beam_struct(source, settings, aperture, d_density, d_terma, d_box):
    """ch_beam for one beam: geometry from `source`, fluence and spectrum
    parameters from the reference's settings dict (example.py:27-58).
    Energies and weights are float32 as in conehead/conehead.py:658-661.
    `aperture` is an ApertureRef or a DeviceArray raster."""
    b = _lib.Beam()
    pos = np.asarray(source.position, dtype=np.float64)
    vy = np.asarray(source.v_y, dtype=np.float64)
    tr = np.asarray(source.transform, dtype=np.float64).reshape(9)
    for i in range(3):
        b.source_position[i] = float(pos[i])
        b.v_y[i] = float(vy[i])
    for i in range(9):
        b.transform[i] = float(tr[i])
    b.sad = float(source.sad)
    for key in SETTINGS_KEYS:
        setattr(b, key, float(settings[key]))
    b.samples = int(settings["fluenceResampling"])
    e, w, mu = spectrum(settings)
    if len(e) > _lib.CH_MAX_ENERGIES:
        raise ValueError("at most %d energy bins" % _lib.CH_MAX_ENERGIES)
    b.n_energy = len(e)
    for i in range(len(e)):
        b.energy[i] = float(e[i])
        b.weight[i] = float(w[i])
        b.mu[i] = float(mu[i])
    if not isinstance(aperture, ApertureRef):
        aperture = ApertureRef(_lib.CH_APERTURE_RASTER, raster=aperture)
    b.aperture = aperture.struct
    b.d_density = d_density.ptr
    b.d_terma = d_terma.ptr
    b.d_box = d_box.ptr if d_box is not None else None
    # the struct holds raw device pointers: it also holds their owners, so
    # an aperture dropped from the engine's cache stays allocated while a
    # beam still points into it
    b.owners = (aperture, d_density, d_terma, d_box)
    return b


def spectrum(settings):
    """(energies, weights, mu) of settings['energy_weights']."""
    ew = settings["energy_weights"]
    e = np.array([np.float32(k) for k in ew.keys()], dtype=np.float32)
    w = np.array([np.float32(v) for v in ew.values()], dtype=np.float32)
    mu = mu_water(e.astype(np.float64))
    return e.astype(np.float64), w.astype(np.float64), mu


def terma(beams, shape, origin, spacing, debug=None, stream=None,
          check=True):
    """Run the fused TERMA kernel for a list of ch_beam structs.  With
    check=False the geometry error of conehead/conehead.py:34-35 (a ray
    parallel to the isocentre plane) is left for terma_check, which reports
    it for every call queued on the stream since the last check."""
    n = len(beams)
    if n == 0:
        return
    arr = (_lib.Beam * n)(*beams)
    shp = (ctypes.c_int32 * 3)(*[int(s) for s in shape])
    org = (ctypes.c_double * 3)(*[float(v) for v in origin])
    spc = (ctypes.c_double * 3)(*[float(v) for v in spacing])
    L = _lib.lib()
    _lib.check(L.ch_terma(n, arr, shp, org, spc,
                          debug.ptr if debug is not None else None,
                          _h(stream)))
    if check:
        terma_check(stream)


class BeamSet:
    """Device-resident TERMA beams on one grid (ch_beamset): the beams'
    kernel parameters live in a device table, so run() queues kernels only
    and can be captured in a Graph; update() rewrites the table for new
    beams with the same outputs.  `keep` holds the Python objects the beams
    point into (apertures, buffers) for as long as the set uses them."""

    def __init__(self, beams, shape, origin, spacing, keep=()):
        n = len(beams)
        if n == 0:
            raise ValueError("a beam set needs beams")
        self.n = n
        arr = (_lib.Beam * n)(*beams)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().ch_beamset_create(
            n, arr, (ctypes.c_int32 * 3)(*[int(v) for v in shape]),
            (ctypes.c_double * 3)(*[float(v) for v in origin]),
            (ctypes.c_double * 3)(*[float(v) for v in spacing]),
            ctypes.byref(h)))
        self.handle = h.value
        self.keep = list(keep)

    def update(self, beams, stream=None, keep=()):
        if len(beams) != self.n:
            raise ValueError("%d beams for a set of %d" % (len(beams), self.n))
        arr = (_lib.Beam * self.n)(*beams)
        _lib.check(_lib.lib().ch_beamset_update(self.handle, self.n, arr,
                                                _h(stream)))
        # the previous beams' objects outlive launches queued before this
        self._prev, self.keep = self.keep, list(keep)

    def run(self, stream=None):
        _lib.check(_lib.lib().ch_terma_beamset(self.handle, _h(stream)))

    def flags(self, stream=None):
        """Per-beam geometry flags (1: a ray in the source plane), cleared."""
        out = np.zeros(self.n, np.int32)
        _lib.check(_lib.lib().ch_beamset_flags(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            _h(stream)))
        return out

    def close(self):
        if self.handle:
            _lib.lib().ch_beamset_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ConvSet:
    """Device-resident FFT-form convolution jobs (ch_convset): the pointer
    tables of convolve_fft_batch written to the device once, so run()
    queues kernels only and can be captured in a Graph.  Jobs are (plan,
    terma, dose, accumulate); missing plan spectra are built here."""

    def __init__(self, jobs, stream=None):
        n = len(jobs)
        if n == 0:
            raise ValueError("a convolution set needs jobs")
        arr = (_lib.ConvJob * n)()
        seen = set()
        for i, (plan, t, d, acc) in enumerate(jobs):
            if plan.handle not in seen:
                plan.spectrum(stream)
                seen.add(plan.handle)
            arr[i].plan = plan.handle
            arr[i].d_terma = t.ptr
            arr[i].d_dose = d.ptr
            arr[i].d_box = None
            arr[i].accumulate = 1 if acc else 0
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().ch_convset_create(n, arr, ctypes.byref(h)))
        self.handle = h.value
        self.keep = [(p, t, d) for p, t, d, _ in jobs]

    def run(self, stream=None):
        _lib.check(_lib.lib().ch_convolve_fft_set(self.handle, _h(stream)))

    def close(self):
        if self.handle:
            _lib.lib().ch_convset_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def terma_check(stream=None):
    """Synchronises `stream`; raises RuntimeError if a ray of any ch_terma
    call on it since the last check lay in the plane through the source
    perpendicular to the beam axis, and clears that flag."""
    _lib.check(_lib.lib().ch_terma_check(_h(stream)))


class Engine:
    """Device-resident caches shared by the API shims, the batched path and
    the request loop (thread-safe)."""

    def __init__(self, device=0):
        self.info = _lib.init(device)
        self._plans = collections.OrderedDict()   # least recently used first
        self._rasters = {}
        self._lock = threading.RLock()

    def trim_spectra(self, budget_bytes, keep=()):
        """Frees the kernel spectra of the least recently used plans, other
        than `keep` and those a live captured graph reads, until the cached
        spectra hold at most budget_bytes of HBM (75 MB per 128^3 plan,
        about 225 MB at 200 x 200 x 160).  Returns the bytes still held."""
        with self._lock:
            keep = {id(p) for p in keep}
            held = sum(getattr(p, "spectrum_bytes", 0)
                       for p in self._plans.values())
            for p in list(self._plans.values()):
                if held <= budget_bytes:
                    break
                b = getattr(p, "spectrum_bytes", 0)
                if b and id(p) not in keep and p.graphs() == 0:
                    p.release_spectrum()
                    held -= b
            return held

    # -- plans -----------------------------------------------------------
    def plan(self, shape, voxel_size, cones, tile=None, accum=None):
        key = plan_key(shape, voxel_size, cones)
        with self._lock:
            p = self._plans.get(key)
            if p is None:
                p = Plan(shape, voxel_size, cones)
                from .autotune import apply_fft, default_tile
                p.set_tile(*(tile or default_tile(p)))
                apply_fft(p)
                self._plans[key] = p
            else:
                self._plans.move_to_end(key)
                if tile is not None:
                    p.set_tile(*tile)
            if accum is not None:
                p.set_accum(accum)
            return p

    # -- apertures -------------------------------------------------------
    @staticmethod
    def _mlc_table(op):
        """Device leaf table and bounding cells of an ("mlc", ...) op."""
        bound, a, b, jx, jy = [np.ascontiguousarray(v, np.int32)
                               for v in op[1:]]
        L = _lib.lib()
        table = DeviceArray((L.ch_mlc_table_len(len(a)),), np.int32)
        bbox = (ctypes.c_int32 * 4)()
        _lib.check(L.ch_mlc_table(
            table.ptr, len(a), bound.ctypes.data_as(_lib.c_i32p),
            a.ctypes.data_as(_lib.c_i32p), b.ctypes.data_as(_lib.c_i32p),
            jx.ctypes.data_as(_lib.c_i32p), jy.ctypes.data_as(_lib.c_i32p),
            bbox))
        return table, tuple(bbox)

    def raster(self, block, stream=None):
        """Device raster of a Block, built on the GPU from its recipe."""
        key = block.recipe()
        with self._lock:
            d = self._rasters.get(("raster", key))
            if d is not None:
                return d
            d = DeviceArray((4000, 4000), np.float32)
            L = _lib.lib()
            if key[0] == "host":
                d.upload(block._values, stream)
            else:
                for op in key:
                    if op[0] == "zeros":
                        d.zero(stream)
                    elif op[0] == "square":
                        _lib.check(L.ch_raster_square(d.ptr, op[1],
                                                      _h(stream)))
                    elif op[0] == "mlc":
                        table, _ = self._mlc_table(op)
                        _lib.check(L.ch_raster_mlc(d.ptr, table.ptr,
                                                   _h(stream)))
                        _lib.check(L.ch_stream_sync(_h(stream)))
            self._cache_put(("raster", key), d)
            return d

    def aperture(self, block, stream=None):
        """ApertureRef of a Block: analytic rectangles for squares, a leaf
        table for an MLC, the raster otherwise (host-edited rasters)."""
        from .block import square_bounds
        key = block.recipe()
        with self._lock:
            ap = self._rasters.get(("aperture", key))
            if ap is not None:
                return ap
            ops = list(key) if key[0] != "host" else None
            if ops is not None and all(op[0] in ("zeros", "square")
                                       for op in ops):
                rects = []
                for op in ops:
                    if op[0] == "zeros":
                        rects = []
                    else:
                        lo, hi = square_bounds(op[1])
                        lo, hi = max(lo, 0), min(hi, 4000)
                        if hi > lo:
                            rects.append((lo, hi, lo, hi))
                if rects:
                    r = np.array(rects, np.int32)
                    table = DeviceArray.from_host(r)
                    bbox = (r[:, 0].min(), r[:, 1].max(), r[:, 2].min(),
                            r[:, 3].max())
                else:
                    table, bbox = None, (0, 0, 0, 0)
                ap = ApertureRef(_lib.CH_APERTURE_RECTS, len(rects),
                                 table=table, bbox=bbox)
            elif ops is not None and len(ops) == 1 and ops[0][0] == "mlc":
                table, bbox = self._mlc_table(ops[0])
                ap = ApertureRef(_lib.CH_APERTURE_MLC, table=table,
                                 bbox=bbox)
            else:
                raster = self.raster(block, stream)
                bbox = (0, 4000, 0, 4000)
                if key[0] == "host":
                    nz = np.nonzero(block._values)
                    bbox = ((nz[0].min(), nz[0].max() + 1, nz[1].min(),
                             nz[1].max() + 1) if len(nz[0]) else
                            (0, 0, 0, 0))
                ap = ApertureRef(_lib.CH_APERTURE_RASTER, raster=raster,
                                 bbox=bbox)
            self._cache_put(("aperture", key), ap)
            return ap

    def _cache_put(self, key, value):
        if len(self._rasters) > 512:
            self._rasters.clear()
        self._rasters[key] = value

    def drop_caches(self):
        with self._lock:
            self._plans.clear()
            self._rasters.clear()
