"""Batched dose engine: many phantoms, beams and cone sets per launch.

A DoseRequest is one phantom (densities on its dose grid), a list of beams
and a cone set.  A batch of requests runs one TERMA launch per beam and
then the convolutions:
  method="fft" (default)  the FFT form, ch_convolve_fft_batch: five passes
                          per beam, the jobs on one plan sharing each
                          launch; TERMA chunks overlap the previous chunk's
                          passes on a companion stream;
  method="gather"         one gather launch over every job of the batch
                          (jobs of different grids, voxel sizes and cone
                          sets share it; ch_convolve_batch), plus one
                          launch for the beams convolved in interleaved
                          groups (beams of a request whose support boxes
                          nearly coincide; ch_convolve_groups).
Two output modes:
  per_beam=True   one dose grid per beam (each beam is its own job);
  per_beam=False  the plan dose of all beams: their TERMA grids are summed
                  (ch_terma accumulate mode) and convolved once -- exact,
                  because convolve_c is linear in TERMA and its cones are
                  fixed to the grid (conehead/convolve_c.pyx:101).
Device buffers come from a size-keyed pool, densities can stay resident
under a caller key, both convolutions are deterministic and groups are
chosen per request from its own beams, so a request's result is bitwise
the same alone or inside any batch.
"""
import threading

import numpy as np

from . import _lib
from . import engine as E
from .device import DeviceArray, Event, HostPool, Stream, _h

PAD_Y = 32   # kYPad of the padded TERMA copy (ch_internal.h)
PAD_Z = 80   # kZPad


def _smooth_at_least(n):
    m = max(int(n), 1)
    while True:
        k = m
        for q in (2, 3, 5):
            while k % q == 0:
                k //= q
        if k == 1:
            return m
        m += 1


def spectral_frame(shape):
    """Frame lengths (smallest 2, 3, 5-smooth F >= 2 n - 1 per axis) and
    half-spectrum pitch of the FFT form, as ch_plan_spectrum picks them."""
    frame = tuple(_smooth_at_least(2 * int(n) - 1) for n in shape)
    kz = frame[2] // 2 + 1
    return frame, (kz + 15) // 16 * 16


SPECTRAL_MAX_FRAME = 1024   # longest transform of ch_plan_spectrum


def spectral_ok(shape):
    """True when the FFT form takes this grid (every frame <= 1024, i.e.
    axes of at most 512 voxels)."""
    return max(spectral_frame(shape)[0]) <= SPECTRAL_MAX_FRAME


class DoseRequest:
    """One dose calculation: phantom + beams + cones."""

    def __init__(self, densities, origin, spacing, beams, cones,
                 per_beam=True, key=None):
        self.densities = np.ascontiguousarray(densities, np.float32)
        self.shape = tuple(int(s) for s in self.densities.shape)
        self.origin = np.asarray(origin, np.float64)
        self.spacing = np.asarray(spacing, np.float64)
        self.beams = list(beams)
        self.cones = cones
        self.per_beam = per_beam
        self.key = key

    @classmethod
    def from_case(cls, case, per_beam=True, key=None):
        return cls(case.densities, case.origin, case.spacing, case.beams,
                   case.cones, per_beam, key)

    @property
    def n_voxels(self):
        return int(np.prod(self.shape))

    @property
    def n_jobs(self):
        return len(self.beams) if self.per_beam else 1

    def device_bytes(self, method="gather"):
        """HBM this request holds while in flight: densities, TERMA, dose
        and box per job, plus the convolution's work space: the padded
        TERMA of the gather workspace (one padded voxel per beam, in
        interleaved groups of four) or, for the FFT form, one complex
        half-spectrum work grid X x Fy x Kp per job."""
        X, Y, Z = self.shape
        nv = self.n_voxels
        base = 4 * (nv + self.n_jobs * (2 * nv + 8))
        if method == "fft" and spectral_ok(self.shape):
            frame, kp = spectral_frame(self.shape)
            return base + 8 * self.n_jobs * X * frame[1] * kp
        padded = X * (Y + 2 * PAD_Y) * (Z + 2 * PAD_Z)
        slots = -(-self.n_jobs // 4) * 4
        return base + 4 * slots * padded

    def voxel_cone_updates(self):
        """Dense work count Nv x C per beam, summed over the beams."""
        return self.n_voxels * self.cones.n_cones * len(self.beams)


class DevicePool:
    """Reusable device buffers keyed by (shape, dtype)."""

    def __init__(self, max_free_bytes=64 << 30):
        self._free = {}
        self._bytes = 0
        self._max = max_free_bytes
        self._lock = threading.Lock()

    def get(self, shape, dtype=np.float32):
        key = (tuple(shape), np.dtype(dtype).str)
        with self._lock:
            lst = self._free.get(key)
            if lst:
                a = lst.pop()
                self._bytes -= a.nbytes
                return a
        return DeviceArray(tuple(shape), dtype)

    def put(self, arr):
        key = (arr.shape, arr.dtype.str)
        with self._lock:
            if self._bytes + arr.nbytes > self._max:
                arr.free()
                return
            self._free.setdefault(key, []).append(arr)
            self._bytes += arr.nbytes

    def clear(self):
        with self._lock:
            for lst in self._free.values():
                for a in lst:
                    a.free()
            self._free.clear()
            self._bytes = 0


class BatchHandle:
    """A launched batch: its completion event and device outputs."""

    def __init__(self, requests, outputs, temps, event, stream,
                 staged=None, flags=None):
        self.requests = requests
        self.outputs = outputs     # per request: list of dose DeviceArrays
        self.temps = temps         # buffers to return to the pool
        self.event = event
        self.stream = stream
        self.staged = staged       # per request: pinned HostBuffer, or None
        self.flags = flags         # per request: TERMA geometry flag (int32)

    def done(self):
        return self.event.done()


def _lib_error(message):
    return _lib.CHError(_lib.CH_ERR_GEOMETRY, message)


# Measured cost of one interleaved group gather relative to one single-beam
# gather over the same support box (four beam slots are always loaded).
GROUP_COST = 2.4


def _volume(lo, hi):
    return float(np.prod(np.maximum(0, np.asarray(hi) - np.asarray(lo) + 1)))


def group_beams(boxes, max_group=4, saving=0.9):
    """Partition beam indices into interleaved groups of 3..max_group
    beams whose union support box makes the group cheaper than single
    gathers (GROUP_COST x union volume <= saving x the sum of their
    volumes); every other beam stays single.  boxes: per beam (lo x, y, z,
    hi x, y, z).  Greedy, largest box first, each step adding the beam that
    grows the union least."""
    B = [np.asarray(b, np.int64) for b in boxes]
    V = [_volume(b[:3], b[3:]) for b in B]
    rem = sorted(range(len(B)), key=lambda i: -V[i])
    parts = []
    while rem:
        g = [rem.pop(0)]
        lo, hi = B[g[0]][:3].copy(), B[g[0]][3:].copy()
        while len(g) < max_group and rem:
            j = min(rem, key=lambda k: _volume(np.minimum(lo, B[k][:3]),
                                               np.maximum(hi, B[k][3:])))
            nlo, nhi = np.minimum(lo, B[j][:3]), np.maximum(hi, B[j][3:])
            if len(g) + 1 >= 3 and GROUP_COST * _volume(nlo, nhi) > \
                    saving * sum(V[k] for k in g + [j]):
                break
            g.append(j)
            rem.remove(j)
            lo, hi = nlo, nhi
        if len(g) >= 3 and GROUP_COST * _volume(lo, hi) <= \
                saving * sum(V[k] for k in g):
            parts.append(sorted(g))
        else:
            parts.extend([k] for k in g)
    return parts


class BatchEngine:
    """Runs batches of DoseRequests on one GPU.

    group_beams: convolve 3-4 beams of a request whose support boxes nearly
    coincide as one interleaved group (ch_convolve_groups: one 16-byte load
    per tap for all of them, 1.7x cheaper per beam than single gathers when
    the boxes coincide), chosen by group_beams(); every other beam is its
    own job of one ch_convolve_batch launch.  Off by default: it pays for
    beams that share a support box (arc control points a few degrees
    apart), not for beams at unrelated gantry angles.

    method: "fft" (the default: the convolution through zero-padded FFTs,
    ch_convolve_fft_batch, five HBM passes per beam whatever the cone
    count, beams on one plan sharing each launch) or "gather" (the direct
    stencil sum, ch_convolve_batch; the default with group_beams).

    pipeline (FFT form): TERMA runs in chunks of this many beams on the
    batch's stream and each chunk's convolution on a companion stream as
    soon as its TERMA is done, beside the next chunk's TERMA: TERMA is
    VALU-bound and the FFT passes are bound by memory latency and LDS, so
    they share the CUs (64 MLC beams: 21.7 -> 19.4 ms with chunks of 4).
    0 runs all TERMA, then all convolutions, on the batch's stream."""

    def __init__(self, engine=None, pool=None, group_beams=False,
                 host_pool=None, method=None, pipeline=4,
                 spectrum_budget=32 << 30):
        if method is None:
            method = "gather" if group_beams else "fft"
        if method not in ("gather", "fft"):
            raise ValueError("method must be 'gather' or 'fft'")
        if group_beams and method != "gather":
            raise ValueError("beam groups are a form of the gather")
        self.engine = engine or E.Engine()
        self.pool = pool or DevicePool()
        self.host_pool = host_pool or HostPool()
        self.group_beams = group_beams
        self.method = method
        # FFT form: TERMA in chunks of `pipeline` beams on the batch's
        # stream, each chunk's convolution on a companion stream behind it
        # (0: all TERMA, then all convolutions, on the one stream)
        self.pipeline = int(pipeline) if method == "fft" else 0
        # HBM the cached plans' kernel spectra may hold (least recently used
        # ones are freed first, never those of the batch being launched)
        self.spectrum_budget = int(spectrum_budget)
        self._companions = {}
        self._resident = {}
        self._default = None
        self._lock = threading.Lock()

    # -- resident phantoms -------------------------------------------------
    def densities(self, req, stream=None):
        """Device densities of a request, kept resident under req.key."""
        if req.key is not None:
            with self._lock:
                d = self._resident.get(req.key)
            if d is not None:
                return d, False
        d = self.pool.get(req.shape)
        d.upload(req.densities, stream)
        if req.key is not None:
            with self._lock:
                self._resident[req.key] = d
            return d, False
        return d, True

    def evict(self, key):
        with self._lock:
            d = self._resident.pop(key, None)
        if d is not None:
            self.pool.put(d)

    # -- one batch ---------------------------------------------------------
    def _default_stream(self):
        with self._lock:
            if self._default is None:
                self._default = Stream()
            return self._default

    def _companion(self, s):
        with self._lock:
            c = self._companions.get(s.handle)
            if c is None:
                c = self._companions[s.handle] = Stream()
            return c

    def close(self):
        """Releases the engine's default stream (its gather workspace with
        it) and the companion streams.  Batches still in flight must be
        collected first."""
        with self._lock:
            st, self._default = self._default, None
            comps, self._companions = list(self._companions.values()), {}
        for c in comps + ([st] if st is not None else []):
            c.sync()
            c.close()

    def launch(self, requests, stream=None, stage=False):
        """Queue TERMA and the convolutions of `requests` on `stream` (the
        engine's default stream if None; FFT-form chunks run on its
        companion stream, joined back before the batch's event); returns a
        BatchHandle (outputs valid once handle.done()).  stage=True also
        queues the copies of the doses into pinned host buffers behind the
        convolutions, so they run on the DMA engine while other streams
        compute and collect() hands them out without a copy of its own."""
        s = stream or self._default_stream()
        eng = self.engine
        jobs, groups, outputs, temps, pending = [], [], [], [], []
        aps = []
        s2 = self._companion(s) if self.pipeline > 0 else None
        events = []
        if self.method == "fft":
            keep = [eng.plan(r.shape, r.spacing, r.cones) for r in requests]
            eng.trim_spectra(self.spectrum_budget, keep=keep)
        # one geometry flag per request, so a ray in the source plane fails
        # only the request whose beam has it
        flags = self.pool.get((max(1, len(requests)),), np.int32)
        flags.zero(s)
        for ri, req in enumerate(requests):
            d_rho, own = self.densities(req, s)
            if own:
                temps.append(d_rho)
            plan = eng.plan(req.shape, req.spacing, req.cones)
            beams, outs = [], []
            n_jobs = req.n_jobs
            # grids beyond the FFT form's frames take the gather
            fft_here = self.method == "fft" and spectral_ok(req.shape)
            t_bufs = [self.pool.get(req.shape) for _ in range(n_jobs)]
            # support boxes only for the gather, which clips to them
            boxes = [None if fft_here else self.pool.get((8,), np.int32)
                     for _ in range(n_jobs)]
            for b, (src, blk, st) in enumerate(req.beams):
                j = b if req.per_beam else 0
                # the aperture's tables stay referenced until collect (the
                # engine's cache may drop them meanwhile)
                ap = eng.aperture(blk, s)
                aps.append(ap)
                bs = E.beam_struct(src, st, ap, d_rho,
                                   t_bufs[j], boxes[j])
                bs.accumulate = 0 if (req.per_beam or b == 0) else 1
                bs.d_status = flags.ptr + 4 * ri
                beams.append(bs)
            for j in range(n_jobs):
                outs.append(self.pool.get(req.shape))
            if s2 is not None and fft_here:
                # a chunk's convolutions start as soon as its TERMA is done,
                # beside the next chunk's TERMA
                step = self.pipeline if req.per_beam else len(beams)
                for b0 in range(0, len(beams), step):
                    E.terma(beams[b0:b0 + step], req.shape, req.origin,
                            req.spacing, stream=s, check=False)
                    ev = Event().record(s)
                    s2.wait(ev)
                    events.append(ev)
                    js = range(b0, min(b0 + step, n_jobs)) if req.per_beam \
                        else [0]
                    E.convolve_fft_batch([(plan, t_bufs[j], outs[j], None,
                                           False) for j in js], stream=s2)
            else:
                E.terma(beams, req.shape, req.origin, req.spacing, stream=s,
                        check=False)
                pending.append((plan, t_bufs, outs, boxes, fft_here))
            temps.extend(t_bufs)
            temps.extend(b for b in boxes if b is not None)
            outputs.append(outs)
        fft_jobs = []
        for plan, t_bufs, outs, boxes, fft_here in pending:
            if fft_here:
                fft_jobs.extend((plan, t_bufs[j], outs[j], None, False)
                                for j in range(len(t_bufs)))
                continue
            parts = [[j] for j in range(len(t_bufs))]
            if self.group_beams and len(t_bufs) >= 3:
                # one sync for the TERMA boxes of this request
                host = [b.download(stream=s)[:6] for b in boxes]
                parts = group_beams(host)
            for part in parts:
                if len(part) == 1:
                    j = part[0]
                    jobs.append((plan, t_bufs[j], outs[j], boxes[j], False))
                else:
                    groups.append((plan, [(t_bufs[j], outs[j], boxes[j])
                                          for j in part], False))
        if s2 is not None:
            ev = Event().record(s2)
            s.wait(ev)
            events.append(ev)
        E.convolve_fft_batch(fft_jobs, stream=s)
        E.convolve_batch(jobs, stream=s)
        E.convolve_groups(groups, stream=s)
        staged = None
        if stage:
            staged = []
            for req, outs in zip(requests, outputs):
                per = 4 * req.n_voxels
                hb = self.host_pool.get(per * len(outs))
                for i, d in enumerate(outs):
                    d.download_async(hb, i * per, stream=s)
                staged.append(hb)
        ev = Event().record(s)
        h = BatchHandle(list(requests), outputs, temps, ev, s, staged, flags)
        h.events = events          # kept alive until the batch is collected
        h.apertures = aps
        return h

    def collect(self, handle, download=True, errors="raise"):
        """Wait for a batch; returns per request a float32 array
        [n_jobs][X][Y][Z] (or the device arrays if download=False) and
        returns the batch's temporaries to the pool.  Arrays of a staged
        batch view pinned buffers, which return to the host pool when the
        arrays are dropped.  A request with a TERMA ray in the plane through
        the source perpendicular to the beam (the reference's RuntimeError,
        conehead/conehead.py:34-35) fails alone: errors="raise" raises a
        CHError naming the failed requests once the batch is cleaned up,
        errors="return" puts the CHError in that request's place."""
        handle.event.sync()
        results = []
        staged = handle.staged or [None] * len(handle.requests)
        handle.staged = None
        failed = []
        try:
            bad = np.zeros(len(handle.requests), bool)
            if handle.flags is not None:
                bad = handle.flags.download(
                    stream=handle.stream)[:len(handle.requests)] != 0
            for i, (req, outs, hb) in enumerate(
                    zip(handle.requests, handle.outputs, staged)):
                if bad[i]:
                    failed.append(i)
                    results.append(_lib_error(
                        "no intersection or line is within plane "
                        "(request %d of the batch)" % i))
                    for d in outs:
                        self.pool.put(d)
                    if hb is not None:
                        self.host_pool.put(hb)
                elif download and hb is not None:
                    results.append(self.host_pool.as_array(
                        hb, (len(outs),) + req.shape))
                    for d in outs:
                        self.pool.put(d)
                elif download:
                    arr = np.empty((len(outs),) + req.shape, np.float32)
                    for j, d in enumerate(outs):
                        d.download(arr[j], stream=handle.stream)
                        self.pool.put(d)
                    results.append(arr)
                else:
                    if hb is not None:
                        self.host_pool.put(hb)
                    results.append(outs)
        finally:
            for t in handle.temps:
                self.pool.put(t)
            if handle.flags is not None:
                self.pool.put(handle.flags)
                handle.flags = None
            handle.temps = []
            handle.events = []
            handle.apertures = []
        if failed and errors == "raise":
            raise _lib_error("no intersection or line is within plane "
                             "(requests %s of the batch)" % failed)
        return results

    def release(self, results):
        """Return device outputs of collect(download=False) to the pool."""
        for outs in results:
            for d in outs:
                self.pool.put(d)

    def run(self, requests, stream=None):
        """Launch and collect one batch."""
        return self.collect(self.launch(requests, stream))


class BatchGraph:
    """A batch of requests captured once and replayed by one graph launch.

    The requests' TERMA runs from device-resident beam sets (one
    E.BeamSet per dose grid: the beams of the grid share each launch) and
    their convolutions from one device-resident E.ConvSet (the FFT form,
    one launch per pass for all beams of a plan), all recorded in one
    graph on the graph's own stream; nothing is staged from the host when
    it is replayed.  Outputs, TERMA grids and densities are fixed device
    buffers: update() loads new requests of the same structure (grids,
    beam counts, output modes and cone sets) for the next replay -- new
    densities, beam geometry and apertures.  The result of a replay is
    bitwise the result of BatchEngine(method="fft") on the same requests.
    """

    def __init__(self, requests, engine=None):
        self.engine = engine or E.Engine()
        eng = self.engine
        self.stream = Stream()
        self.requests = list(requests)
        for r in self.requests:
            if not spectral_ok(r.shape):
                raise ValueError("grid %s exceeds the FFT form's frames" %
                                 (r.shape,))
        s = self.stream
        self.rho = [DeviceArray(r.shape) for r in self.requests]
        self.t_bufs = [[DeviceArray(r.shape) for _ in range(r.n_jobs)]
                       for r in self.requests]
        self.outs = [[DeviceArray(r.shape) for _ in range(r.n_jobs)]
                     for r in self.requests]
        jobs = []
        for ri, r in enumerate(self.requests):
            plan = eng.plan(r.shape, r.spacing, r.cones)
            jobs.extend((plan, t, d, False)
                        for t, d in zip(self.t_bufs[ri], self.outs[ri]))
        self.sets = []
        beams, keep = self._load(self.requests)
        for (key, idx), (bl, kp) in zip(self._grids(), zip(beams, keep)):
            self.sets.append(E.BeamSet(bl, *key, keep=kp))
        self.conv = E.ConvSet(jobs, stream=s)
        # one eager pass sizes the stream's FFT work grids, then capture
        self._run_eager()
        s.sync()
        self._check()
        with E.Graph.capture(s) as cap:
            for bs in self.sets:
                bs.run(s)
            self.conv.run(s)
        self.graph = cap.graph

    def _grids(self):
        """(shape, origin, spacing) -> request indices, in first-seen order."""
        out = {}
        for ri, r in enumerate(self.requests):
            key = (r.shape, tuple(float(v) for v in r.origin),
                   tuple(float(v) for v in r.spacing))
            out.setdefault(key, []).append(ri)
        return list(out.items())

    def _load(self, requests):
        """Uploads the requests' densities; ch_beam lists per grid."""
        eng, s = self.engine, self.stream
        beams, keep = [], []
        for key, idx in self._grids():
            bl, kp = [], []
            for ri in idx:
                r = requests[ri]
                self.rho[ri].upload(r.densities, s)
                for b, (src, blk, st) in enumerate(r.beams):
                    j = b if r.per_beam else 0
                    ap = eng.aperture(blk, s)
                    bs = E.beam_struct(src, st, ap, self.rho[ri],
                                       self.t_bufs[ri][j], None)
                    bs.accumulate = 0 if (r.per_beam or b == 0) else 1
                    bl.append(bs)
                    kp.append(ap)
            beams.append(bl)
            keep.append(kp)
        return beams, keep

    def _run_eager(self):
        for bs in self.sets:
            bs.run(self.stream)
        self.conv.run(self.stream)

    def _check(self):
        for bs in self.sets:
            if bs.flags(self.stream).any():
                raise _lib_error("no intersection or line is within plane")

    def update(self, requests):
        """Loads new requests of the same structure for the next launch()."""
        requests = list(requests)
        if len(requests) != len(self.requests):
            raise ValueError("update needs %d requests" % len(self.requests))
        for a, b in zip(self.requests, requests):
            if (a.shape, a.n_jobs, len(a.beams), a.per_beam, a.cones) != \
                    (b.shape, b.n_jobs, len(b.beams), b.per_beam, b.cones) \
                    or not (np.array_equal(a.origin, b.origin) and
                            np.array_equal(a.spacing, b.spacing)):
                raise ValueError("update must keep grids, beam counts, "
                                 "output modes and cone sets")
        beams, keep = self._load(requests)
        for bs, bl, kp in zip(self.sets, beams, keep):
            bs.update(bl, self.stream, keep=kp)
        self.requests = requests

    def launch(self):
        """Queues one replay; returns its event."""
        self.graph.launch(self.stream)
        return Event().record(self.stream)

    def results(self):
        """Waits for the stream; per request float32 [n_jobs][X][Y][Z].
        Raises the reference's geometry error if a replay since the last
        call had a ray in the source plane."""
        self.stream.sync()
        self._check()
        out = []
        for r, outs in zip(self.requests, self.outs):
            arr = np.empty((len(outs),) + r.shape, np.float32)
            for j, d in enumerate(outs):
                d.download(arr[j], stream=self.stream)
            out.append(arr)
        return out

    def close(self):
        self.stream.sync()
        if getattr(self, "graph", None) is not None:
            self.graph.close()
            self.graph = None
        for bs in self.sets:
            bs.close()
        self.conv.close()
        self.sets = []
