"""Autotuner for the convolutions and the batch pipeline.

Gather tiles.  Two gathers compete.  In the direct one a wavefront owns a
1 x TY x TZ output tile (TY * TZ = 64 voxels) and walks the union of its
voxels' clipped tap ranges with one global load per tap, so taller tiles
share tap loads across more y-rows but widen that union.  In the staged one
(4 x 16 x 16) a workgroup stages the TERMA box of each chunk of taps in LDS
and reads the taps from there, trading global loads for LDS traffic and a
wider union.  tune() times every candidate on the real plan and a synthetic
TERMA of the given support.

FFT form.  Each of the five passes (z forward, y forward, x convolution,
y inverse, z inverse) transforms lines of its axis, W lines per workgroup:
more lines share each twiddle and kernel-spectrum read and fill the 256
threads on short frames, fewer lines mean smaller LDS buffers and more
workgroups in flight.  tune_fft() picks W per pass by coordinate descent
(each pass in turn takes its fastest W with the others fixed), timed on the
batched passes of n beams.  The choice moves no bits: every W computes the
same butterflies on every line.

Batch pipeline.  tune_pipeline() picks the BatchEngine's TERMA chunk (the
beams whose convolutions start as soon as their TERMA is done, beside the
next chunk's TERMA) for a set of requests.

Winners are cached per key in memory and in a JSON file.
"""
import json
import os
import threading

import numpy as np

from . import _lib
from .device import DeviceArray, Event, Stream

CANDIDATES = [(1, 1, 64), (1, 2, 32), (1, 4, 16), (1, 8, 8), (4, 16, 16)]

CACHE_PATH = os.environ.get(
    "CONEHEAD_AMD_TUNE_CACHE",
    os.path.join(os.path.expanduser("~"), ".cache", "conehead_amd",
                 "tiles.json"))

_lock = threading.Lock()
_memo = {}


def default_tile(plan):
    """Tile before (or without) tuning: the last tuned tile of this grid
    shape and cone count, else 1 x 2 x 32."""
    key = _shape_key(plan)
    if key in _memo:
        return _memo[key]
    return (1, 2, 32)


def _shape_key(plan):
    return "%dx%dx%d/%d" % (plan.shape + (plan.n_cones,))


def _load_cache():
    try:
        with open(CACHE_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(d):
    try:
        os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
        tmp = CACHE_PATH + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1)
        os.replace(tmp, CACHE_PATH)
    except OSError:
        pass


def tune(plan, box=None, candidates=CANDIDATES, reps=3, stream=None,
         use_cache=True, terma=None, batch=False):
    """Time every candidate tile on `plan` and keep the fastest.

    box: (lo0, lo1, lo2, hi0, hi1, hi2) support of the TERMA to tune for
    (default: the whole grid); terma: a device TERMA grid to tune on (by
    default the box is filled with random values); batch: time the batched
    gather (ch_convolve_batch, which uses its own workgroup size) instead of
    the single-grid one.  Returns (best_tile, {tile: ms})."""
    key = _shape_key(plan)
    if box is not None:
        key += "/" + ",".join(str(int(b)) for b in box)
    if batch:
        key += "/batch"
    if use_cache:
        with _lock:
            cached = _load_cache().get(key)
        if cached:
            tile = tuple(cached["tile"])
            plan.set_tile(*tile)
            _memo[_shape_key(plan)] = tile
            return tile, {tuple(map(int, k.split("x"))): v
                          for k, v in cached["ms"].items()}
    shape = plan.shape
    b = box or (0, 0, 0, shape[0] - 1, shape[1] - 1, shape[2] - 1)
    s = stream or Stream()
    if terma is not None:
        d_t = terma
    else:
        rng = np.random.default_rng(0)
        t = np.zeros(shape, np.float32)
        t[b[0]:b[3] + 1, b[1]:b[4] + 1, b[2]:b[5] + 1] = rng.random(
            (b[3] - b[0] + 1, b[4] - b[1] + 1, b[5] - b[2] + 1),
            dtype=np.float32)
        d_t = DeviceArray.from_host(t)
    d_d = DeviceArray(shape, np.float32)
    d_box = DeviceArray.from_host(np.asarray(b, np.int32))
    from .engine import convolve, convolve_batch

    def run():
        if batch:
            convolve_batch([(plan, d_t, d_d, d_box, False)], stream=s)
        else:
            convolve(plan, d_t, d_d, d_box, stream=s)

    e0, e1 = Event(), Event()
    times = {}
    for tile in candidates:
        plan.set_tile(*tile)
        run()  # warm-up
        e0.record(s)
        for _ in range(reps):
            run()
        e1.record(s)
        e1.sync()
        times[tile] = e0.elapsed_ms(e1) / reps
    best = min(times, key=times.get)
    plan.set_tile(*best)
    _memo[_shape_key(plan)] = best
    if use_cache:
        with _lock:
            d = _load_cache()
            d[key] = {"tile": list(best),
                      "ms": {"x".join(map(str, k)): v
                             for k, v in times.items()}}
            _save_cache(d)
    return best, times


# -- FFT form ---------------------------------------------------------------
FFT_PASSES = ("z_fwd", "y_fwd", "x_conv", "y_inv", "z_inv")
timings = 0          # candidates timed by tune_fft / tune_pipeline so far


def fft_key(plan, n_beams):
    # "fft2": the choices reach 16 lines wherever 16 frames fit 4096 points
    return "fft2/%dx%dx%d/%d/%d" % (plan.shape + (plan.n_cones, n_beams))


def tune_fft(plan, n_beams=1, reps=3, stream=None, use_cache=True):
    """Lines per workgroup of the FFT form's five passes for `plan`, timed
    on the batched passes (convolve_fft_batch) of n_beams random TERMA
    grids.  Sets the winner on the plan; returns (lines, {lines: ms})."""
    global timings
    from .engine import convolve_fft_batch
    key = fft_key(plan, n_beams)
    if key in _memo:
        plan.set_fft_lines(_memo[key])
        return _memo[key], {}
    if use_cache:
        with _lock:
            cached = _load_cache().get(key)
        if cached:
            lines = tuple(cached["lines"])
            plan.set_fft_lines(lines)
            _memo[key] = lines
            return lines, {tuple(map(int, k.split("x"))): v
                           for k, v in cached["ms"].items()}
    s = stream or Stream()
    plan.spectrum(s)
    rng = np.random.default_rng(0)
    ts = [DeviceArray.from_host(rng.random(plan.shape, dtype=np.float32))
          for _ in range(n_beams)]
    ds = [DeviceArray(plan.shape) for _ in range(n_beams)]
    jobs = [(plan, t, d, None, False) for t, d in zip(ts, ds)]
    plan.set_fft_lines((16,) * 5)
    most = plan.fft_lines()              # the most each pass takes
    plan.set_fft_lines((0,) * 5)
    top = plan.fft_lines()               # the defaults
    e0, e1 = Event(), Event()
    times = {}

    def timed(lines):
        global timings
        lines = tuple(lines)
        if lines not in times:
            plan.set_fft_lines(lines)
            convolve_fft_batch(jobs, stream=s)       # warm-up
            e0.record(s)
            for _ in range(reps):
                convolve_fft_batch(jobs, stream=s)
            e1.record(s)
            e1.sync()
            times[lines] = e0.elapsed_ms(e1) / reps
            timings += 1
        return times[lines]

    best = list(top)
    timed(best)
    for k in range(5):
        for w in (2, 4, 8, 16):
            if w > most[k] or w == best[k]:
                continue
            trial = list(best)
            trial[k] = w
            if timed(trial) < timed(best):
                best = trial
    best = tuple(best)
    plan.set_fft_lines(best)
    _memo[key] = best
    if use_cache:
        with _lock:
            d = _load_cache()
            d[key] = {"lines": list(best),
                      "ms": {"x".join(map(str, k)): v
                             for k, v in times.items()}}
            _save_cache(d)
    return best, times


def apply_fft(plan, n_beams=1):
    """Sets the tuned lines of a plan if this process or the cache file has
    them (no timing); returns them or None."""
    key = fft_key(plan, n_beams)
    lines = _memo.get(key)
    if lines is None:
        with _lock:
            cached = _load_cache().get(key)
        if cached:
            lines = _memo[key] = tuple(cached["lines"])
    if lines is not None:
        plan.set_fft_lines(lines)
    return lines


# -- batch pipeline ---------------------------------------------------------
PIPELINE_CANDIDATES = (0, 4, 8, 16, 32, 64)


def tune_pipeline(batch_engine, requests, candidates=PIPELINE_CANDIDATES,
                  reps=3, stream=None):
    """TERMA chunk of an FFT-form BatchEngine for these requests: times a
    launch and collect of the batch for every candidate (0: all TERMA, then
    all convolutions), sets the fastest on the engine; returns (chunk,
    {chunk: ms})."""
    global timings
    import time
    be = batch_engine
    if be.method != "fft":
        raise ValueError("the pipeline chunk is a setting of the FFT form")
    s = stream or Stream()
    times = {}
    for c in candidates:
        be.pipeline = int(c)
        be.release(be.collect(be.launch(requests, s), download=False))
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            be.release(be.collect(be.launch(requests, s), download=False))
            t.append((time.perf_counter() - t0) * 1e3)
        times[int(c)] = float(np.median(t))
        timings += 1
    best = min(times, key=times.get)
    be.pipeline = best
    return best, times
