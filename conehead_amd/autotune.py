"""Tile autotuner for the collapsed-cone gather.

Two gathers compete.  In the direct one a wavefront owns a 1 x TY x TZ
output tile (TY * TZ = 64 voxels) and walks the union of its voxels'
clipped tap ranges with one global load per tap, so taller tiles share tap
loads across more y-rows but widen that union.  In the staged one
(4 x 16 x 16) a workgroup stages the TERMA box of each chunk of taps in LDS
and reads the taps from there, trading global loads for LDS traffic and a
wider union.  The best
choice depends on the grid, the cone set and the TERMA support, so it is
measured: every candidate is timed with HIP events on the real plan and a
synthetic TERMA of the given support, and the winner is cached per plan
key in memory and in a JSON file.
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
