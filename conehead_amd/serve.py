"""Continuous-batching dose server on one GPU.

Requests (DoseRequest) arrive at any time through submit().  A scheduler
thread keeps up to `max_inflight` batches on the device, each on its own
stream: whenever a batch finishes (its event completes) its results are
delivered, its HBM is released, and the waiting requests that now fit the
HBM budget are admitted as the next batch -- so requests that arrive while
others run join the next free slot and their work overlaps on the device.
Admission is first-come first-served by the budget; a request that could
never fit is rejected.  Every result is bitwise identical to the same
request computed alone (both convolutions have no atomics and a fixed
order of operations per voxel), which tests/test_gpu_serve.py checks.
"""
import collections
import threading
import time
from concurrent.futures import Future

import numpy as np


class AdmissionQueue:
    """FIFO of pending requests with an HBM budget (no GPU calls)."""

    def __init__(self, budget_bytes, max_batch_jobs=64, method="gather"):
        self.budget = int(budget_bytes)
        self.in_use = 0
        self.peak = 0         # largest in_use seen
        self.deferred = 0     # admissions held back by the budget
        self.max_batch_jobs = max_batch_jobs
        self.method = method      # convolution whose work space is counted
        self._q = collections.deque()

    def push(self, item):
        """item: (request, future, t_arrival).  Rejects requests that
        exceed the whole budget."""
        req, fut, _ = item
        need = req.device_bytes(self.method)
        if need > self.budget:
            fut.set_exception(MemoryError(
                "request needs %d bytes of HBM; budget is %d" %
                (need, self.budget)))
            return False
        self._q.append(item)
        return True

    def __len__(self):
        return len(self._q)

    def take_batch(self):
        """Longest FIFO prefix that fits the free budget and job limit."""
        batch, jobs = [], 0
        while self._q:
            req = self._q[0][0]
            need = req.device_bytes(self.method)
            if self.in_use + need > self.budget:
                self.deferred += 1
                break
            if batch and jobs + req.n_jobs > self.max_batch_jobs:
                break
            batch.append(self._q.popleft())
            self.in_use += need
            self.peak = max(self.peak, self.in_use)
            jobs += req.n_jobs
        return batch

    def release(self, items):
        for req, _, _ in items:
            self.in_use -= req.device_bytes(self.method)


class DoseServer:
    """Request loop around a BatchEngine."""

    def __init__(self, batch_engine=None, hbm_budget=None, max_inflight=3,
                 max_batch_jobs=16, download=True):
        """max_inflight batches on the device at once, each of at most
        max_batch_jobs dose grids: a request is delivered when its whole
        batch is done, so small batches keep latency low; three slots of
        16 jobs measured best for the uneven workload (p50 6.0 ms at
        500 req/s, against 14.0 ms with two slots of 64)."""
        from .batch import BatchEngine
        from .device import Stream, mem_info
        self.be = batch_engine or BatchEngine()
        if hbm_budget is None:
            free, _ = mem_info()
            hbm_budget = int(free * 0.8)
        self.queue = AdmissionQueue(hbm_budget, max_batch_jobs,
                                    method=self.be.method)
        self.max_inflight = max_inflight
        # download=False delivers each result as its device dose grids
        # (the caller returns them with batch_engine.release([grids]))
        self.download = download
        self.streams = [Stream() for _ in range(max_inflight)]
        self._free_streams = list(self.streams)
        self._inflight = []
        self._cv = threading.Condition()
        self._stop = False
        self._thread = None
        self.records = []     # (t_arrival, t_done, n_beams, voxel_cone)

    # -- client side -------------------------------------------------------
    def submit(self, request):
        fut = Future()
        with self._cv:
            self.queue.push((request, fut, time.perf_counter()))
            self._cv.notify()
        return fut

    def start(self):
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()
        return self

    def stop(self):
        with self._cv:
            self._stop = True
            self._cv.notify()
        if self._thread:
            self._thread.join()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # -- scheduler ---------------------------------------------------------
    def _retire(self):
        still = []
        for handle, items, stream in self._inflight:
            if handle.done():
                try:
                    results = self.be.collect(handle,
                                              download=self.download,
                                              errors="return")
                    t = time.perf_counter()
                    for (req, fut, t0), res in zip(items, results):
                        if isinstance(res, Exception):
                            fut.set_exception(res)   # this request only
                            continue
                        self.records.append((t0, t, len(req.beams),
                                             req.voxel_cone_updates()))
                        fut.set_result(res)
                except Exception as e:  # deliver the failure
                    for _, fut, _ in items:
                        if not fut.done():
                            fut.set_exception(e)
                with self._cv:
                    self.queue.release(items)
                self._free_streams.append(stream)
            else:
                still.append((handle, items, stream))
        self._inflight = still

    def _admit(self):
        while self._free_streams:
            with self._cv:
                items = self.queue.take_batch()
            if not items:
                return
            stream = self._free_streams.pop()
            try:
                handle = self.be.launch([it[0] for it in items], stream,
                                        stage=self.download)
            except Exception as e:
                for _, fut, _ in items:
                    fut.set_exception(e)
                with self._cv:
                    self.queue.release(items)
                self._free_streams.append(stream)
                continue
            self._inflight.append((handle, items, stream))

    def _loop(self):
        while True:
            with self._cv:
                if self._stop and not self._inflight and not len(self.queue):
                    return
                if not self._inflight and not len(self.queue):
                    self._cv.wait(0.05)
                    continue
            self._retire()
            self._admit()
            if self._inflight:
                # wait for the oldest batch without spinning the host
                oldest = self._inflight[0][0]
                if not oldest.done():
                    time.sleep(0.0002)

    # -- statistics --------------------------------------------------------
    def stats(self):
        if not self.records:
            return {}
        r = np.array([(t1 - t0, nb, vc) for t0, t1, nb, vc in self.records])
        t_first = min(t0 for t0, _, _, _ in self.records)
        t_last = max(t1 for _, t1, _, _ in self.records)
        span = max(t_last - t_first, 1e-9)
        return {
            "requests": len(self.records),
            "beams": int(r[:, 1].sum()),
            "latency_ms_p50": float(np.percentile(r[:, 0], 50) * 1e3),
            "latency_ms_p99": float(np.percentile(r[:, 0], 99) * 1e3),
            "beams_per_s": float(r[:, 1].sum() / span),
            "voxel_cone_updates_per_s": float(r[:, 2].sum() / span),
            "span_s": span,
        }


def poisson_arrivals(rate_per_s, n, seed=99):
    """n arrival offsets (s) of a Poisson process of the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))


def replay(server, requests, arrivals, out=None):
    """Submit requests at the given offsets (s from now); returns futures.
    With `out` (a list), each future is appended to it as it is submitted,
    so another thread can take results while the stream is still arriving
    (see consume)."""
    t0 = time.perf_counter()
    futs = [] if out is None else out
    for req, t in zip(requests, arrivals):
        dt = t0 + t - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        futs.append(server.submit(req))
    return futs


def consume(server, requests, arrivals, take=None, timeout=300):
    """Replays the stream (in a thread) while this thread takes each result
    as soon as it is done, in order, hands it to take(i, result) if given
    and drops it: a consumer that does not hold every result returns the
    pinned buffers of delivered doses to their pool as it goes, as a
    serving client would."""
    out = []
    th = threading.Thread(target=replay, args=(server, requests, arrivals,
                                               out), daemon=True)
    th.start()
    for i in range(len(requests)):
        while len(out) <= i:
            time.sleep(0.0002)
        r = out[i].result(timeout=timeout)
        out[i] = None
        if take is not None:
            take(i, r)
        del r
    th.join()
