"""Host logic of the batched path and the request loop (no GPU calls):
request accounting, HBM-budget admission and Poisson arrivals."""
from concurrent.futures import Future

import numpy as np
import pytest

from conehead_amd import workloads as W
from conehead_amd.batch import (DoseRequest, group_beams, spectral_frame,
                                 spectral_ok)
from conehead_amd.serve import AdmissionQueue, poisson_arrivals


class FakeRequest:
    def __init__(self, nbytes, n_jobs=1):
        self.nbytes = nbytes
        self.n_jobs = n_jobs

    def device_bytes(self, method="gather"):
        return self.nbytes


def item(nbytes, n_jobs=1):
    return (FakeRequest(nbytes, n_jobs), Future(), 0.0)


def test_request_accounting():
    c = W.uneven(n_phantoms=1, seed=11, max_beams=3, xy=(16, 36),
                 z=(14, 40), n_theta=6, n_phis=(4,))[0]
    r = DoseRequest.from_case(c)
    X, Y, Z = c.shape
    assert r.n_jobs == len(c.beams)
    assert DoseRequest.from_case(c, per_beam=False).n_jobs == 1
    padded = X * (Y + 64) * (Z + 160)
    slots = -(-r.n_jobs // 4) * 4
    assert r.device_bytes() == 4 * (X * Y * Z + r.n_jobs *
                                    (2 * X * Y * Z + 8) + slots * padded)
    assert r.voxel_cone_updates() == X * Y * Z * c.cones.n_cones * \
        len(c.beams)
    # FFT form: one complex X x Fy x Kp work grid per job instead
    frame, kp = spectral_frame(c.shape)
    assert all(f >= 2 * n - 1 for f, n in zip(frame, c.shape))
    assert r.device_bytes("fft") == 4 * (X * Y * Z + r.n_jobs *
                                         (2 * X * Y * Z + 8)) + \
        8 * r.n_jobs * X * frame[1] * kp


def test_spectral_frame():
    """Smallest 2,3,5-smooth lengths >= 2n - 1 and the 16-aligned pitch of
    the Fz / 2 + 1 half-spectrum bins."""
    assert spectral_frame((128, 128, 128)) == ((256, 256, 256), 144)
    assert spectral_frame((200, 200, 160)) == ((400, 400, 320), 176)
    assert spectral_frame((24, 26, 22)) == ((48, 54, 45), 32)
    assert spectral_frame((20, 13, 1)) == ((40, 25, 1), 16)
    assert spectral_ok((512, 64, 64)) and not spectral_ok((513, 64, 64))


def test_admission_fifo_and_budget():
    q = AdmissionQueue(100, max_batch_jobs=4)
    items = [item(30), item(30), item(30), item(30)]
    for it in items:
        assert q.push(it)
    b1 = q.take_batch()
    assert [x is y for x, y in zip(b1, items[:3])] == [True] * 3
    assert q.in_use == 90 and len(q) == 1
    assert q.take_batch() == []           # 90 + 30 > 100: waits
    q.release(b1)
    assert q.in_use == 0
    b2 = q.take_batch()
    assert b2 == [items[3]] and q.in_use == 30


def test_admission_rejects_oversized():
    q = AdmissionQueue(100)
    it = item(101)
    assert not q.push(it)
    with pytest.raises(MemoryError):
        it[1].result(timeout=0)
    assert len(q) == 0


def test_admission_job_limit_and_order():
    q = AdmissionQueue(10 ** 9, max_batch_jobs=4)
    a, b, c = item(1, 3), item(1, 2), item(1, 1)
    for it in (a, b, c):
        q.push(it)
    assert q.take_batch() == [a]          # a + b would be 5 jobs
    assert q.take_batch() == [b, c]       # FIFO: c never overtakes b
    big = item(1, 9)                      # alone it may exceed the limit
    q.push(big)
    assert q.take_batch() == [big]


def test_poisson_arrivals():
    t = poisson_arrivals(200.0, 20000, seed=3)
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 200.0) < 0.05 / 200.0
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05   # exponential CV=1


def test_group_beams_cost_model():
    same = [(0, 0, 0, 9, 9, 9)] * 4
    assert group_beams(same) == [[0, 1, 2, 3]]
    # two far-apart clusters of three: each cluster groups, nothing mixes
    a = [(0, 0, 0, 9, 9, 9), (0, 0, 1, 9, 9, 9), (1, 0, 0, 9, 9, 9)]
    b = [(50, 50, 50, 59, 59, 59)] * 3
    parts = group_beams(a + b)
    assert sorted(map(sorted, parts)) == [[0, 1, 2], [3, 4, 5]]
    # disjoint boxes stay single; two beams never form a group
    far = [(0, 0, 0, 4, 4, 4), (20, 20, 20, 24, 24, 24),
           (40, 40, 40, 44, 44, 44)]
    assert sorted(group_beams(far)) == [[0], [1], [2]]
    assert sorted(group_beams(same[:2])) == [[0], [1]]
    parts = group_beams([(0, 0, 0, 9, 9, 9)] * 6)
    assert sorted(len(p) for p in parts) in ([1, 1, 4], [3, 3])
    assert sorted(i for p in parts for i in p) == list(range(6))


def test_pinned_view_ownership():
    """A HostPool array owns its buffer: views keep it out of the pool, and
    it returns when the last of them is dropped."""
    import gc
    from conehead_amd.device import HostPool

    class Buf:                  # host memory standing in for a pinned one
        def __init__(self, n):
            self.mem = np.zeros(n // 4, np.float32)
            self.ptr = self.mem.ctypes.data
            self.nbytes = n

    pool = HostPool()
    b = Buf(64)
    a = pool.as_array(b, (2, 8))
    a[1, 3] = 5.0
    assert b.mem[11] == 5.0
    v = a[1]
    del a
    gc.collect()
    assert pool._bytes == 0
    del v
    gc.collect()
    assert pool._bytes == 64 and pool.get(64) is b
    with pytest.raises(ValueError):
        pool.as_array(b, (4, 8))
