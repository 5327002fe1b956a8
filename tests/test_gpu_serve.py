"""The continuous-batching request loop: requests arriving while others run
are admitted into free stream slots; every result is bitwise the result of
the same request computed alone."""
import numpy as np
import pytest

from conehead_amd import workloads as W
from conehead_amd.batch import BatchEngine, DoseRequest
from conehead_amd.device import Stream
from conehead_amd.serve import DoseServer, poisson_arrivals, replay

pytestmark = pytest.mark.gpu


def test_served_equals_alone(engine):
    cases = W.uneven(n_phantoms=8, seed=21, max_beams=2, xy=(16, 32),
                     z=(14, 30), spacing=(0.4, 0.8), n_theta=6,
                     n_phis=(4, 6))
    be = BatchEngine(engine)
    reqs = [DoseRequest.from_case(c) for c in cases]
    s = Stream()
    alone = [be.run([r], s)[0] for r in reqs]
    with DoseServer(be, hbm_budget=1 << 30, max_inflight=2) as srv:
        futs = replay(srv, reqs, poisson_arrivals(400.0, len(reqs), seed=5))
        got = [f.result(timeout=60) for f in futs]
    for i, (a, b) in enumerate(zip(got, alone)):
        assert np.array_equal(a, b), i
    st = srv.stats()
    assert st["requests"] == len(reqs)
    assert st["latency_ms_p99"] >= st["latency_ms_p50"] > 0


def test_budget_serialises_and_rejects(engine):
    cases = W.uneven(n_phantoms=3, seed=22, max_beams=1, xy=(16, 24),
                     z=(14, 20), n_theta=6, n_phis=(4,))
    be = BatchEngine(engine)
    reqs = [DoseRequest.from_case(c) for c in cases]
    one = max(r.device_bytes() for r in reqs)
    with DoseServer(be, hbm_budget=one, max_inflight=2) as srv:
        futs = [srv.submit(r) for r in reqs]
        big = DoseRequest.from_case(cases[0])
        big.device_bytes = lambda method="gather": one + 1
        rej = srv.submit(big)
        got = [f.result(timeout=60) for f in futs]
    with pytest.raises(MemoryError):
        rej.result(timeout=0)
    s = Stream()
    for r, g in zip(reqs, got):
        assert np.array_equal(g, be.run([r], s)[0])
