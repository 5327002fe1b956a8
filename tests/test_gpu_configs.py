"""The batched and served configurations at their full sizes.

  * Batched control points: the 128^3 heterogeneous phantom, 64 MLC beams
    at gantry angles U[0, 360) (seed 2024), 96 cones, as one BatchEngine
    batch with each convolution.  Beams 0 and 63 against the FP64 oracle
    chain (oracle TERMA with the beam's leaf raster -> oracle convolve_c),
    within 1e-4 of Dmax; the batch equals each beam computed alone, bit for
    bit.
  * Serving: every request of the uneven workload (32 phantoms from 64^3 to
    200 x 200 x 160, 2-5 mm voxels, 1-9 beams, 48-144 cones) fed to the
    DoseServer as a Poisson stream under an HBM budget of 1.5 times the
    largest request, so admissions wait for memory: served == alone bit for
    bit for every request, the budget is never exceeded, and the smallest
    phantom's first beam is checked against the oracle chain."""
import numpy as np
import pytest

from conehead_amd import engine as E
from conehead_amd import workloads as W
from conehead_amd.batch import BatchEngine, DoseRequest
from conehead_amd.device import Stream
from conehead_amd.serve import DoseServer, poisson_arrivals, replay

pytestmark = pytest.mark.gpu

TOL = 1e-4


def oracle_chain(oracle, case, beam, n_threads=16):
    src, blk, st = beam
    e, w, mu = E.spectrum(st)
    rb = oracle.make_beam(src.position, src.v_y, src.transform, src.sad, st,
                          e, w, mu)
    t = oracle.terma(rb, case.shape, case.origin, case.spacing,
                     case.densities, blk.block_values,
                     n_threads=n_threads)["terma"]
    ref = np.zeros(case.shape)
    oracle.convolve_dirs(t, ref, case.spacing, case.cones.dirs,
                         case.cones.rows, case.cones.cum,
                         n_threads=n_threads)
    return ref


@pytest.fixture(scope="module")
def c3(oracle):
    case = W.batched()
    assert len(case.beams) == 64 and case.shape == (128, 128, 128)
    assert case.cones.n_cones == 96
    refs = {b: oracle_chain(oracle, case, case.beams[b]) for b in (0, 63)}
    return case, refs


@pytest.mark.parametrize("method", ["fft", "gather"])
def test_c3_batched_control_points(engine, c3, method):
    case, refs = c3
    be = BatchEngine(engine, method=method)
    s = Stream()
    dose = be.run([DoseRequest.from_case(case)], s)[0]
    assert dose.shape == (64,) + case.shape
    assert np.isfinite(dose).all()
    for b, ref in refs.items():
        err = np.abs(dose[b] - ref).max() / ref.max()
        assert err < TOL, (method, b, err)
        alone = be.run([DoseRequest(case.densities, case.origin,
                                    case.spacing, [case.beams[b]],
                                    case.cones)], s)[0]
        assert np.array_equal(alone[0], dose[b]), (method, b)
    # every beam delivers dose (no MLC aperture is closed)
    assert (dose.reshape(64, -1).max(axis=1) > 0).all()


def test_c5_serving_uneven_distribution(engine, oracle):
    cases = W.uneven()
    assert len(cases) == 32
    reqs = [DoseRequest.from_case(c) for c in cases]
    be = BatchEngine(engine)
    s = Stream()
    alone = [be.run([r], s)[0] for r in reqs]
    need = [r.device_bytes(be.method) for r in reqs]
    budget = int(1.5 * max(need))
    with DoseServer(be, hbm_budget=budget, max_inflight=2) as srv:
        futs = replay(srv, reqs, poisson_arrivals(200.0, len(reqs), seed=99))
        got = [f.result(timeout=120) for f in futs]
    for i, (a, b) in enumerate(zip(got, alone)):
        assert a.shape == (len(cases[i].beams),) + cases[i].shape
        assert np.array_equal(a, b), i
    st = srv.stats()
    assert st["requests"] == 32
    assert st["beams"] == sum(len(c.beams) for c in cases)
    assert st["latency_ms_p99"] >= st["latency_ms_p50"] > 0
    assert 0 < srv.queue.peak <= budget and srv.queue.in_use == 0
    # all 32 at once: together they need 6x the budget, so admissions wait
    # for memory to free up, and the results do not change
    with DoseServer(be, hbm_budget=budget, max_inflight=2) as srv:
        futs = replay(srv, reqs, np.zeros(len(reqs)))
        burst = [f.result(timeout=120) for f in futs]
    for i, (a, b) in enumerate(zip(burst, alone)):
        assert np.array_equal(a, b), i
    q = srv.queue
    assert sum(need) > 4 * budget
    assert 0 < q.peak <= budget and q.in_use == 0
    assert q.deferred > 0
    # the smallest phantom's first served beam against the oracle chain
    i = int(np.argmin([c.n_voxels for c in cases]))
    ref = oracle_chain(oracle, cases[i], cases[i].beams[0])
    err = np.abs(got[i][0] - ref).max() / ref.max()
    assert err < TOL, err
