"""HIP-graph capture of the single-beam pipeline (zero, TERMA, then the
gather, or the FFT form's five passes accumulating onto the zeroed dose).

Every replay is checked bit for bit against the eager result, with eager
work and host synchronisation between replays; a replay reads the buffers'
current contents (new densities give the new dose); a capture on a stream
whose scratch does not exist yet fails with an error and leaves the stream
usable."""
import numpy as np
import pytest

from conehead_amd import _lib
from conehead_amd import engine as E
from conehead_amd import workloads as W
from conehead_amd.device import DeviceArray, Stream

pytestmark = pytest.mark.gpu


def setup(engine, s):
    case = W.hetero(n=40, spacing=0.6, n_theta=8, n_phi=6)
    shape = case.shape
    src, blk, st = case.beams[0]
    rho = DeviceArray.from_host(case.densities, stream=s)
    t = DeviceArray(shape)
    d = DeviceArray(shape)
    box = DeviceArray((8,), np.int32)
    beam = E.beam_struct(src, st, engine.aperture(blk, s), rho, t, box)
    plan = engine.plan(shape, case.spacing, case.cones, tile=(1, 4, 16))
    return case, rho, t, d, box, beam, plan


def run(case, t, d, box, beam, plan, s, method="gather"):
    d.zero(s)
    E.terma([beam], case.shape, case.origin, case.spacing, stream=s,
            check=False)
    if method == "fft":
        E.convolve_fft(plan, t, d, accumulate=True, stream=s)
    else:
        E.convolve(plan, t, d, box, stream=s)


@pytest.mark.parametrize("method", ["gather", "fft"])
def test_graph_replays_match_eager(engine, method):
    s = Stream()
    case, rho, t, d, box, beam, plan = setup(engine, s)
    run(case, t, d, box, beam, plan, s, method)   # eager, warms scratch
    ref = d.download(stream=s)
    assert ref.max() > 0
    with E.Graph.capture(s) as cap:
        run(case, t, d, box, beam, plan, s, method)
    g = cap.graph
    for _ in range(4):
        d.zero(s)                                  # eager work in between
        t.zero(s)
        s.sync()
        g.launch(s)
        assert np.array_equal(d.download(stream=s), ref)
    # the graph holds pointers, not data: new densities, new dose
    rho.upload(case.densities * 1.1, s)
    g.launch(s)
    replay = d.download(stream=s)
    run(case, t, d, box, beam, plan, s, method)
    assert np.array_equal(replay, d.download(stream=s))
    assert not np.array_equal(replay, ref)
    g.close()


def test_capture_refuses_unwarmed_stream(engine):
    s0 = Stream()
    case, rho, t, d, box, beam, plan = setup(engine, s0)
    s = Stream()                                   # no scratch yet
    with pytest.raises(_lib.CHError):
        with E.Graph.capture(s):
            run(case, t, d, box, beam, plan, s)
    run(case, t, d, box, beam, plan, s)            # the stream still works
    s.sync()
    assert d.download(stream=s).max() > 0


def test_graph_holds_its_plan_and_stream(engine):
    """A live graph pins the plans it reads and the capture stream's
    scratch: trim_spectra skips the plan, releasing its spectrum is refused,
    destroying the plan is deferred to the graph's destruction, a call that
    would grow the stream's work grid is refused and so is destroying the
    stream; every replay still gives the eager bits.  Once the graph is
    gone, all of that works again."""
    s = Stream()
    case, rho, t, d, box, beam, plan = setup(engine, s)
    # a private plan (not the engine's cached one) to destroy under the graph
    own = E.Plan(case.shape, case.spacing, case.cones)
    run(case, t, d, box, beam, own, s, "fft")
    ref = d.download(stream=s)
    with E.Graph.capture(s) as cap:
        run(case, t, d, box, beam, own, s, "fft")
    g = cap.graph
    assert own.graphs() == 1
    with pytest.raises(_lib.CHError, match="live graph"):
        own.release_spectrum()
    # the engine's LRU trim leaves a pinned plan's spectrum alone
    engine._plans["pinned-test"] = own
    try:
        engine.trim_spectra(0)
        assert own.spectrum_bytes > 0
    finally:
        del engine._plans["pinned-test"]
    # a bigger grid on the capture stream would regrow its work grid
    big = W.hetero(n=64, spacing=0.5, n_theta=4, n_phi=4)
    pb = engine.plan(big.shape, big.spacing, big.cones)
    tb, db = DeviceArray(big.shape, zero=True), DeviceArray(big.shape)
    with pytest.raises(_lib.CHError, match="live graph"):
        E.convolve_fft(pb, tb, db, stream=s)
    s2 = Stream()
    E.convolve_fft(pb, tb, db, stream=s2)          # another stream is fine
    s2.sync()
    assert _lib.lib().ch_stream_destroy(s.handle) == _lib.CH_ERR_ARG
    handle = own.handle
    own.close()                                    # deferred: graph lives
    for _ in range(3):
        d.zero(s)
        s.sync()
        g.launch(s)
        assert np.array_equal(d.download(stream=s), ref)
    g.close()                                      # frees the plan as well
    E.convolve_fft(pb, tb, db, stream=s)           # the stream may grow now
    s.sync()
    assert handle is not None
    p2 = E.Plan(case.shape, case.spacing, case.cones)
    p2.spectrum(s)
    run(case, t, d, box, beam, p2, s, "fft")
    assert p2.graphs() == 0
    p2.release_spectrum()
    assert np.array_equal(d.download(stream=s), ref)
