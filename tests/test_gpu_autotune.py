"""The autotuner: every choice it can make gives the same dose bits (FFT
lines per workgroup, batch pipeline chunk) or the same dose to FP32
rounding (gather tiles, whose tap unions differ), the winner is set on the
plan or engine, and a cached choice is reused without timing again."""
import numpy as np
import pytest

from conehead_amd import autotune
from conehead_amd import engine as E
from conehead_amd import workloads as W
from conehead_amd.batch import BatchEngine, DoseRequest
from conehead_amd.device import DeviceArray, Stream

pytestmark = pytest.mark.gpu


@pytest.fixture
def cache(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE_PATH", str(tmp_path / "t.json"))
    monkeypatch.setattr(autotune, "_memo", {})
    return tmp_path / "t.json"


def fft_dose(plan, t, s):
    d = DeviceArray(plan.shape)
    E.convolve_fft_batch([(plan, t, d, None, False)], stream=s)
    return d.download(stream=s)


def test_fft_lines_same_bits_and_cache(engine, cache):
    case = W.hetero(n=44, spacing=0.6, n_theta=6, n_phi=4)
    plan = E.Plan(case.shape, case.spacing, case.cones)
    s = Stream()
    plan.spectrum(s)
    rng = np.random.default_rng(1)
    t = DeviceArray.from_host(rng.random(case.shape, dtype=np.float32))
    plan.set_fft_lines((16,) * 5)
    most = plan.fft_lines()
    plan.set_fft_lines((0,) * 5)
    ref = fft_dose(plan, t, s)
    # every lines choice of every pass, one pass at a time
    for k in range(5):
        for w in (2, 4, 8, 16):
            if w > most[k]:
                continue
            lines = [0] * 5
            lines[k] = w
            plan.set_fft_lines(lines)
            assert plan.fft_lines()[k] == w
            assert np.array_equal(fft_dose(plan, t, s), ref), (k, w)
    best, times = autotune.tune_fft(plan, n_beams=2, reps=1, stream=s)
    assert plan.fft_lines() == best and best in times
    assert all(w in (2, 4, 8, 16) for w in best)
    assert np.array_equal(fft_dose(plan, t, s), ref)
    assert cache.exists()
    # cached: a new plan of the same key takes the choice without timing
    n0 = autotune.timings
    autotune._memo.clear()
    p2 = E.Plan(case.shape, case.spacing, case.cones)
    again, t2 = autotune.tune_fft(p2, n_beams=2, stream=s)
    assert again == best and autotune.timings == n0
    assert p2.fft_lines() == best and set(t2) == set(times)
    # a value whose frames do not fit falls back to the default
    plan.set_fft_lines((16, 16, 16, 16, 16))
    assert all(a <= b for a, b in zip(plan.fft_lines(), (16,) * 5))
    with pytest.raises(Exception):
        plan.set_fft_lines((3, 0, 0, 0, 0))


def test_gather_tiles_agree(engine, cache):
    case = W.hetero(n=36, spacing=0.6, n_theta=6, n_phi=4)
    plan = E.Plan(case.shape, case.spacing, case.cones)
    s = Stream()
    rng = np.random.default_rng(2)
    tt = np.zeros(case.shape, np.float32)
    tt[4:30, 2:33, 5:31] = rng.random((26, 31, 26), dtype=np.float32)
    t = DeviceArray.from_host(tt)
    box = DeviceArray.from_host(np.array([4, 2, 5, 29, 32, 30, 0, 0],
                                         np.int32))
    outs = {}
    for tile in autotune.CANDIDATES:
        plan.set_tile(*tile)
        d = DeviceArray(case.shape)
        E.convolve(plan, t, d, box, stream=s)
        outs[tile] = d.download(stream=s)
    ref = outs[autotune.CANDIDATES[0]]
    for tile, o in outs.items():
        assert np.abs(o - ref).max() <= 1e-6 * ref.max(), tile
    best, times = autotune.tune(plan, box=(4, 2, 5, 29, 32, 30), reps=1,
                                stream=s)
    assert plan.tile() == best and set(times) == set(autotune.CANDIDATES)
    n = len(autotune._load_cache())
    best2, _ = autotune.tune(plan, box=(4, 2, 5, 29, 32, 30), stream=s)
    assert best2 == best and len(autotune._load_cache()) == n


def test_pipeline_chunk_same_bits(engine, cache):
    cases = W.uneven(n_phantoms=3, seed=5, max_beams=5, xy=(20, 36),
                     z=(16, 36), spacing=(0.4, 0.8), n_theta=6, n_phis=(4,))
    reqs = [DoseRequest.from_case(c) for c in cases]
    be = BatchEngine(engine, method="fft", pipeline=0)
    s = Stream()
    ref = be.run(reqs, s)
    chunk, times = autotune.tune_pipeline(be, reqs, candidates=(0, 1, 2, 4),
                                          reps=1, stream=s)
    assert be.pipeline == chunk and set(times) == {0, 1, 2, 4}
    for c in (0, 1, 2, 4):
        be.pipeline = c
        for a, b in zip(be.run(reqs, s), ref):
            assert np.array_equal(a, b), c
