"""The gather and the FFT form at the benchmark's full size (128^3 at
2.5 mm, 12 x 12 cones) through properties that hold at any size, where the
FP64 oracle would take minutes:
  - scaling T by 2 scales every product and sum exactly (FP32), so the dose
    doubles bit for bit;
  - linearity: dose(T1 + T2) = dose(T1) + dose(T2) within FP32 rounding;
  - a batch holding the beam twice gives the single launch's result twice,
    bit for bit;
  - the x-slab support boxes change nothing: a caller's box or none at all
    gives the same dose within FP32 rounding.
The single-beam dose at this size is checked against the FP64 oracle by
bench.py (error_vs_fp64_oracle)."""
import numpy as np
import pytest

from conehead_amd import engine as E
from conehead_amd.device import DeviceArray
from conehead_amd.workloads import hetero

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case_terma(engine):
    case = hetero()
    src, blk, st = case.beams[0]
    d_rho = DeviceArray.from_host(case.densities)
    d_t = DeviceArray(case.shape, np.float32)
    d_box = DeviceArray((6,), np.int32)
    beam = E.beam_struct(src, st, engine.aperture(blk), d_rho, d_t, d_box)
    E.terma([beam], case.shape, case.origin, case.spacing)
    return case, d_t.download()


def gather(engine, case, T, box=True):
    plan = engine.plan(case.shape, case.spacing, case.cones)
    d_t = DeviceArray.from_host(T)
    d_b = None
    if box:
        d_b = DeviceArray((6,), np.int32)
        E.support_box(d_t, case.shape, d_b)
    d_d = DeviceArray(case.shape, np.float32)
    E.convolve(plan, d_t, d_d, d_b)
    return d_d.download()


def test_scaling_is_exact(engine, case_terma):
    case, T = case_terma
    d1 = gather(engine, case, T)
    d2 = gather(engine, case, 2.0 * T)
    assert d1.max() > 0
    assert np.array_equal(d2, 2.0 * d1)


def test_linearity(engine, case_terma):
    case, T = case_terma
    rng = np.random.default_rng(17)
    T2 = np.zeros_like(T)
    T2[40:80, 20:60, 50:70] = rng.random((40, 40, 20)).astype(np.float32)
    a = gather(engine, case, T)
    b = gather(engine, case, T2)
    ab = gather(engine, case, T + T2)
    assert np.abs(ab - (a + b)).max() / ab.max() < 1e-6


def test_batch_repeats_single_bitwise(engine, case_terma):
    case, T = case_terma
    single = gather(engine, case, T)
    plan = engine.plan(case.shape, case.spacing, case.cones)
    d_t = DeviceArray.from_host(T)
    d_b = DeviceArray((6,), np.int32)
    E.support_box(d_t, case.shape, d_b)
    outs = [DeviceArray(case.shape, np.float32) for _ in range(2)]
    E.convolve_batch([(plan, d_t, o, d_b, False) for o in outs])
    for o in outs:
        assert np.array_equal(o.download(), single)


def test_box_or_none(engine, case_terma):
    case, T = case_terma
    a = gather(engine, case, T, box=True)
    b = gather(engine, case, T, box=False)
    assert np.abs(a - b).max() / a.max() < 1e-6


def fft(engine, case, T):
    plan = engine.plan(case.shape, case.spacing, case.cones)
    d_t = DeviceArray.from_host(T)
    d_d = DeviceArray(case.shape, np.float32)
    E.convolve_fft(plan, d_t, d_d)
    return d_d.download()


def test_fft_form_fullsize(engine, case_terma):
    """The FFT form at 128^3 (frame 256^3): the same exact doubling (every
    FFT operation commutes with a power-of-two scale), linearity, batch ==
    single bitwise, and agreement with the independent gather within 1e-5
    of Dmax (both are within 5e-7 of the FP64 oracle in bench.py)."""
    case, T = case_terma
    d1 = fft(engine, case, T)
    assert d1.max() > 0
    assert np.array_equal(fft(engine, case, 2.0 * T), 2.0 * d1)
    rng = np.random.default_rng(17)
    T2 = np.zeros_like(T)
    T2[40:80, 20:60, 50:70] = rng.random((40, 40, 20)).astype(np.float32)
    b = fft(engine, case, T2)
    ab = fft(engine, case, T + T2)
    assert np.abs(ab - (d1 + b)).max() / ab.max() < 1e-6
    plan = engine.plan(case.shape, case.spacing, case.cones)
    d_t = DeviceArray.from_host(T)
    outs = [DeviceArray(case.shape, np.float32) for _ in range(2)]
    E.convolve_fft_batch([(plan, d_t, o, None, False) for o in outs])
    for o in outs:
        assert np.array_equal(o.download(), d1)
    g = gather(engine, case, T)
    assert np.abs(d1 - g).max() / g.max() < 1e-5
