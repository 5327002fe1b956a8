"""CT numbers to densities on the GPU (ch_ct_to_density) against
numpy.interp, the routine the calibration curve is defined by, bit for bit;
density overrides by label; and a CtPhantom through Conehead.calculate
giving the same dose as a Phantom holding numpy's densities."""
import numpy as np
import pytest

from conehead_amd import _lib
from conehead_amd.block import Block
from conehead_amd.conehead import Conehead
from conehead_amd.cones import default_thetas
from conehead_amd.density import DEFAULT_CT_CURVE, ct_to_density
from conehead_amd.device import DeviceArray
from conehead_amd.kernel2 import PolyenergeticKernel, merged_kernel
from conehead_amd.phantom import CtPhantom, Phantom
from conehead_amd.source import Source
from conehead_amd.workloads import SETTINGS

pytestmark = pytest.mark.gpu

CURVE5 = ((-1024.0, 0.0), (-950.0, 0.05), (-100.0, 0.93), (120.0, 1.09),
          (2500.0, 2.45))


def numpy_density(ct, curve, labels=None, overrides=None):
    c = np.asarray(curve, dtype=np.float64)
    rho = np.interp(ct.astype(np.float64), c[:, 0], c[:, 1]).astype(
        np.float32)
    if labels is not None:
        for lab, v in overrides.items():
            rho[labels == lab] = np.float32(v)
    return rho


@pytest.mark.parametrize("curve", [DEFAULT_CT_CURVE, CURVE5,
                                   ((40.0, 1.05),)])
def test_int16_matches_numpy_interp(curve):
    rng = np.random.default_rng(11)
    ct = rng.integers(-1500, 4000, size=(33, 17, 29)).astype(np.int16)
    nodes = np.asarray(curve)[:, 0].astype(np.int16)
    ct.flat[:len(nodes)] = nodes                      # exactly on a node
    ct.flat[len(nodes):2 * len(nodes)] = nodes + 1
    got = ct_to_density(DeviceArray.from_host(ct), curve).download()
    want = numpy_density(ct, curve)
    assert got.dtype == np.float32 and got.shape == ct.shape
    assert np.array_equal(got, want)
    if len(curve) > 1:                                # both clamps and slopes
        assert (ct < nodes[0]).any() and (ct > nodes[-1]).any()


def test_float32_matches_numpy_interp():
    rng = np.random.default_rng(12)
    ct = (rng.random(5000) * 5000 - 1500).astype(np.float32)
    got = ct_to_density(DeviceArray.from_host(ct), CURVE5).download()
    assert np.array_equal(got, numpy_density(ct, CURVE5))


def test_label_overrides():
    rng = np.random.default_rng(13)
    ct = rng.integers(-1100, 3000, size=(20, 21, 22)).astype(np.int16)
    labels = rng.choice(np.array([0, 1, 2, 7, 255], np.uint8), size=ct.shape)
    overrides = {1: 1.0, 7: 0.25, 255: 11.3}          # label 2 keeps its CT
    got = ct_to_density(DeviceArray.from_host(ct), CURVE5,
                        DeviceArray.from_host(labels), overrides).download()
    want = numpy_density(ct, CURVE5, labels, overrides)
    assert np.array_equal(got, want)
    assert (got[labels == 7] == np.float32(0.25)).all()
    assert np.array_equal(got[labels == 2],
                          numpy_density(ct[labels == 2], CURVE5))


def test_bad_curves_are_refused():
    d = DeviceArray.from_host(np.zeros(8, np.int16))
    for curve in [((0.0, 1.0), (0.0, 2.0)), ((10.0, 1.0), (-10.0, 2.0)),
                  tuple((float(k), 1.0) for k in range(40)),
                  ((0.0, np.inf), (1.0, 1.0))]:
        with pytest.raises(_lib.CHError):
            ct_to_density(d, curve)
    with pytest.raises(TypeError):
        ct_to_density(DeviceArray.from_host(np.zeros(4, np.float64)))
    with pytest.raises(ValueError):
        ct_to_density(d, overrides={1: 1.0})          # no label map


def test_calculate_with_ct_phantom(engine):
    """The CT phantom's dose equals the dose of a Phantom holding numpy's
    densities bit for bit, on its own grid and on a separate dose grid."""
    n, s = 24, 0.5
    shape = (n, n, n)
    origin = np.array([-(n - 1) / 2 * s, s / 2, -(n - 1) / 2 * s])
    rng = np.random.default_rng(14)
    ct = rng.integers(-300, 600, size=shape).astype(np.int16)
    labels = np.zeros(shape, np.uint8)
    labels[8:14, :, 8:14] = 3                          # a structure at water
    ctp = CtPhantom(ct, origin, [s] * 3, CURVE5, labels, {3: 1.0})
    ref = Phantom(numpy_density(ct, CURVE5, labels, {3: 1.0}), origin,
                  [s] * 3)
    assert np.array_equal(ctp.densities, ref.densities)
    src = Source("varian_clinac_6MV")
    blk = Block()
    blk.set_square(6)
    k = merged_kernel(PolyenergeticKernel(), 4)
    settings = dict(SETTINGS, kernel=k, thetas=default_thetas(6),
                    phis=np.array(k.keys(), float))
    ch = Conehead(engine)
    a = ch.calculate(src, blk, ctp, settings).copy()
    b = ch.calculate(src, blk, ref, settings).copy()
    assert np.array_equal(a, b) and a.max() > 0
    grid = ((16, 16, 16), origin + 0.1, np.array([0.7] * 3))
    a = ch.calculate(src, blk, ctp, dict(settings, dose_grid=grid)).copy()
    b = ch.calculate(src, blk, ref, dict(settings, dose_grid=grid)).copy()
    assert np.array_equal(a, b) and a.max() > 0
