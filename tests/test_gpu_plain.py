"""The plain single-beam path (scatter convolution with FP32 atomics, host
round trips per stage): an independent GPU evaluation of convolve_c, so it
checks the gather engine as much as it is checked by it.

Tolerances: against the FP64 oracle 1e-4 of Dmax; against the gather
1e-5 of Dmax (both FP32, different summation orders)."""
import numpy as np
import pytest

from conehead_amd import workloads as W
from conehead_amd import engine as E
from conehead_amd.cones import ConeSet
from conehead_amd.device import DeviceArray
from conehead_amd.plain import plain_beam, plain_case, scatter

pytestmark = pytest.mark.gpu


def small_hetero():
    return W.hetero(n=40, spacing=0.6, n_theta=8, n_phi=6)


def gather_dose(engine, case, terma):
    plan = engine.plan(case.shape, case.spacing, case.cones)
    d_t = DeviceArray.from_host(terma)
    d_d = DeviceArray(case.shape, zero=True)
    E.convolve(plan, d_t, d_d)
    return d_d.download()


def test_scatter_matches_oracle_and_gather(oracle, engine):
    case = small_hetero()
    dose, terma = plain_case(case)
    ref = np.zeros(case.shape)
    oracle.convolve_dirs(terma.astype(np.float64), ref, case.spacing,
                         case.cones.dirs, case.cones.rows, case.cones.cum,
                         n_threads=8)
    dmax = ref.max()
    assert dmax > 0
    assert np.abs(dose - ref).max() / dmax < 1e-4
    g = gather_dose(engine, case, terma)
    assert np.abs(dose - g).max() / dmax < 1e-5


def test_plain_terma_equals_engine_terma(engine):
    """Raster aperture (plain) and analytic rectangles (engine) give the
    same TERMA bit for bit."""
    case = small_hetero()
    src, blk, st = case.beams[0]
    _, t_plain = plain_beam(case.densities, case.origin, case.spacing, src,
                            blk, st, case.cones)
    rho = DeviceArray.from_host(case.densities)
    d_t = DeviceArray(case.shape)
    b = E.beam_struct(src, st, engine.aperture(blk), rho, d_t, None)
    E.terma([b], case.shape, case.origin, case.spacing)
    assert np.array_equal(t_plain, d_t.download())


def test_scatter_cuda_dose_semantics(engine):
    """Offset 5 and the clamp at 5995 (conehead/conehead.py:453-463): the
    scatter and the gather agree on a table whose clamp matters."""
    i = np.arange(5996, dtype=np.float64)
    cum = np.stack([1e-3 * np.sqrt(i + 1) * (j + 1) for j in range(3)])
    cs = ConeSet([0.0, 90.0, 180.0, 270.0], [10.0, 60.0, 120.0], cum,
                 semantics="cuda_dose")
    shape = (10, 200, 10)
    vs = (0.5, 0.5, 0.5)
    rng = np.random.default_rng(3)
    terma = (rng.random(shape) * (rng.random(shape) < 0.2)).astype(
        np.float32)
    plan = engine.plan(shape, vs, cs)
    d_t = DeviceArray.from_host(terma)
    g = DeviceArray(shape, zero=True)
    E.convolve(plan, d_t, g)
    sc = DeviceArray(shape, zero=True)
    scatter(d_t, sc, shape, vs, DeviceArray.from_host(cs.dirs),
            DeviceArray.from_host(cs.rows.astype(np.int32)),
            DeviceArray.from_host(cs.cum), cs.cum.shape[1], cs.n_cones,
            "cuda_dose")
    gd, sd = g.download(), sc.download()
    assert np.abs(gd - sd).max() / np.abs(gd).max() < 1e-5


def test_scatter_rejects_bad_arguments():
    from conehead_amd import _lib
    d = DeviceArray((8,), np.float32)
    t = DeviceArray((8,), np.float64)
    r = DeviceArray((1,), np.int32)
    with pytest.raises(_lib.CHError):
        scatter(d, d, (2, 2, 2), (1, 1, 1), t, r, t, 100, 1)  # short rows
    with pytest.raises(_lib.CHError):
        scatter(d, d, (2, 2, 2), (0, 1, 1), t, r, t, 5996, 1)


def test_edk_polyenergetic_cuda_dose(engine):
    """The reference GPU pipeline's convolution (cuda_dose semantics,
    conehead/conehead.py:385-488) with the polyenergetic EDKnrc kernel of
    :686-704: the gather and the independent scatter agree."""
    from conehead_amd.cones import default_thetas
    from conehead_amd.kernel import EdkSet
    poly = EdkSet.load().poly(W.SETTINGS["energy_weights"])
    cs = ConeSet(default_thetas(4), poly.angles.astype(np.float64),
                 poly.kernel_cum, semantics="cuda_dose")
    assert cs.n_cones == 4 * 48
    shape, vs = (20, 24, 18), (0.5, 0.5, 0.5)
    rng = np.random.default_rng(8)
    terma = np.zeros(shape, np.float32)
    terma[6:14, 2:20, 5:13] = rng.random((8, 18, 8)).astype(np.float32)
    plan = engine.plan(shape, vs, cs)
    d_t = DeviceArray.from_host(terma)
    g = DeviceArray(shape, zero=True)
    E.convolve(plan, d_t, g)
    sc = DeviceArray(shape, zero=True)
    scatter(d_t, sc, shape, vs, DeviceArray.from_host(cs.dirs),
            DeviceArray.from_host(cs.rows.astype(np.int32)),
            DeviceArray.from_host(np.ascontiguousarray(cs.cum)),
            cs.cum.shape[1], cs.n_cones, "cuda_dose")
    gd, sd = g.download(), sc.download()
    assert gd.max() > 0
    assert np.abs(gd - sd).max() / gd.max() < 1e-5
