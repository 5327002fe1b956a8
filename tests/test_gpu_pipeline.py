"""End-to-end Conehead.calculate on the GPU against the oracle chain
(oracle TERMA -> oracle convolve_c), within 0.1 % of Dmax (the stated
target); measured errors are ~1e-6."""
import numpy as np
import pytest

from conehead_amd import engine as E
from conehead_amd.block import Block
from conehead_amd.conehead import Conehead, dose_grid_densities
from conehead_amd.kernel2 import PolyenergeticKernel, merged_kernel
from conehead_amd.cones import default_thetas
from conehead_amd.phantom import Phantom
from conehead_amd.source import Source
from conehead_amd.workloads import SETTINGS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["fft", "gather"])
def test_calculate_matches_oracle_chain(oracle, engine, method):
    n, s = 30, 0.5
    shape = (n, n, n)
    origin = np.array([-(n - 1) / 2 * s, s / 2, -(n - 1) / 2 * s])
    rng = np.random.default_rng(5)
    dens = (1.0 + 0.3 * rng.random(shape)).astype(np.float32)
    phantom = Phantom(dens, origin, [s] * 3)
    src = Source("varian_clinac_6MV")
    blk = Block()
    blk.set_square(8)
    k = merged_kernel(PolyenergeticKernel(), 6)
    settings = dict(SETTINGS, kernel=k, thetas=default_thetas(8),
                    phis=np.array(k.keys(), float), convolution=method)
    ch = Conehead(engine)
    dose = ch.calculate(src, blk, phantom, settings)
    assert np.array_equal(ch.dose_grid.dose, dose)
    # oracle chain on the dose-grid densities (the reference's xy swap)
    dd = dose_grid_densities(phantom)
    e, w, mu = E.spectrum(SETTINGS)
    rb = oracle.make_beam(src.position, src.v_y, src.transform, src.sad,
                          SETTINGS, e, w, mu)
    ref_t = oracle.terma(rb, shape, origin, [s] * 3, dd, blk.block_values,
                         n_threads=8)["terma"]
    ref = np.zeros(shape)
    oracle.convolve(ref_t, ref, [s] * 3, settings["thetas"],
                    settings["phis"], k, n_threads=8)
    err = np.abs(dose - ref).max() / ref.max()
    assert err < 1e-3, err
    assert err < 1e-4, err


def test_density_swap_quirk():
    d = np.arange(2 * 2 * 3, dtype=np.float32).reshape(2, 2, 3)
    p = Phantom(d, [0, 0, 0], [1, 1, 1])
    out = dose_grid_densities(p)
    assert out[1, 0, 2] == d[0, 1, 2]
    with pytest.raises(ValueError):
        dose_grid_densities(Phantom(np.ones((2, 3, 4)), [0, 0, 0],
                                    [1, 1, 1]))


def test_calculate_on_separate_dose_grid(engine):
    """settings['dose_grid']: densities resampled on the GPU.  Naming the
    phantom's own grid reproduces the default calculation bit for bit; a
    coarser grid gives a finite dose with its maximum near the phantom
    default's."""
    n, s = 24, 0.5
    shape = (n, n, n)
    origin = np.array([-(n - 1) / 2 * s, s / 2, -(n - 1) / 2 * s])
    rng = np.random.default_rng(6)
    dens = (1.0 + 0.2 * rng.random(shape)).astype(np.float32)
    phantom = Phantom(dens, origin, [s] * 3)
    src = Source("varian_clinac_6MV")
    blk = Block()
    blk.set_square(6)
    k = merged_kernel(PolyenergeticKernel(), 4)
    settings = dict(SETTINGS, kernel=k, thetas=default_thetas(6),
                    phis=np.array(k.keys(), float))
    ch = Conehead(engine)
    base = ch.calculate(src, blk, phantom, settings)
    same = ch.calculate(src, blk, phantom, dict(
        settings, dose_grid=(list(shape), phantom.origin, phantom.spacing)))
    assert np.array_equal(base, same)
    coarse = ch.calculate(src, blk, phantom, dict(
        settings, dose_grid=([16, 16, 16], phantom.origin,
                             np.float32([0.75, 0.75, 0.75]))))
    assert coarse.shape == (16, 16, 16) and np.isfinite(coarse).all()
    assert 0.5 < coarse.max() / base.max() < 2.0


def test_reference_example_config(oracle, engine):
    """The reference's CPU-runnable configuration (example.py): 50^3 water
    at 4 mm, a 10 x 10 cm field at SSD 100, 12 x 8 cones of the varian_6
    kernel -- TERMA and dose against the FP64 oracle chain on the whole
    grid (within 1e-4 of the maximum; measured ~1e-7)."""
    from conehead_amd.device import DeviceArray
    from conehead_amd.workloads import plain_cpu
    case = plain_cpu()
    src, blk, st = case.beams[0]
    d_rho = DeviceArray.from_host(case.densities)
    d_t = DeviceArray(case.shape, np.float32)
    d_box = DeviceArray((6,), np.int32)
    beam = E.beam_struct(src, st, engine.aperture(blk), d_rho, d_t, d_box)
    E.terma([beam], case.shape, case.origin, case.spacing)
    plan = engine.plan(case.shape, case.spacing, case.cones)
    d_d = DeviceArray(case.shape, np.float32)
    E.convolve(plan, d_t, d_d, d_box)
    T, dose = d_t.download(), d_d.download()
    e, w, mu = E.spectrum(st)
    rb = oracle.make_beam(src.position, src.v_y, src.transform, src.sad, st,
                          e, w, mu)
    ref_t = oracle.terma(rb, case.shape, case.origin, case.spacing,
                         case.densities, blk.block_values,
                         n_threads=8)["terma"]
    assert np.abs(T - ref_t).max() / ref_t.max() < 1e-4
    ref = np.zeros(case.shape)
    oracle.convolve_dirs(ref_t, ref, case.spacing, case.cones.dirs,
                         case.cones.rows, case.cones.cum, n_threads=8)
    assert ref.max() > 0
    assert np.abs(dose - ref).max() / ref.max() < 1e-4


@pytest.mark.parametrize("method", ["fft", "gather"])
def test_calculate_cuda_dose_reference_settings(oracle, engine, method):
    """The reference's live dose call (conehead/conehead.py:686-738) from
    its own settings alone: semantics="cuda_dose" builds the EDKnrc
    KernelPoly of settings['energy_weights'] (48 x 5996), its 48 polar
    angles and 12 float32 azimuths, y-polar cones and the |floor(100 t - 5)|
    index clamped at 5995.  Checked on the reference's example grid (50^3
    water at 4 mm, 10 x 10 cm field) against the oracle chain: oracle TERMA,
    then the oracle's geometric gather with the cuda_dose index
    (convolve_rad, radiological=False), within 1e-4 of Dmax."""
    from conehead_amd.cones import ConeSet
    from conehead_amd.kernel import EdkSet
    from conehead_amd.workloads import plain_cpu
    case = plain_cpu()
    src, blk, st = case.beams[0]
    phantom = Phantom(case.densities, case.origin, case.spacing)
    settings = dict(SETTINGS, semantics="cuda_dose", convolution=method)
    dose = Conehead(engine).calculate(src, blk, phantom, settings)
    poly = EdkSet.load().poly(SETTINGS["energy_weights"])
    assert poly.kernel_cum.shape == (48, 5996)
    cs = ConeSet(np.linspace(0, 330, 12, dtype=np.float32), poly.angles,
                 poly.kernel_cum, semantics="cuda_dose")
    assert cs.n_cones == 576
    e, w, mu = E.spectrum(SETTINGS)
    rb = oracle.make_beam(src.position, src.v_y, src.transform, src.sad,
                          SETTINGS, e, w, mu)
    dd = dose_grid_densities(phantom)
    # the dose grid's own geometry: Phantom holds float32 origin and
    # spacing (conehead/phantom.py:13-14), so 0.4 cm is 0.4000000059604645
    # cm here as in the reference, and kernel indices floor(100 t - 5) sit
    # where that spacing puts them
    org = np.asarray(phantom.origin, np.float64)
    spc = np.asarray(phantom.spacing, np.float64)
    assert spc[0] != 0.4
    ref_t = oracle.terma(rb, case.shape, org, spc, dd, blk.block_values,
                         n_threads=16)["terma"]
    ref = np.zeros(case.shape)
    oracle.convolve_rad(ref_t, None, ref, spc, cs.dirs, cs.rows, cs.cum,
                        radiological=False, n_threads=16)
    assert ref.max() > 0
    err = np.abs(dose - ref).max() / ref.max()
    assert err < 1e-4, err
