"""GPU density resampling against scipy's RegularGridInterpolator, the
routine the reference calls (conehead/conehead.py:525-533, :561-568):
linear, fill 0 outside, axes linspace(origin, origin + n * spacing, n),
meshgrid 'xy' layout.  FP64 coordinates on both sides; the float32 results
agree to one unit in the last place."""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from conehead_amd.conehead import dose_grid_densities
from conehead_amd.density import resample
from conehead_amd.device import DeviceArray
from conehead_amd.phantom import Phantom

pytestmark = pytest.mark.gpu


def axes(origin, spacing, size):
    return [np.linspace(origin[a], origin[a] + size[a] * spacing[a], size[a])
            for a in range(3)]


def scipy_ref(dens, so, ss, dshape, do, ds, swap):
    f = RegularGridInterpolator(axes(so, ss, dens.shape), dens,
                                method="linear", bounds_error=False,
                                fill_value=0)
    lx, ly, lz = axes(do, ds, dshape)
    if swap:
        xx, yy, zz = np.meshgrid(lx, ly, lz)          # indexing='xy'
    else:
        xx, yy, zz = np.meshgrid(lx, ly, lz, indexing="ij")
    return np.float32(f((xx, yy, zz)))


@pytest.mark.parametrize("swap", [True, False])
def test_resample_matches_scipy(swap):
    rng = np.random.default_rng(4)
    dens = rng.random((20, 24, 18)).astype(np.float32) * 2
    so, ss = np.array([-5.0, -3.0, -4.5]), np.array([0.5, 0.4, 0.55])
    dshape = (16, 16, 12) if swap else (17, 13, 12)
    do, ds = np.array([-6.0, -2.0, -4.0]), np.array([0.71, 0.63, 0.8])
    want = scipy_ref(dens, so, ss, dshape, do, ds, swap)
    got = resample(DeviceArray.from_host(dens), dens.shape, so, ss, dshape,
                   do, ds, swap_xy=swap).download()
    assert got.shape == want.shape
    assert (want == 0).any() and (want > 0).any()      # both regimes hit
    np.testing.assert_array_max_ulp(got, want, maxulp=1)


def test_same_grid_is_the_reference_transpose():
    """On the phantom's own grid every point is a node: the result is the
    xy-transposed densities, bit for bit."""
    rng = np.random.default_rng(5)
    dens = rng.random((14, 14, 9)).astype(np.float32)
    ph = Phantom(dens, [-3.0, 0.25, -2.0], [0.5, 0.5, 0.5])
    got = resample(DeviceArray.from_host(dens), dens.shape, ph.origin,
                   ph.spacing, dens.shape, ph.origin, ph.spacing).download()
    assert np.array_equal(got, dose_grid_densities(ph))
