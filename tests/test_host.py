"""Host-side mirrors of the reference API: beam frame, plane hit, NIST
attenuation, aperture rasters, kernel tables and cone sets (CPU only).
Expected values follow the reference's own tests where they exist
(tests/test_source.py, test_geometry.py, test_nist.py, test_block.py)."""
import numpy as np
import pytest

from conehead_amd.block import Block, mlc_raster, square_bounds
from conehead_amd.cones import ConeSet, convolve_c_direction, default_thetas
from conehead_amd.geometry import line_block_plane_collision
from conehead_amd.kernel2 import PolyenergeticKernel, merged_kernel
from conehead_amd.nist import mu_water
from conehead_amd.source import Source

R2 = np.sqrt(2.0)


@pytest.mark.parametrize("gantry,coll,pos,vx,vy,vz", [
    (0, 0, [0, -100, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]),
    (45, 0, [100 / R2, -100 / R2, 0], [1 / R2, 1 / R2, 0],
     [-1 / R2, 1 / R2, 0], [0, 0, 1]),
    (225, 0, [-100 / R2, 100 / R2, 0], [-1 / R2, -1 / R2, 0],
     [1 / R2, -1 / R2, 0], [0, 0, 1]),
    (270, 0, [-100, 0, 0], [0, -1, 0], [1, 0, 0], [0, 0, 1]),
    (0, 90, None, [0, 0, 1], [0, 1, 0], [-1, 0, 0]),
    (0, 270, None, [0, 0, -1], [0, 1, 0], [1, 0, 0]),
    (90, 90, [100, 0, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0]),
    (45, 270, [100 / R2, -100 / R2, 0], [0, 0, -1], [-1 / R2, 1 / R2, 0],
     [1 / R2, 1 / R2, 0]),
])
def test_source_frames(gantry, coll, pos, vx, vy, vz):
    s = Source("varian_clinac_6MV")
    if gantry:
        s.gantry = np.float32(gantry)
    if coll:
        s.collimator = np.float32(coll)
    if pos is not None:
        np.testing.assert_array_almost_equal(s.position, pos, decimal=5)
    np.testing.assert_array_almost_equal(s.v_x, vx, decimal=5)
    np.testing.assert_array_almost_equal(s.v_y, vy, decimal=5)
    np.testing.assert_array_almost_equal(s.v_z, vz, decimal=5)


def test_source_transform_and_errors():
    s = Source("varian_clinac_6MV")
    s.gantry = np.float32(90)
    for p, q in [([10, 0, 0], [0, -10, 0]), ([-5, -5, 0], [-5, 5, 0]),
                 ([0, 20, -20], [20, 0, -20])]:
        np.testing.assert_array_almost_equal(np.matmul(s.transform, p), q,
                                             decimal=5)
    with pytest.raises(NotImplementedError):
        Source("varian_clinac_10MV")
    assert Source("varian_clinac_6MV", sad=np.float32(50)).sad == 50
    with pytest.raises(AssertionError):
        s.gantry = 360


@pytest.mark.parametrize("gantry,coll,direction,expect", [
    (0, 0, [0, 1, 0], [0, 0, 0]),
    (90, 0, [-1, 0, 0.5], [0, 0, 50]),
    (0, 270, [1, 1, 1], [100, 0, 100]),
    (315, 45, [1, 1, 1], [0, 0, 100 / R2]),
])
def test_plane_collision(gantry, coll, direction, expect):
    s = Source("varian_clinac_6MV")
    s.gantry = np.float32(gantry)
    if coll:
        s.collimator = np.float32(coll)
    p = line_block_plane_collision(s.position, np.array(direction), s.v_y)
    np.testing.assert_array_almost_equal(p, expect)


def test_plane_collision_parallel_raises():
    s = Source("varian_clinac_6MV")
    with pytest.raises(RuntimeError):
        line_block_plane_collision(s.position, np.array([1, 0, 0]), s.v_y)


def test_mu_water_nodes_and_interpolation():
    assert mu_water(2.0) == np.float32(4.942E-02)
    assert mu_water(6.0) == np.float32(2.770E-02)
    # linear between nodes 2 and 3 MeV
    lo, hi = float(np.float32(4.942E-02)), float(np.float32(3.969E-02))
    assert mu_water(2.5) == pytest.approx((lo + hi) / 2, rel=1e-15)
    with pytest.raises(ValueError):
        mu_water(30.0)


def test_square_block_matches_reference_slice():
    b = Block()
    b.set_square(np.float32(10))
    expect = np.zeros((4000, 4000), np.float32)
    expect[1500:2500, 1500:2500] = 1.0
    assert np.array_equal(b.block_values, expect)
    assert square_bounds(10.0) == (1500, 2500)
    # transmission: floor(100 p) + 2000 - 1
    assert b.transmission(np.array([0.0, 0.0])) == 1.0
    assert b.transmission(np.array([4.995, -4.99])) == 1.0
    assert b.transmission(np.array([5.02, 0.0])) == 0.0
    assert b.transmission(np.array([25.0, 0.0])) == 0.0


def test_block_recipe_union_of_squares():
    b = Block()
    b.set_square(4.0)
    b.set_square(2.0)
    assert b.recipe() == (("zeros",), ("square", 4.0), ("square", 2.0))
    v = b.block_values
    assert v.sum() == 400 * 400
    assert b.recipe()[0] == "host"


def test_mlc_raster_leaf_semantics():
    bound = np.arange(-50, 51, 10)            # 10 leaves of 1 mm (tenths)
    a = np.full(10, -300)
    b = np.full(10, 300)
    r = mlc_raster(bound, a, b, (-1000, 1000), (-50, 50))
    W = 100
    r0 = 2000 - W // 2
    # open between the tips, blocked under the leaves far from the tip
    assert r[r0 + 15, 2000] == 1.0
    assert r[r0 + 15, 1000] == 0.0 and r[r0 + 15, 3000] == 0.0
    # leaf edge rows transmit 0.8 / 0.5 / 0.25 under the leaf
    assert r[r0 + 10, 1000] == np.float32(1) - np.float32(0.2)
    assert r[r0 + 11, 1000] == np.float32(1) - np.float32(0.5)
    assert r[r0 + 12, 1000] == np.float32(1) - np.float32(0.75)
    # tip taper over the last 15 columns of bank A
    cmax = 2000 - 300
    taper = r[r0 + 15, cmax - 15:cmax]
    assert taper[0] == 0.0 and np.all(np.diff(taper) > 0)
    # rows outside the bank and outside the jaws are blocked
    assert np.all(r[:r0] == 0.0) and np.all(r[r0 + W:] == 0.0)
    assert np.all(r[:, :1000] == 0.0) and np.all(r[:, 3000:] == 0.0)


def test_polyenergetic_kernel_tables():
    k = PolyenergeticKernel()
    keys = k.keys()
    assert len(keys) == 48 and keys[0] == "0.00" and keys[-1] == "176.25"
    total = 0.0
    for key in keys:
        c = k.cumulative[key]
        assert c.shape == (5996,)
        assert np.all(np.diff(c) >= 0)
        # np.interp at a node is the node value: r = 0.05 -> first bin,
        # r = 60 -> the full cumulative sum
        assert c[0] == k.differential[key][0]
        assert c[-1] == pytest.approx(np.sum(k.differential[key]),
                                      rel=1e-15)
        total += c[-1]
    # the CSV sums to 0.303080 but is normalised by 0.289622754141
    assert total == pytest.approx(0.303080 / 0.289622754141, rel=1e-4)
    # r = 1.0 cm is resample index 95 and a table node
    assert k.cumulative["90.00"][95] == pytest.approx(
        np.sum(k.differential["90.00"][:10]), rel=1e-14)


def test_merged_kernel_groups():
    k = PolyenergeticKernel()
    m = merged_kernel(k, 8)
    assert m.keys() == ["11.25", "33.75", "56.25", "78.75", "101.25",
                        "123.75", "146.25", "168.75"]
    np.testing.assert_allclose(
        m.cumulative["11.25"],
        sum(k.cumulative[x] for x in ["0.00", "3.75", "7.50", "11.25",
                                      "15.00", "18.75"]), rtol=1e-14)
    assert merged_kernel(k, 12).keys()[0] == "7.50"
    with pytest.raises(ValueError):
        merged_kernel(k, 7)


def test_cone_set_order_and_missing_key():
    k = merged_kernel(PolyenergeticKernel(), 4)
    cs = ConeSet(default_thetas(3), np.array(k.keys(), float), k)
    assert cs.n_cones == 12
    assert cs.rows.tolist() == [0, 1, 2, 3] * 3
    np.testing.assert_array_equal(cs.dirs[5],
                                  convolve_c_direction(120.0, 67.5))
    with pytest.raises(KeyError):
        ConeSet([0.0], [12.0], k)


def test_beam_aligned_cones():
    """Physics cones: the cuda_dose directions in the beam frame.  Gantry 0
    is the identity; at any gantry a cone's angle to the beam direction v_y
    is its polar angle phi."""
    from conehead_amd.cones import (ConeSet, beam_frame,
                                    beam_aligned_direction,
                                    cuda_dose_direction)
    from conehead_amd.source import Source
    src = Source("varian_clinac_6MV")
    for th in (0.0, 30.0, 250.0):
        for ph in (3.75, 90.0, 172.5):
            assert np.array_equal(
                beam_aligned_direction(th, ph, beam_frame(src)),
                cuda_dose_direction(th, ph))
    for g in (37.0, 90.0, 181.0):
        src.gantry = g
        f = beam_frame(src)
        vy = f[1] / np.linalg.norm(f[1])
        cs = ConeSet([0.0, 120.0], [7.5, 67.5], np.zeros((2, 5996)),
                     semantics="physics", frame=f)
        for c, ph in enumerate([7.5, 67.5] * 2):
            d = cs.dirs[c]
            ang = np.degrees(np.arccos(np.dot(d, vy) / np.linalg.norm(d)))
            assert ang == pytest.approx(ph, abs=1e-9)
    with pytest.raises(ValueError):
        ConeSet([0.0], [7.5], np.zeros((1, 5996)), semantics="cuda_dose",
                frame=np.eye(3))


def test_convolve_c_dropin_argument_errors():
    """The drop-in raises where the reference's typed memoryviews and kernel
    lookup raise (conehead/convolve_c.pyx:17-20, :269-273), before any GPU
    work: wrong ndim, wrong dtype, read-only dose, mismatched grids, and a
    missing kernel key."""
    from conehead_amd.convolve_c import convolve_c
    k = merged_kernel(PolyenergeticKernel(), 4)
    phis = np.array(k.keys(), float)
    t = np.zeros((4, 4, 4))
    d = np.zeros((4, 4, 4))
    dims = np.array([0.5, 0.5, 0.5])
    with pytest.raises(ValueError, match="dimensions"):
        convolve_c(np.zeros((4, 4)), d, dims, default_thetas(2), phis, k)
    with pytest.raises(ValueError, match="dtype"):
        convolve_c(t.astype(np.float32), d, dims, default_thetas(2), phis, k)
    with pytest.raises(ValueError, match="dtype"):
        convolve_c(t, d.astype(np.float32), dims, default_thetas(2), phis, k)
    ro = np.zeros((4, 4, 4))
    ro.flags.writeable = False
    with pytest.raises(ValueError, match="read-only"):
        convolve_c(t, ro, dims, default_thetas(2), phis, k)
    with pytest.raises(ValueError, match="does not match"):
        convolve_c(t, np.zeros((4, 4, 5)), dims, default_thetas(2), phis, k)
    with pytest.raises(KeyError):
        convolve_c(t, d, dims, default_thetas(2), [12.0], k)
