"""GPU fused TERMA against the FP64 oracle restatement of the reference's
TERMA chain (conehead/conehead.py:71-377, :641-682).

Tolerances: blocked fraction bit-exact (FP64 geometry on both sides);
radiological depth, OAD and fluence within 2e-6 relative; TERMA within
1e-5 of its maximum (FP32 arithmetic on the GPU)."""
import numpy as np
import pytest

from conehead_amd import engine as E
from conehead_amd.block import Block
from conehead_amd.device import DeviceArray
from conehead_amd.source import Source
from conehead_amd.workloads import SETTINGS, hetero_densities

pytestmark = pytest.mark.gpu


def compare(oracle, engine, shape, spacing, origin, dens, gantry, block,
            collimator=0.0, settings=SETTINGS):
    src = Source("varian_clinac_6MV")
    if gantry:
        src.gantry = gantry
    if collimator:
        src.collimator = collimator
    e, w, mu = E.spectrum(settings)
    rb = oracle.make_beam(src.position, src.v_y, src.transform, src.sad,
                          settings, e, w, mu)
    ref = oracle.terma(rb, shape, origin, spacing, dens, block.block_values,
                       n_threads=8)
    d_rho = DeviceArray.from_host(dens.astype(np.float32))
    d_t = DeviceArray(shape, np.float32)
    d_box = DeviceArray((6,), np.int32)
    d_dbg = DeviceArray((4,) + tuple(shape), np.float32)
    # analytic aperture (rectangles / leaf table) with the bounding-box
    # early out, and the materialised raster: identical TERMA bits
    ap = engine.aperture(block)
    beam = E.beam_struct(src, settings, ap, d_rho, d_t, d_box)
    E.terma([beam], shape, origin, spacing, debug=d_dbg)
    T = d_t.download()
    dbg = d_dbg.download()
    box = d_box.download()
    d_t2 = DeviceArray(shape, np.float32)
    beam2 = E.beam_struct(src, settings, engine.raster(block), d_rho, d_t2,
                          None)
    E.terma([beam2], shape, origin, spacing)
    assert np.array_equal(d_t2.download(), T)
    assert np.array_equal(dbg[0], ref["blocked"].astype(np.float32))
    m = ref["blocked"] > 0
    np.testing.assert_allclose(dbg[1][m], ref["d_eff"][m], rtol=2e-6,
                               atol=1e-6)
    np.testing.assert_allclose(dbg[2], ref["oad"], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(dbg[3], ref["fluence"], rtol=2e-6,
                               atol=1e-7)
    err = np.abs(T - ref["terma"]).max() / ref["terma"].max()
    assert err < 1e-5, err
    nz = np.argwhere(ref["terma"] != 0)
    assert box[:3].tolist() == nz.min(0).tolist()
    assert box[3:].tolist() == nz.max(0).tolist()
    return T, ref


def test_off_axis_boundaries(oracle, engine):
    """The annular window (rOuter moved to 0.4 so both ends lie in the
    grid), the softening limit (8 cm) and the exponential clamp switch on
    the same voxels as in the oracle: voxel (30, j, 20) images exactly at
    OAD 5 (OAD zAnn / SAD == rInner), (40, 0, 20) at OAD 10 (== rOuter),
    (36, 0, 20) at OAD 8 (== softLimit).  The oracle's off-axis arithmetic
    is pinned by closed forms in test_oracle_terma_offaxis."""
    shape = (41, 9, 41)
    origin = np.array([-10.0, 0.0, -10.0])
    rng = np.random.default_rng(11)
    dens = (1.0 + 0.3 * rng.standard_normal(shape)).clip(0.1, None)
    blk = Block()
    blk.set_square(40)
    st = dict(SETTINGS, rOuter=0.4, softLimit=8.0)
    compare(oracle, engine, shape, [0.5] * 3, origin,
            dens.astype(np.float32), 0.0, blk, settings=st)


def test_water_square_field(oracle, engine):
    n, s = 40, 0.4
    shape = (n, n, n)
    origin = np.array([-(n - 1) / 2 * s, s / 2, -(n - 1) / 2 * s])
    blk = Block()
    blk.set_square(10)
    compare(oracle, engine, shape, [s] * 3, origin,
            np.ones(shape, np.float32), 0.0, blk)


def test_long_axis_walk(oracle, engine):
    """A grid with an axis above 512 voxels (600 along the beam axis y):
    the depth walk keeps one counter per axis there (grids up to 512 per
    axis pack them into one word), and its walks cross hundreds of voxels
    toward the source."""
    shape = (10, 600, 10)
    spacing = [0.5, 0.05, 0.5]
    origin = np.array([-2.25, 0.025, -2.25])
    rng = np.random.default_rng(13)
    dens = (1.0 + 0.3 * rng.standard_normal(shape)).clip(0.1, None)
    blk = Block()
    blk.set_square(4)
    compare(oracle, engine, shape, spacing, origin,
            dens.astype(np.float32), 15.0, blk)


def test_samples_on_cell_boundaries(oracle, engine):
    """Sub-samples whose plane images fall on raster cell boundaries: with
    3 mm voxels and 3 sub-samples per axis the samples sit on a 1 mm grid,
    and the layer at y = 0 (the isocentre plane) images onto multiples of
    10 cells, the field edges of set_square(10) included.  The fast cell
    path declines every such sample (within 1e-6 cell of an integer) and
    the definition's plane image decides: the blocked fraction is still
    bitwise the oracle's."""
    shape = (40, 8, 40)
    origin = np.array([-6.0, 0.1, -6.0])
    blk = Block()
    blk.set_square(10)
    rng = np.random.default_rng(12)
    dens = (1.0 + 0.2 * rng.standard_normal(shape)).clip(0.1, None)
    compare(oracle, engine, shape, [0.3] * 3, origin,
            dens.astype(np.float32), 0.0, blk)


def test_heterogeneous_oblique_mlc(oracle, engine):
    n, s = 48, 0.5
    shape = (n, n, n)
    origin = np.array([-(n - 1) / 2 * s] * 3)
    dens = hetero_densities(n, s)
    blk = Block()
    bound = np.arange(-150, 151, 5.0)
    rng = np.random.default_rng(0)
    a = rng.uniform(-80, -10, 60)
    b = rng.uniform(10, 80, 60)
    blk.set_mlc(bound, a, b, (-100, 100), (-60, 60))
    compare(oracle, engine, shape, [s] * 3, origin, dens, 37.5, blk,
            collimator=15.0)


def test_device_mlc_raster_matches_host(engine):
    blk = Block()
    bound = np.arange(-150, 151, 5.0)
    rng = np.random.default_rng(1)
    blk.set_mlc(bound, rng.uniform(-90, 5, 60), rng.uniform(-5, 90, 60),
                (-95.5, 101.2), (-70.3, 66.6))
    d = engine.raster(blk).download()
    assert np.array_equal(d, blk.block_values)


def test_device_square_raster_matches_host(engine):
    blk = Block()
    blk.set_square(7.3)
    blk.set_square(2.0)
    d = engine.raster(blk).download()
    assert np.array_equal(d, blk.block_values)


def test_dicom_plan_mlc(oracle, engine):
    """The MLC and jaws of the reference's sample RT plan
    (RP.3DCRT.dcm, conehead/block.py:87-215): the analytic leaf table, the
    materialised raster and the oracle agree."""
    import os
    from conehead_amd import dicom
    plan = dicom.read_file(os.path.join(os.path.dirname(__file__), "golden",
                                        "RP.3DCRT.dcm"))
    blk = Block()
    blk._set_from_plan(plan)
    n, s = 40, 0.4
    shape = (n, n, n)
    origin = np.array([-(n - 1) / 2 * s, s / 2, -(n - 1) / 2 * s])
    dens = hetero_densities(n, s)
    T, ref = compare(oracle, engine, shape, [s] * 3, origin, dens, 0.0, blk)
    assert T.max() > 0 and (T == 0).any()


def test_geometry_error_sticky_until_checked(engine):
    """The geometry error of conehead/conehead.py:34-35 (the grid of the
    oracle test that raises: the source plane y = -100 cuts it) is reported
    by the next check after any number of TERMA calls, and cleared by it."""
    from conehead_amd.device import Stream
    shape, spacing = (8, 8, 8), (0.5, 0.5, 0.5)
    bad_origin, good_origin = (2.0, -102.0, -2.0), (-2.0, -2.0, -2.0)
    src = Source("varian_clinac_6MV")
    blk = Block()
    blk.set_square(10.0)
    d_rho = DeviceArray.from_host(np.ones(shape, np.float32))
    d_t = DeviceArray(shape, np.float32)
    beam = E.beam_struct(src, SETTINGS, engine.aperture(blk), d_rho, d_t,
                         None)
    s = Stream()
    with pytest.raises(RuntimeError, match="within plane"):
        E.terma([beam], shape, bad_origin, spacing, stream=s)
    E.terma([beam], shape, good_origin, spacing, stream=s)    # cleared
    # a failing call followed by a good one: one check reports it
    E.terma([beam], shape, bad_origin, spacing, stream=s, check=False)
    E.terma([beam], shape, good_origin, spacing, stream=s, check=False)
    with pytest.raises(RuntimeError, match="within plane"):
        E.terma_check(s)
    E.terma_check(s)
    s.sync()
    s.close()


def test_repeated_calls_keep_their_counters(engine):
    """Phase 1's cost-class counters are kept between calls: they only
    grow, each call counts its bricks from the totals the previous call
    left, and no launch clears them.  Calls on one stream that alternate two
    grids (different list capacities), two beams sharing a launch, and a
    captured graph replayed between eager calls all give the bits of each
    grid's first call."""
    from conehead_amd.device import Stream
    src = Source("varian_clinac_6MV")
    src.gantry = 30.0
    blk = Block()
    blk.set_square(6.0)
    rng = np.random.default_rng(5)
    s = Stream()
    cases = []
    for shape, d in (((40, 36, 44), 0.25), ((24, 20, 28), 0.5)):
        spacing = (d, d, d)
        origin = tuple(-0.5 * n * d for n in shape)
        dens = rng.uniform(0.2, 1.8, shape).astype(np.float32)
        d_rho = DeviceArray.from_host(dens)
        d_t = DeviceArray(shape, np.float32)
        d_t2 = DeviceArray(shape, np.float32)
        ap = engine.aperture(blk)
        beams = [E.beam_struct(src, SETTINGS, ap, d_rho, t, None)
                 for t in (d_t, d_t2)]
        cases.append((shape, origin, spacing, beams, d_t, d_t2, d_rho, ap))
    first = []
    for shape, origin, spacing, beams, d_t, *_ in cases:
        E.terma(beams[:1], shape, origin, spacing, stream=s)
        first.append(d_t.download(stream=s))
        assert first[-1].max() > 0
    for rep in range(3):
        for (shape, origin, spacing, beams, d_t, d_t2, *_), ref in zip(
                cases, first):
            E.terma(beams[:1], shape, origin, spacing, stream=s)
            assert np.array_equal(d_t.download(stream=s), ref)
            E.terma(beams, shape, origin, spacing, stream=s)  # one launch
            assert np.array_equal(d_t.download(stream=s), ref)
            assert np.array_equal(d_t2.download(stream=s), ref)
    (shape_a, origin_a, spacing_a, beams_a, t_a, *_) = cases[0]
    (shape_b, origin_b, spacing_b, beams_b, t_b, *_) = cases[1]
    with E.Graph.capture(s) as cap:
        E.terma(beams_a[:1], shape_a, origin_a, spacing_a, stream=s,
                check=False)
    g = cap.graph
    for rep in range(3):
        E.terma(beams_b[:1], shape_b, origin_b, spacing_b, stream=s)
        assert np.array_equal(t_b.download(stream=s), first[1])
        g.launch(s)
        s.sync()
        E.terma_check(s)
        assert np.array_equal(t_a.download(stream=s), first[0])
    g.close()
    s.sync()
    s.close()
