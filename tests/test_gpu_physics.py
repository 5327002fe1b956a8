"""Physics mode on the GPU (ch_convolve_rad): beam-aligned cones and the
kernel looked up at the radiological distance, against the oracle's FP64
restatement (ch_ref_convolve_rad).

Tolerance: max |D_gpu - D_ref| / max(D_ref) <= 1e-5.  R, the kernel index
and the increments round like the oracle's (FP64, no contraction); the
remaining difference is the FP32 compensated sum (measured ~1e-7)."""
import numpy as np
import pytest

from conehead_amd import engine as E
from conehead_amd.cones import ConeSet, beam_frame, default_thetas
from conehead_amd.device import DeviceArray
from conehead_amd.kernel import EdkSet, merge_polar
from conehead_amd.source import Source
from conehead_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_err(d, ref):
    return float(np.abs(d - ref).max() / np.abs(ref).max())


@pytest.fixture(scope="module")
def edk():
    poly = EdkSet.load().poly(W.SETTINGS["energy_weights"])
    return merge_polar(poly.angles, poly.kernel_cum, 6)


def physics_cones(edk, gantry, n_theta=6):
    src = Source("varian_clinac_6MV")
    src.gantry = gantry
    phis, cum = edk
    return ConeSet(default_thetas(n_theta), phis, cum, semantics="physics",
                   frame=beam_frame(src))


def heterogeneous(shape, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.2, 1.9, shape).astype(np.float32)
    T = np.zeros(shape, np.float32)
    sl = tuple(slice(s // 4, 3 * s // 4) for s in shape)
    T[sl] = rng.random(T[sl].shape).astype(np.float32)
    return T, rho


def gpu_rad(engine, shape, vs, cs, T, rho, box=True, accumulate=False,
            init=None):
    plan = engine.plan(shape, vs, cs)
    d_t = DeviceArray.from_host(T)
    d_r = DeviceArray.from_host(rho)
    d_d = DeviceArray.from_host(init) if init is not None else \
        DeviceArray(shape, zero=True)
    d_box = None
    if box:
        d_box = DeviceArray((6,), np.int32)
        E.support_box(d_t, shape, d_box)
    E.convolve_rad(plan, d_t, d_r, d_d, d_box, accumulate=accumulate)
    return d_d.download()


@pytest.mark.parametrize("gantry,shape", [(0.0, (20, 24, 18)),
                                          (37.0, (21, 19, 33)),
                                          (250.0, (13, 40, 17))])
def test_matches_oracle(oracle, engine, edk, gantry, shape):
    vs = (0.5, 0.4, 0.6)
    cs = physics_cones(edk, gantry)
    T, rho = heterogeneous(shape, int(gantry) + 1)
    ref = np.zeros(shape)
    oracle.convolve_rad(T, rho, ref, vs, cs.dirs, cs.rows, cs.cum,
                        radiological=True, n_threads=8)
    assert ref.max() > 0
    for box in (True, False):
        d = gpu_rad(engine, shape, vs, cs, T, rho, box=box)
        assert rel_err(d, ref) < TOL, (box, rel_err(d, ref))


def test_density_one_is_the_cuda_dose_gather(engine, edk):
    """rho = 1 and the gantry at 0: the same dose as the geometric cuda_dose
    gather (same cones, same increments up to the R = t rounding)."""
    shape, vs = (18, 22, 20), (0.5, 0.5, 0.5)
    phis, cum = edk
    cs = physics_cones(edk, 0.0)
    geo = ConeSet(default_thetas(6), phis, cum, semantics="cuda_dose")
    assert np.array_equal(cs.dirs, geo.dirs)
    T, _ = heterogeneous(shape, 4)
    rho = np.ones(shape, np.float32)
    d = gpu_rad(engine, shape, vs, cs, T, rho)
    plan = engine.plan(shape, vs, geo)
    d_t = DeviceArray.from_host(T)
    d_g = DeviceArray(shape, zero=True)
    E.convolve(plan, d_t, d_g)
    assert rel_err(d, d_g.download()) < TOL


def test_accumulate_and_empty(engine, edk):
    shape, vs = (9, 10, 11), (0.5, 0.5, 0.5)
    cs = physics_cones(edk, 90.0)
    T, rho = heterogeneous(shape, 6)
    base = gpu_rad(engine, shape, vs, cs, T, rho)
    init = np.full(shape, 0.25, np.float32)
    acc = gpu_rad(engine, shape, vs, cs, T, rho, accumulate=True, init=init)
    np.testing.assert_allclose(acc, base + 0.25, rtol=1e-6, atol=1e-6)
    zero = gpu_rad(engine, shape, vs, cs, np.zeros(shape, np.float32), rho)
    assert np.all(zero == 0.0)


def test_geometric_plans_are_refused(engine, edk):
    from conehead_amd import _lib
    shape, vs = (8, 8, 8), (0.5, 0.5, 0.5)
    cs = physics_cones(edk, 0.0)
    plan = engine.plan(shape, vs, cs)
    d = DeviceArray(shape, zero=True)
    with pytest.raises(_lib.CHError):
        E.convolve(plan, d, d)
    phis, cum = edk
    geo = engine.plan(shape, vs, ConeSet([0.0], phis, cum,
                                         semantics="cuda_dose"))
    with pytest.raises(_lib.CHError):
        E.convolve_rad(geo, d, d, d)


def test_gantry_90_is_gantry_0_rotated(engine):
    """A cubic water phantom: turning the gantry by 90 degrees turns the
    TERMA and the cone set with it, so the physics dose turns too (the 90
    degree rotation maps the grid onto itself)."""
    from conehead_amd.block import Block
    from conehead_amd.conehead import Conehead
    from conehead_amd.phantom import Phantom
    n, s = 24, 0.5
    origin = np.array([-(n - 1) / 2 * s] * 3)
    phantom = Phantom(np.ones((n, n, n), np.float32), origin, [s] * 3)
    blk = Block()
    blk.set_square(6)
    settings = dict(W.SETTINGS, semantics="physics", n_phi=6,
                    thetas=default_thetas(8))
    doses = {}
    for g in (0.0, 90.0):
        src = Source("varian_clinac_6MV")
        src.gantry = g
        doses[g] = Conehead(engine).calculate(src, blk, phantom, settings)
    # gantry 0: beam along +y; gantry 90: beam along -x.  The grid index
    # (i, j) at gantry 90 sees what (n-1-j, i) sees at gantry 0.
    rot = np.transpose(doses[0.0], (1, 0, 2))[::-1, :, :]
    assert rel_err(doses[90.0], rot) < 1e-3, rel_err(doses[90.0], rot)


def test_water_pdd_against_tank(engine):
    """Validation, not parity: the physics-mode central-axis depth dose of
    a 10 x 10 cm field in water (SSD 100, 12 x 12 cones, EDKnrc kernel)
    against the measured tank depth dose the reference carries
    (conehead/conehead.py:1047-1061, tests/golden/tank_10x10.json).
    Beyond the build-up region the two agree within the bound below."""
    import json
    import os
    from conehead_amd.block import Block
    from conehead_amd.conehead import Conehead
    from conehead_amd.phantom import Phantom
    tank = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "tank_10x10.json")))["pdd"]
    nx, ny, s = 40, 64, 0.5
    origin = np.array([-(nx - 1) / 2 * s, s / 2, -(nx - 1) / 2 * s])
    phantom = Phantom(np.ones((nx, ny, nx), np.float32), origin, [s] * 3)
    blk = Block()
    blk.set_square(10)
    settings = dict(W.SETTINGS, semantics="physics", density_xy_swap=False)
    dose = Conehead(engine).calculate(Source("varian_clinac_6MV"), blk,
                                      phantom, settings)
    cax = dose[nx // 2 - 1:nx // 2 + 1, :, nx // 2 - 1:nx // 2 + 1].mean(
        axis=(0, 2))
    depth = origin[1] + s * np.arange(ny)
    calc = 100.0 * cax / cax.max()
    meas = np.interp(depth, tank["depth_cm"], tank["dose_pct"])
    sel = (depth >= 2.0) & (depth <= 25.0)
    dev = np.abs(calc[sel] - meas[sel]).max()
    print("max |calc - tank| over 2-25 cm: %.2f %% of Dmax" % dev)
    print("depth  calc  tank")
    for d, c, m in zip(depth[::4], calc[::4], meas[::4]):
        print("%5.2f %6.2f %6.2f" % (d, c, m))
    assert dev < 2.5


def test_water_profiles_against_tank(engine):
    """Validation, not parity: physics-mode cross-profiles of a 10 x 10 cm
    field in water (SSD 100, 12 x 12 cones, EDKnrc kernel) at 5, 10, 20
    and 30 cm depth against the measured tank profiles the reference plots
    beside its own (conehead/conehead.py:955-1014,
    tests/golden/tank_10x10.json), each normalised to its central value as
    the reference does.  Bounds, in % of the central value: inside the
    field (|x| <= 0.8 of the field half-width projected to that depth)
    within IN_FIELD, outside it (|x| >= 1.25 of the half-width) within
    OUT_FIELD, and the 50 % crossing of each edge within EDGE_MM."""
    import json
    import os

    from conehead_amd.block import Block
    from conehead_amd.conehead import Conehead
    from conehead_amd.phantom import Phantom
    IN_FIELD, OUT_FIELD, EDGE_MM = 4.0, 4.0, 3.0
    prof = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "tank_10x10.json")))["profiles"]
    nx, ny, s = 56, 66, 0.5
    origin = np.array([-(nx - 1) / 2 * s, s / 2, -(nx - 1) / 2 * s])
    phantom = Phantom(np.ones((nx, ny, nx), np.float32), origin, [s] * 3)
    blk = Block()
    blk.set_square(10)
    settings = dict(W.SETTINGS, semantics="physics", density_xy_swap=False)
    dose = Conehead(engine).calculate(Source("varian_clinac_6MV"), blk,
                                      phantom, settings)
    xs = origin[0] + s * np.arange(nx)
    ys = origin[1] + s * np.arange(ny)
    zc = nx // 2 - 1                  # the voxel pair around z = 0
    worst = {}
    for depth_key, data in prof.items():
        depth = float(depth_key)
        x_m = np.asarray(data["x_cm"], float)
        d_m = np.asarray(data["dose_pct"], float)
        # dose along x at this depth: linear in y, mean of the z pair
        line = dose[:, :, zc:zc + 2].mean(axis=2)
        calc = np.array([np.interp(depth, ys, line[i]) for i in range(nx)])
        calc = 100.0 * calc / np.interp(0.0, xs, calc)
        c_at = np.interp(x_m, xs, calc)
        half = 5.0 * (100.0 + depth) / 100.0
        inside = np.abs(x_m) <= 0.8 * half
        outside = (np.abs(x_m) >= 1.25 * half) & (np.abs(x_m) <= xs[-1])
        dev_in = np.abs(c_at - d_m)[inside].max()
        dev_out = np.abs(c_at - d_m)[outside].max()

        def crossing(x, d, sign):
            sel = (x * sign > 0)
            xx, dd = x[sel] * sign, d[sel]
            o = np.argsort(xx)
            xx, dd = xx[o], dd[o]
            k = np.nonzero(dd < 50.0)[0][0]
            return xx[k - 1] + (50.0 - dd[k - 1]) * (xx[k] - xx[k - 1]) / \
                (dd[k] - dd[k - 1])
        edge = max(abs(crossing(xs, calc, sg) - crossing(x_m, d_m, sg))
                   for sg in (-1, 1)) * 10.0
        worst[depth] = (dev_in, dev_out, edge)
        print("depth %4.1f cm: in-field %.2f %%, out-of-field %.2f %%, "
              "50%% edge %.2f mm" % (depth, dev_in, dev_out, edge))
    for depth, (dev_in, dev_out, edge) in worst.items():
        assert dev_in < IN_FIELD, (depth, dev_in)
        assert dev_out < OUT_FIELD, (depth, dev_out)
        assert edge < EDGE_MM, (depth, edge)
