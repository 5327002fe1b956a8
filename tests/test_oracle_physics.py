"""The oracle's physics-mode convolution (ch_ref_convolve_rad) pinned by
closed forms and by reductions to the geometric cuda_dose convolution.

The reference has no physics mode (kernel tilting is a README milestone,
README.md:22), so the definition is the build's own:
    R_m(v) = sum_{j<=m} rho(v - o_j) (t_j - t_{j-1}),
    dose[v] += sum_c sum_m T(v - o_m) (cum[i(R_m)] - cum[i(R_{m-1})]),
with i(r) = min(|floor(100 r - 5)|, 5995) (conehead/conehead.py:453-463).
What pins it:
  - radiological=False against a forward walk of the reference-faithful
    DDA (ch_ref_dda_3d) from an impulse, i.e. the cuda_dose scatter;
  - rho = 1 reduces to the geometric mode, rho = 2 to the geometric mode on
    voxels twice as large (t scales exactly by 2);
  - a 1-D slab closed form that fixes which voxel's density weights which
    path segment (a forward-path variant fails it).
"""
import numpy as np
import pytest

from conehead_amd.cones import cuda_dose_direction


def table(n_rows=3):
    i = np.arange(5996, dtype=np.float64)
    return np.stack([1e-3 * np.sqrt(i + 1.0) * (j + 1) + 1e-7 * i * i
                     for j in range(n_rows)])


def cuda_index(t):
    return np.minimum(np.abs(np.floor(np.asarray(t) * 100.0 - 5.0)),
                      5995).astype(np.int64)


def cones(seed=0, n=6):
    rng = np.random.default_rng(seed)
    dirs = [cuda_dose_direction(th, ph)
            for th, ph in zip(rng.uniform(0, 360, n), rng.uniform(0, 180, n))]
    dirs += [cuda_dose_direction(0.0, 90.0),   # (1, ~0, 0)
             np.array([0.0, -1.0, 0.0]),       # exact axis, negative step
             np.array([-0.0, 0.6, 0.8])]       # -0.0 component
    dirs = np.array(dirs)
    rows = np.arange(len(dirs), dtype=np.int32) % 3
    return dirs, rows


def test_geometric_impulse_equals_forward_walk(oracle):
    """An impulse T = delta_s: the gather form gives, along each cone, the
    cuda_dose increments of a forward DDA walk from s."""
    shape, dims = (9, 11, 10), (0.5, 0.4, 0.3)
    s = (4, 5, 3)
    cum = table()
    dirs, rows = cones()
    for c in range(len(dirs)):
        T = np.zeros(shape)
        T[s] = 2.0
        dose = np.zeros(shape)
        oracle.convolve_rad(T, None, dose, dims, dirs[c:c + 1],
                            rows[c:c + 1], cum, radiological=False)
        vox, tv, _ = oracle.dda_3d(dirs[c], shape, s, dims)
        idx = cuda_index(tv)
        row = cum[rows[c]]
        k = np.concatenate([[row[idx[0]]], np.diff(row[idx])])
        ref = np.zeros(shape)
        for v, kv in zip(vox, k):
            ref[tuple(v)] += 2.0 * kv
        np.testing.assert_allclose(dose, ref, rtol=1e-13, atol=1e-18)


def test_density_one_is_geometric(oracle):
    shape, dims = (8, 9, 7), (0.4, 0.4, 0.4)
    rng = np.random.default_rng(1)
    T = rng.random(shape) * (rng.random(shape) < 0.3)
    cum = table()
    dirs, rows = cones(2)
    geo = np.zeros(shape)
    rad = np.zeros(shape)
    oracle.convolve_rad(T, None, geo, dims, dirs, rows, cum, False)
    oracle.convolve_rad(T, np.ones(shape, np.float32), rad, dims, dirs, rows,
                        cum, True)
    assert geo.max() > 0
    np.testing.assert_allclose(rad, geo, rtol=1e-12, atol=1e-16)


def test_density_two_is_geometric_on_doubled_voxels(oracle):
    """rho = 2 doubles every radiological length; doubling the voxel size
    doubles every t exactly, so the kernel indices agree."""
    shape, dims = (8, 9, 7), np.array([0.3, 0.25, 0.35])
    rng = np.random.default_rng(3)
    T = rng.random(shape) * (rng.random(shape) < 0.3)
    cum = table()
    dirs, rows = cones(4)
    geo = np.zeros(shape)
    rad = np.zeros(shape)
    oracle.convolve_rad(T, None, geo, 2 * dims, dirs, rows, cum, False)
    oracle.convolve_rad(T, np.full(shape, 2.0, np.float32), rad, dims, dirs,
                        rows, cum, True)
    np.testing.assert_allclose(rad, geo, rtol=1e-12, atol=1e-16)
    # and differs from the geometric result on the original voxels
    geo1 = np.zeros(shape)
    oracle.convolve_rad(T, None, geo1, dims, dirs, rows, cum, False)
    assert np.abs(geo1 - rad).max() > 1e-3 * np.abs(geo1).max()


@pytest.mark.parametrize("sign", [1, -1])
def test_slab_closed_form(oracle, sign):
    """One cone along +-y through a random density column: the dose m voxels
    downstream of an impulse is cum[i(R_m)] - cum[i(R_{m-1})], where R_m
    weights the receiving voxel by half a voxel and every voxel between it
    and the source (source included) by a whole one."""
    shape, s = (3, 24, 3), 0.5
    rng = np.random.default_rng(5)
    col = rng.uniform(0.2, 1.9, shape[1]).astype(np.float32)
    rho = np.broadcast_to(col[None, :, None], shape).copy()
    y0 = 4 if sign > 0 else 19
    T = np.zeros(shape)
    T[1, y0, 1] = 1.0
    dose = np.zeros(shape)
    cum = table(1)
    oracle.convolve_rad(T, rho, dose, (s, s, s),
                        np.array([[0.0, float(sign), 0.0]]),
                        np.zeros(1, np.int32), cum, True)
    ref = np.zeros(shape[1])
    for m in range(0, shape[1]):
        yv = y0 + sign * m
        if not 0 <= yv < shape[1]:
            break
        # backward from the receiver yv toward the source y0
        seg = [col[yv] * (s / 2)] + [col[yv - sign * j] * s
                                     for j in range(1, m + 1)]
        r_m = 0.0
        for x in seg:
            r_m = r_m + float(np.float64(x))
        r_prev = 0.0
        for x in seg[:-1]:
            r_prev = r_prev + float(np.float64(x))
        c_m = cum[0][cuda_index(r_m)]
        c_prev = cum[0][cuda_index(r_prev)] if m > 0 else 0.0
        ref[yv] = c_m - c_prev
    np.testing.assert_allclose(dose[1, :, 1], ref, rtol=1e-12, atol=1e-18)
    assert np.count_nonzero(dose) == np.count_nonzero(ref)


def test_cone_path_matches_dda(oracle):
    """The path table is the reference DDA from the centre of the grid."""
    shape, dims = (6, 7, 8), (0.5, 0.4, 0.3)
    dirs, _ = cones(7)
    for d in dirs:
        off, t = oracle.cone_path(d, shape, dims)
        big = tuple(2 * n + 1 for n in shape)
        vox, tv, _ = oracle.dda_3d(d, big, list(shape), dims)
        n = len(off)
        assert np.array_equal(off, vox[:n] - np.array(shape))
        assert np.array_equal(t, tv[:n])
        assert np.all(np.abs(off).max(axis=0) < np.array(shape))
