"""The FP64 oracle's 3D-DDA (oracle/ch_ref.c, ch_ref_dda_3d) pinned by the
reference's notebook goldens, brute force and closed forms."""
import json
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dda_goldens.json")


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as f:
        return json.load(f)


def ray_a_direction():
    r = math.sqrt(14.0)
    return np.array([1.0 / r, 2.0 / r, 3.0 / r])


def test_ray_a_bitwise(oracle, gold):
    """dda_benchmarking.ipynb rays A: 102 t-values, bit for bit."""
    g = gold["ray_A"]
    vox, t, _ = oracle.dda_3d(ray_a_direction(), g["grid_shape"],
                              g["first_voxel"], g["voxel_size"])
    expect = np.array([float(v) for v in g["t"]])
    assert len(t) == len(expect) == 102
    assert np.array_equal(t.view(np.uint64), expect.view(np.uint64))
    # invariants: one unit step per entry, ending on the boundary
    steps = np.abs(np.diff(vox, axis=0)).sum(axis=1)
    assert np.all(steps == 1)
    assert vox[0].tolist() == g["first_voxel"]


def test_ray_c_bitwise(oracle, gold):
    """Ray C exercises the zero-component sentinel and a negative step."""
    g = gold["ray_C"]
    vox, t, dabs = oracle.dda_3d(g["direction"], g["grid_shape"],
                                 g["first_voxel"], g["voxel_size"])
    assert vox.tolist() == g["voxels"]
    states = np.array([[float(v) for v in s] for s in g["t_states"]])
    expect = states.min(axis=1)          # the appended t is the state's min
    assert np.array_equal(t.view(np.uint64), expect.view(np.uint64))
    assert dabs.tolist() == [0.997859, 0.0, 0.065403]


def brute_force(direction, shape, first, size):
    """Merge of the per-axis boundary crossings (k + 1/2) s_a / |d_a|; ties
    broken z over y over x, as the traversal's comparisons do."""
    events = []
    for a in range(3):
        d = abs(direction[a])
        if d == 0:
            continue
        for k in range(0, 2 * max(shape) + 2):
            events.append((((k + 0.5) * size[a]) / d, -a, a))
    events.sort()
    cur = list(first)
    step = [-1 if direction[a] < 0 else 1 for a in range(3)]
    vox, ts = [], []
    for t, _, a in events:
        if not all(0 <= cur[i] < shape[i] for i in range(3)):
            break
        vox.append(list(cur))
        ts.append(t)
        cur[a] += step[a]
    return vox, ts


def test_brute_force_small_grids(oracle):
    rng = np.random.default_rng(5)
    for _ in range(300):
        shape = rng.integers(1, 9, size=3)
        first = [int(rng.integers(0, s)) for s in shape]
        d = rng.normal(size=3)
        d[rng.random(3) < 0.15] = 0.0
        if not np.any(d):
            d[2] = 1.0
        size = rng.uniform(0.1, 2.0, size=3)
        vox, t, _ = oracle.dda_3d(d, shape, first, size)
        bv, bt = brute_force(d, shape, first, size)
        assert len(vox) == len(bv)
        # voxel sequence exact except where two crossings tie within 1 ulp
        np.testing.assert_allclose(t, bt, rtol=4e-15, atol=0)
        assert vox.tolist() == bv or np.any(np.diff(bt) < 1e-12 * max(bt))


def test_axis_aligned_closed_form(oracle):
    vox, t, _ = oracle.dda_3d([0.0, 0.0, -1.0], [5, 5, 20], [2, 2, 10],
                              [0.3, 0.3, 0.5])
    assert len(t) == 11
    assert vox[:, 2].tolist() == list(range(10, -1, -1))
    expect = [0.25]
    for _ in range(10):
        expect.append(expect[-1] + 0.5)      # repeated addition, not k*s
    assert t.tolist() == expect


def test_exact_ties_step_z_then_y_then_x(oracle):
    """Direction (1, 1, 1) on unit voxels: every crossing is an exact
    three-way tie (t = k + 1/2 exactly in binary), so the traversal order
    is fixed by the tie rule alone: z beats y beats x
    (conehead/dda_3d_c.pyx:80-99)."""
    vox, t, _ = oracle.dda_3d([1.0, 1.0, 1.0], [3, 3, 3], [0, 0, 0],
                              [1.0, 1.0, 1.0])
    expect_v, expect_t, cur = [], [], [0, 0, 0]
    for k in range(3):
        for a in (2, 1, 0):
            if max(cur) >= 3:
                break
            expect_v.append(list(cur))
            expect_t.append(k + 0.5)
            cur[a] += 1
    assert vox.tolist() == expect_v[:len(vox)]
    assert len(vox) == 7
    assert t.tolist() == expect_t[:7]
    # x beats nothing: with d = (1, 1, 0) the y step comes first
    vox, _, _ = oracle.dda_3d([1.0, 1.0, 0.0], [3, 3, 1], [0, 0, 0],
                              [1.0, 1.0, 1.0])
    assert vox[:3].tolist() == [[0, 0, 0], [0, 1, 0], [1, 1, 0]]


def test_negative_zero_is_a_zero_component(oracle):
    a = oracle.dda_3d([-0.0, 1.0, 0.0], [4, 6, 4], [1, 1, 1], [1, 1, 1])
    b = oracle.dda_3d([0.0, 1.0, 0.0], [4, 6, 4], [1, 1, 1], [1, 1, 1])
    assert a[0].tolist() == b[0].tolist()
    assert a[1].tolist() == b[1].tolist()
    # step sign of -0.0 is +1; its t stays at the 1e9 sentinel
    assert len(a[1]) == 5


def test_outside_start_gives_empty_path(oracle):
    vox, t, _ = oracle.dda_3d([1.0, 0.0, 0.0], [4, 4, 4], [4, 0, 0],
                              [1, 1, 1])
    assert len(t) == 0 and len(vox) == 0
