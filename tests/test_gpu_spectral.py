"""Spectral (FFT) form of the convolution against the FP64 oracle.

convolve_c is a fixed stencil per cone, so summed over cones the dose is one
linear convolution of T with the plan's kernel K; ch_convolve_fft computes
it through zero-padded FP32 FFTs.  Checks:
  - the oracle's convolve_c (scatter form) on shapes whose frame lengths
    need radices 2, 3, 4, 5 and 8, odd y (an unpaired z line), a one-voxel
    z axis and both translation-invariant semantics;
  - an impulse T = delta_s: the dose is the plan's summed tap table placed
    at s (a closed form that a dropped tap, a wrong sign or a transposed
    axis fails);
  - accumulate mode, and radiological plans refused;
  - the zero-support skips: a sparse T after dense ones on the same work
    grid, alone and in a batch, and an all-zero T;
  - the direct gather on the same plan.
Tolerance: max |D - D_ref| / max(D_ref) <= 1e-5 (target 1e-3; FP32 FFTs
land near 3e-7)."""
import numpy as np
import pytest

from conehead_amd import engine as E
from conehead_amd.cones import ConeSet, default_thetas
from conehead_amd.device import DeviceArray
from conehead_amd.kernel2 import PolyenergeticKernel, merged_kernel

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_err(d, ref):
    return float(np.abs(d - ref).max() / np.abs(ref).max())


@pytest.fixture(scope="module")
def kern():
    return PolyenergeticKernel()


def cone_set(kern, n_theta, n_phi, semantics="convolve_c"):
    k = merged_kernel(kern, n_phi)
    phis = np.array(k.keys(), float)
    if semantics != "convolve_c":
        # table rows instead of the kernel dict
        k = np.stack([k.cumulative["%.2f" % p] for p in phis])
    return ConeSet(default_thetas(n_theta), phis, k, semantics=semantics)


def beam_like(shape, rng, frac=0.5):
    T = np.zeros(shape)
    sl = tuple(slice(int(s * (1 - frac) / 2), int(s * (1 + frac) / 2) + 1)
               for s in shape)
    T[sl] = 1.0 + rng.random(T[sl].shape)
    return T


def run_fft(plan, T, accumulate_onto=None):
    d_t = DeviceArray.from_host(T.astype(np.float32))
    if accumulate_onto is None:
        d_d = DeviceArray(T.shape, np.float32)
    else:
        d_d = DeviceArray.from_host(accumulate_onto.astype(np.float32))
    E.convolve_fft(plan, d_t, d_d, accumulate=accumulate_onto is not None)
    return d_d.download()


@pytest.mark.parametrize("shape,vs,semantics", [
    ((24, 26, 22), (0.4, 0.4, 0.4), "convolve_c"),    # frames 48, 54, 45
    ((17, 9, 31), (0.3, 0.5, 0.25), "convolve_c"),    # odd y: 36, 18, 64
    ((20, 13, 1), (0.4, 0.4, 0.4), "convolve_c"),     # one z voxel: 40, 25, 1
    ((30, 28, 40), (0.35, 0.3, 0.3), "cuda_dose"),    # 60, 60, 80
    ((6, 5, 3), (0.5, 0.5, 0.5), "convolve_c"),       # 12, 9, 5: one-stage z
])
def test_matches_oracle(oracle, engine, kern, shape, vs, semantics):
    rng = np.random.default_rng(sum(shape))
    T = beam_like(shape, rng)
    cs = cone_set(kern, 6, 4, semantics)
    ref = np.zeros(shape)
    if semantics == "convolve_c":
        oracle.convolve_dirs(T, ref, vs, cs.dirs, cs.rows, cs.cum,
                             n_threads=8)
    else:
        # the cuda_dose index |floor(100 t - 5)| clamped at 5995, geometric
        oracle.convolve_rad(T, np.ones(shape, np.float32), ref, vs, cs.dirs,
                            cs.rows, cs.cum, radiological=False, n_threads=8)
    plan = engine.plan(shape, vs, cs)
    frame, pitch, work = plan.spectrum()
    assert all(f >= 2 * n - 1 for f, n in zip(frame, shape))
    assert pitch >= frame[2] // 2 + 1
    d = run_fft(plan, T)
    assert rel_err(d, ref) < TOL


def test_impulse_gives_the_tap_table(engine, kern):
    """T = delta at s: dose(s + o) = sum of the increments of every tap at
    offset o (over all cones), zero elsewhere."""
    shape, vs = (21, 18, 25), (0.3, 0.3, 0.3)
    cs = cone_set(kern, 6, 4)
    plan = engine.plan(shape, vs, cs)
    s = np.array([7, 11, 9])
    T = np.zeros(shape)
    T[tuple(s)] = 1.0
    want = np.zeros(shape)
    for c in range(cs.n_cones):
        off, k = plan.cone_taps(c)
        v = off + s
        ok = np.all((v >= 0) & (v < np.array(shape)), axis=1)
        np.add.at(want, tuple(v[ok].T), k[ok])
    d = run_fft(plan, T)
    assert rel_err(d, want) < TOL
    far = want == 0
    assert np.abs(d[far]).max() < 1e-5 * want.max()


def test_accumulate_and_direct_gather(engine, kern):
    rng = np.random.default_rng(5)
    shape, vs = (40, 36, 44), (0.3, 0.35, 0.25)
    cs = cone_set(kern, 12, 8)
    plan = engine.plan(shape, vs, cs)
    T = beam_like(shape, rng)
    d_t = DeviceArray.from_host(T.astype(np.float32))
    d_box = DeviceArray((6,), np.int32)
    E.support_box(d_t, shape, d_box)
    d_g = DeviceArray(shape, np.float32)
    E.convolve(plan, d_t, d_g, d_box)
    gather = d_g.download()
    fft = run_fft(plan, T)
    assert rel_err(fft, gather) < TOL
    base = rng.random(shape).astype(np.float32)
    both = run_fft(plan, T, accumulate_onto=base)
    assert rel_err(both - base, gather) < TOL


def test_zero_support_leaves_no_stale_values(engine, kern):
    """The forward passes skip zero line pairs and zero slabs of T (support
    flags behind the work grid) and leave those parts of the work grid
    unwritten.  A sparse T convolved right after dense ones on the same
    work grid must give the same bits whatever the dense T was, agree with
    the gather, and an all-zero T must give a dose of exact zeros."""
    rng = np.random.default_rng(17)
    shape, vs = (24, 17, 20), (0.35, 0.3, 0.4)   # odd y: an unpaired line
    cs = cone_set(kern, 6, 4)
    plan = engine.plan(shape, vs, cs)
    dense = (0.5 + rng.random(shape)).astype(np.float32)
    sparse = np.zeros(shape, np.float32)
    sparse[3] = 1.0 + rng.random(shape[1:])        # a full slab
    sparse[11, 4:9] = 1.0 + rng.random((5, shape[2]))
    sparse[11, 16] = 2.0                            # the unpaired line only
    sparse[19, 0:2, 7] = 3.0                        # one line pair, one z
    after_a = (run_fft(plan, dense), run_fft(plan, sparse))[1]
    after_b = (run_fft(plan, 3.0 * dense), run_fft(plan, sparse))[1]
    assert np.array_equal(after_a, after_b)
    d_t = DeviceArray.from_host(sparse)
    d_box = DeviceArray((6,), np.int32)
    E.support_box(d_t, shape, d_box)
    d_g = DeviceArray(shape, np.float32)
    E.convolve(plan, d_t, d_g, d_box)
    assert rel_err(after_a, d_g.download()) < TOL
    # in a batch each job has its own work grid and flags
    jobs = []
    for T in (dense, sparse, np.zeros(shape, np.float32), sparse[::-1]):
        jobs.append((plan, DeviceArray.from_host(np.ascontiguousarray(T)),
                     DeviceArray(shape, np.float32), None, False))
    E.convolve_fft_batch(jobs)
    assert np.array_equal(jobs[1][2].download(), after_a)
    assert not jobs[2][2].download().any()
    assert np.array_equal(jobs[3][2].download(),
                          run_fft(plan, np.ascontiguousarray(sparse[::-1])))
    run_fft(plan, dense)
    assert not run_fft(plan, np.zeros(shape, np.float32)).any()


def test_radiological_plan_refused(engine, kern):
    cs = cone_set(kern, 4, 4, "physics")
    plan = engine.plan((16, 16, 16), (0.4, 0.4, 0.4), cs)
    with pytest.raises(RuntimeError, match="radiological"):
        plan.spectrum()


def test_batch_equals_single_bitwise(engine, kern):
    """Jobs on three plans (shared launches per plan, accumulate flags
    mixed) give the same bits as one ch_convolve_fft call each."""
    rng = np.random.default_rng(9)
    jobs, want = [], []
    specs = [((20, 30, 40), 4), ((33, 17, 29), 8), ((20, 30, 40), 4),
             ((48, 40, 36), 6), ((33, 17, 29), 8)]
    for i, (shape, nphi) in enumerate(specs):
        plan = engine.plan(shape, (0.3, 0.25, 0.35), cone_set(kern, 4, nphi))
        T = beam_like(shape, rng).astype(np.float32)
        base = rng.random(shape).astype(np.float32)
        acc = i % 2 == 1
        want.append(run_fft(plan, T, accumulate_onto=base if acc else None))
        d_t = DeviceArray.from_host(T)
        d_d = DeviceArray.from_host(base)
        jobs.append((plan, d_t, d_d, None, acc))
    E.convolve_fft_batch(jobs)
    for (_, _, d_d, _, _), w in zip(jobs, want):
        assert np.array_equal(d_d.download(), w)


def test_batch_engine_fft_method(engine):
    from conehead_amd import workloads as W
    from conehead_amd.batch import BatchEngine, DoseRequest
    from conehead_amd.device import Stream
    cases = W.uneven(n_phantoms=3, seed=11, max_beams=3, xy=(16, 36),
                     z=(14, 40), spacing=(0.4, 0.8), n_theta=6,
                     n_phis=(4, 6))
    reqs = [DoseRequest.from_case(c) for c in cases]
    s = Stream()
    gather = BatchEngine(engine, method="gather").run(reqs, s)
    fft_be = BatchEngine(engine, method="fft", pipeline=2)
    fft = fft_be.run(reqs, s)
    alone = [fft_be.run([r], s)[0] for r in reqs]
    serial = BatchEngine(engine, method="fft", pipeline=0).run(reqs, s)
    for g, f, a, o in zip(gather, fft, alone, serial):
        assert rel_err(f, g) < TOL
        assert np.array_equal(f, a)
        assert np.array_equal(f, o)   # chunks on a companion stream
    fft_be.close()


def test_batch_engine_large_axis_takes_the_gather(engine, kern):
    """A grid whose frame would exceed 1024 points (an axis of 520 voxels)
    is convolved by the gather inside an FFT-form batch: the same bits as a
    gather engine, next to a request the FFT form takes."""
    from conehead_amd import workloads as W
    from conehead_amd.batch import BatchEngine, DoseRequest
    from conehead_amd.device import Stream
    small = W.uneven(n_phantoms=1, seed=3, max_beams=1, xy=(16, 20),
                     z=(14, 16), n_theta=4, n_phis=(4,))[0]
    big = W.uneven(n_phantoms=1, seed=4, max_beams=1, xy=(16, 20),
                   z=(14, 16), n_theta=4, n_phis=(4,))[0]
    dens = np.ones((520, 12, 10), np.float32)
    req_big = DoseRequest(dens, big.origin, big.spacing, big.beams,
                          big.cones)
    with pytest.raises(RuntimeError, match="transform"):
        engine.plan(dens.shape, big.spacing, big.cones).spectrum()
    req_small = DoseRequest.from_case(small)
    s = Stream()
    fft_be = BatchEngine(engine, method="fft")
    got = fft_be.run([req_small, req_big], s)
    want_big = BatchEngine(engine, method="gather").run([req_big], s)[0]
    want_small = fft_be.run([req_small], s)[0]
    assert np.array_equal(got[1], want_big)
    assert np.array_equal(got[0], want_small)
    fft_be.close()


def test_uneven_workload_fft_equals_gather(engine):
    """Every beam of the uneven workload (32 phantoms from 64^3 to
    200 x 200 x 160, 1-9 beams each, 48-144 cones; frames of every radix,
    one- and two-buffer LDS paths): the FFT form against the gather, each
    beam within 1e-5 of its Dmax.  The gather itself is checked against
    the oracle on small grids (test_gpu_convolve.py)."""
    from conehead_amd import workloads as W
    from conehead_amd.batch import BatchEngine, DoseRequest
    from conehead_amd.device import Stream
    cases = W.uneven()
    reqs = [DoseRequest.from_case(c) for c in cases]
    s = Stream()
    worst = 0.0
    for lo in range(0, len(reqs), 8):          # bounded host memory
        part = reqs[lo:lo + 8]
        fft = BatchEngine(engine, method="fft").run(part, s)
        gat = BatchEngine(engine, method="gather").run(part, s)
        for f, g in zip(fft, gat):
            for b in range(len(f)):
                worst = max(worst, rel_err(f[b], g[b]))
    assert worst < TOL, worst


def test_spectrum_budget(engine, kern):
    """Cached plans' spectra are freed least recently used first, never
    those a batch is about to use, and rebuilt on the next use with the same
    bits."""
    from conehead_amd import engine as Emod
    from conehead_amd.batch import BatchEngine, DoseRequest
    from conehead_amd.device import Stream
    from conehead_amd import workloads as W
    eng = Emod.Engine()
    cs = cone_set(kern, 4, 4)
    rng = np.random.default_rng(12)
    shapes = [(20, 22, 24), (24, 20, 22), (22, 24, 20)]
    plans, Ts, first = [], [], []
    for shp in shapes:
        p = eng.plan(shp, (0.4, 0.4, 0.4), cs)
        T = beam_like(shp, rng)
        plans.append(p)
        Ts.append(T)
        first.append(run_fft(p, T))
    one = plans[0].spectrum_bytes
    assert one > 0
    eng.plan(shapes[0], (0.4, 0.4, 0.4), cs)        # plan 0 most recent
    # LRU order: plan 1, plan 2, plan 0; the budget of one plan 0 frees
    # plan 1, skips plan 2 (kept) and frees plan 0
    held = eng.trim_spectra(one, keep=[plans[2]])
    assert plans[1].spectrum_bytes == 0 and plans[0].spectrum_bytes == 0
    assert plans[2].spectrum_bytes > 0              # kept despite the budget
    assert held == plans[2].spectrum_bytes
    for p, T, d in zip(plans, Ts, first):           # rebuilt, same bits
        assert np.array_equal(run_fft(p, T), d)
    c = W.uneven(n_phantoms=2, seed=5, max_beams=2, xy=(16, 20), z=(14, 16),
                 n_theta=4, n_phis=(4,))
    reqs = [DoseRequest.from_case(x) for x in c]
    s = Stream()
    tight = BatchEngine(eng, method="fft", spectrum_budget=1)
    got = tight.run(reqs, s)
    want = BatchEngine(eng, method="fft").run(reqs, s)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    tight.close()
