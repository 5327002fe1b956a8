"""Multi-beam TERMA in one grid of workgroups, the device-resident beam and
convolution sets, and a whole batch captured in one graph.

ch_terma runs beams with distinct TERMA grids in one launch; each brick is
computed by the same code as in a single-beam launch, so every beam's TERMA,
support box and geometry flag must be bitwise those of running it alone.
Beams adding into one grid must give the sequential sum bit for bit.  A
captured batch (BatchGraph) must reproduce the eager BatchEngine bitwise on
every replay, with eager work between replays, and after update() with new
beams."""
import numpy as np
import pytest

from conehead_amd import _lib
from conehead_amd import engine as E
from conehead_amd import workloads as W
from conehead_amd.batch import BatchEngine, BatchGraph, DoseRequest
from conehead_amd.device import DeviceArray, Stream

pytestmark = pytest.mark.gpu


def mlc_beams(n, seed):
    rng = np.random.default_rng(seed)
    return [(W.source_at(float(g)), W.mlc_block(rng), W.SETTINGS)
            for g in rng.uniform(0.0, 360.0, size=n)]


class Held(list):
    """The device objects the beam structs point into, kept alive."""


def setup_beams(engine, s, beams, n=48, spacing=0.5, boxes=True):
    case = W.hetero(n=n, spacing=spacing, n_theta=6, n_phi=4)
    rho = DeviceArray.from_host(case.densities, stream=s)
    ts = [DeviceArray(case.shape) for _ in beams]
    bx = [DeviceArray((8,), np.int32) if boxes else None for _ in beams]
    aps = [engine.aperture(blk, s) for _, blk, _ in beams]
    structs = Held(E.beam_struct(src, st, ap, rho, t, b)
                   for (src, _, st), ap, t, b in zip(beams, aps, ts, bx))
    structs.keep = (rho, ts, bx, aps)
    return case, rho, ts, bx, structs


def test_multi_beam_terma_equals_single(engine):
    s = Stream()
    beams = mlc_beams(7, 3) + [(W.source_at(33.0), W.square_block(12),
                                W.SETTINGS)]
    case, rho, ts, bx, structs = setup_beams(engine, s, beams)
    E.terma(structs, case.shape, case.origin, case.spacing, stream=s)
    multi = [t.download(stream=s) for t in ts]
    mbox = [b.download(stream=s)[:6] for b in bx]
    for i, b in enumerate(structs):
        ts[i].zero(s)
        E.terma([b], case.shape, case.origin, case.spacing, stream=s)
        assert np.array_equal(ts[i].download(stream=s), multi[i]), i
        assert np.array_equal(bx[i].download(stream=s)[:6], mbox[i]), i
    assert all(m.max() > 0 for m in multi)


def test_accumulating_beams_stay_ordered(engine):
    """Beams summed into one grid (accumulate) split the launches; the sum
    is bitwise the sequential one."""
    s = Stream()
    beams = mlc_beams(5, 4)
    case, rho, ts, bx, structs = setup_beams(engine, s, beams, boxes=False)
    acc = DeviceArray(case.shape)
    for k, b in enumerate(structs):
        b.d_terma = acc.ptr
        b.accumulate = 0 if k == 0 else 1
    # interleave with beams of their own grids: one launch holds both kinds
    own = mlc_beams(3, 5)
    _, _, ts2, _, structs2 = setup_beams(engine, s, own, boxes=False)
    mixed = [structs[0], structs2[0], structs[1], structs2[1], structs[2],
             structs2[2], structs[3], structs[4]]
    E.terma(mixed, case.shape, case.origin, case.spacing, stream=s)
    got = acc.download(stream=s)
    acc.zero(s)
    for b in structs:
        E.terma([b], case.shape, case.origin, case.spacing, stream=s)
    seq = acc.download(stream=s)
    assert np.array_equal(got, seq)
    one = ts2[1].download(stream=s)
    ts2[1].zero(s)
    E.terma([structs2[1]], case.shape, case.origin, case.spacing, stream=s)
    assert np.array_equal(ts2[1].download(stream=s), one)


def test_beamset_replays_and_updates(engine):
    s = Stream()
    beams = mlc_beams(6, 6)
    case, rho, ts, bx, structs = setup_beams(engine, s, beams)
    E.terma(structs, case.shape, case.origin, case.spacing, stream=s)
    want = [t.download(stream=s) for t in ts]
    wbox = [b.download(stream=s)[:6] for b in bx]
    bs = E.BeamSet(structs, case.shape, case.origin, case.spacing)
    for t in ts:
        t.zero(s)
    bs.run(s)                                    # eager
    assert all(np.array_equal(t.download(stream=s), w)
               for t, w in zip(ts, want))
    with E.Graph.capture(s) as cap:
        bs.run(s)
    g = cap.graph
    for rep in range(3):
        for t in ts:
            t.zero(s)
        for b in bx:
            b.zero(s)
        g.launch(s)
        s.sync()
        for i, (t, b) in enumerate(zip(ts, bx)):
            assert np.array_equal(t.download(stream=s), want[i]), (rep, i)
            assert np.array_equal(b.download(stream=s)[:6], wbox[i])
    # new geometry, same outputs: the replay gives the new beams' TERMA
    beams2 = mlc_beams(6, 7)
    aps = [engine.aperture(blk, s) for _, blk, _ in beams2]
    s2 = [E.beam_struct(src, st, ap, rho, t, b)
          for (src, _, st), ap, t, b in zip(beams2, aps, ts, bx)]
    bs.update(s2, s, keep=aps)
    g.launch(s)
    s.sync()
    got2 = [t.download(stream=s) for t in ts]
    E.terma(s2, case.shape, case.origin, case.spacing, stream=s)
    for i, t in enumerate(ts):
        assert np.array_equal(t.download(stream=s), got2[i]), i
    assert not np.array_equal(got2[0], want[0])
    assert not bs.flags(s).any()
    # destroying the set while the graph lives defers it
    bs.close()
    g.launch(s)
    s.sync()
    assert np.array_equal(ts[0].download(stream=s), got2[0])
    g.close()


def test_beamset_flags_per_beam(engine):
    """A ray in the source plane flags only its own beam."""
    s = Stream()
    case = W.hetero(n=24, spacing=0.5, n_theta=6, n_phi=4)
    good = W.source_at(0.0)
    rho = DeviceArray.from_host(case.densities, stream=s)
    ts = [DeviceArray(case.shape) for _ in range(3)]
    ap = engine.aperture(W.square_block(10), s)
    structs = [E.beam_struct(good, W.SETTINGS, ap, rho, t, None) for t in ts]
    bad_origin = (2.0, -102.0, -2.0)   # the grid crosses the source plane
    bs = E.BeamSet(structs, case.shape, case.origin, case.spacing)
    bad = E.BeamSet(structs[:1], case.shape, bad_origin, case.spacing)
    bs.run(s)
    bad.run(s)
    assert list(bs.flags(s)) == [0, 0, 0]
    assert list(bad.flags(s)) == [1]
    assert list(bad.flags(s)) == [0]             # cleared by the read


def requests(seed, n_beams=6):
    base = W.hetero(n=40, spacing=0.6, n_theta=6, n_phi=4)
    reqs = [DoseRequest(base.densities, base.origin, base.spacing,
                        mlc_beams(n_beams, seed), base.cones)]
    small = W.uneven(n_phantoms=2, seed=seed, max_beams=3, xy=(16, 30),
                     z=(14, 32), spacing=(0.4, 0.8), n_theta=6, n_phis=(4,))
    reqs += [DoseRequest.from_case(c) for c in small]
    reqs.append(DoseRequest.from_case(small[0], per_beam=False))
    return reqs


def test_batch_graph_matches_eager(engine):
    reqs = requests(21)
    be = BatchEngine(engine, method="fft")
    s = Stream()
    want = be.run(reqs, s)
    bg = BatchGraph(reqs, engine)
    try:
        for rep in range(3):
            bg.launch()
            got = bg.results()
            for i, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), (rep, i)
            be.run(reqs[1:2], s)               # eager work between replays
        # new beams and densities of the same structure
        reqs2 = requests(22)
        reqs2 = [DoseRequest(r2.densities if r2.shape == r.shape
                             else r.densities, r.origin, r.spacing,
                             (r2.beams + r.beams)[:len(r.beams)], r.cones,
                             r.per_beam)
                 for r, r2 in zip(reqs, reqs2)]
        bg.update(reqs2)
        bg.launch()
        got = bg.results()
        want2 = be.run(reqs2, s)
        for i, (a, b) in enumerate(zip(got, want2)):
            assert np.array_equal(a, b), i
        assert not np.array_equal(got[0], want[0])
        with pytest.raises(ValueError):
            bg.update(reqs2[:-1])
    finally:
        bg.close()


def test_capture_refuses_staged_tables(engine):
    """Eager-only entry points refuse capture instead of recording a
    host-staged table."""
    s = Stream()
    case, rho, ts, bx, structs = setup_beams(engine, s, mlc_beams(2, 8),
                                             n=24)
    plan = engine.plan(case.shape, case.spacing, case.cones)
    d = DeviceArray(case.shape)
    E.convolve_fft_batch([(plan, ts[0], d, None, False)], stream=s)
    with pytest.raises(_lib.CHError):
        with E.Graph.capture(s):
            E.convolve_fft_batch([(plan, ts[0], d, None, False)], stream=s)
    s.sync()
