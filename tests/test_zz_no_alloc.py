"""ch_terma_noalloc, the allocation-free TERMA entry point, against
ch_terma through the C-ABI: the same kernels on a fixed scratch array, so
TERMA and the support box must be bitwise those of ch_terma.  Two calls in
a row, the second on a smaller grid (the array keeps the first call's
contents); a grid beyond the array's size is refused with a clear error; a
beam's geometry error reaches its own flag."""
import ctypes

import numpy as np
import pytest

from conehead_amd import _lib
from conehead_amd import engine as E
from conehead_amd import workloads as W
from conehead_amd.device import DeviceArray, Stream

pytestmark = pytest.mark.gpu


def beams_for(engine, case, n, d_rho, s):
    """n beams of `case` (repeating its beam list), each with its own TERMA
    grid, support box and zeroed geometry flag."""
    out = []
    for i in range(n):
        src, blk, st = case.beams[i % len(case.beams)]
        t = DeviceArray(case.shape, zero=True, stream=s)
        box = DeviceArray((8,), np.int32, zero=True, stream=s)
        flag = DeviceArray((1,), np.int32, zero=True, stream=s)
        b = E.beam_struct(src, st, engine.aperture(blk, s), d_rho, t, box)
        b.d_status = flag.ptr
        out.append((b, t, box, flag))
    return out


def call(name, beams, case, s):
    arr = (_lib.Beam * len(beams))(*[b for b, _, _, _ in beams])
    shp = (ctypes.c_int32 * 3)(*case.shape)
    org = (ctypes.c_double * 3)(*[float(v) for v in case.origin])
    spc = (ctypes.c_double * 3)(*[float(v) for v in case.spacing])
    return getattr(_lib.lib(), name)(len(beams), arr, shp, org, spc, None,
                                     s.handle)


def outputs(beams, s):
    s.sync()
    return [(t.download(stream=s), box.download(stream=s)[:6],
             int(flag.download(stream=s)[0])) for _, t, box, flag in beams]


def test_noalloc_equals_terma_twice(engine):
    """The batched workload's MLC beams on a 96^3 phantom, then the
    heterogeneous beam on 40^3: the second call is smaller."""
    s = Stream()
    big = W.batched(n_beams=4, n=96, spacing=0.3)
    small = W.hetero(n=40, spacing=0.6)
    for c in (big, small):
        d_rho = DeviceArray.from_host(c.densities, stream=s)
        ref = beams_for(engine, c, 3, d_rho, s)
        got = beams_for(engine, c, 3, d_rho, s)
        _lib.check(call("ch_terma", ref, c, s))
        _lib.check(call("ch_terma_noalloc", got, c, s))
        r, g = outputs(ref, s), outputs(got, s)
        for (rt, rb, rf), (gt, gb, gf) in zip(r, g):
            assert rt.max() > 0
            assert np.array_equal(rt.view(np.uint32), gt.view(np.uint32))
            assert np.array_equal(rb, gb)
            assert rf == 0 and gf == 0


def test_noalloc_refuses_and_flags(engine):
    s = Stream()
    c = W.hetero(n=24, spacing=0.5)
    d_rho = DeviceArray.from_host(c.densities, stream=s)
    (b, t, box, flag), = beams_for(engine, c, 1, d_rho, s)
    # a grid whose brick lists exceed the fixed array: refused before any
    # buffer is touched
    shp = (ctypes.c_int32 * 3)(600, 600, 500)
    org = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    spc = (ctypes.c_double * 3)(0.1, 0.1, 0.1)
    arr = (_lib.Beam * 1)(b)
    st = _lib.lib().ch_terma_noalloc(1, arr, shp, org, spc, None, s.handle)
    assert st == _lib.CH_ERR_ARG
    assert "scratch" in _lib.last_error()
    # a beam without its own flag is refused too
    nb = E.beam_struct(c.beams[0][0], c.beams[0][2],
                       engine.aperture(c.beams[0][1], s), d_rho, t, None)
    arr = (_lib.Beam * 1)(nb)
    shp = (ctypes.c_int32 * 3)(*c.shape)
    org = (ctypes.c_double * 3)(*[float(v) for v in c.origin])
    spc = (ctypes.c_double * 3)(*[float(v) for v in c.spacing])
    st = _lib.lib().ch_terma_noalloc(1, arr, shp, org, spc, None, s.handle)
    assert st == _lib.CH_ERR_ARG and "d_status" in _lib.last_error()
    # the source plane cuts this grid (gantry 0, y = -100 cm): the beam's
    # own flag is set, as ch_terma's stream flag would be
    # (every device buffer a beam points to stays referenced until the
    # stream has synchronised)
    rho8 = DeviceArray.from_host(np.ones((8, 8, 8), np.float32), stream=s)
    t8 = DeviceArray((8, 8, 8))
    flag8 = DeviceArray((1,), np.int32, zero=True, stream=s)
    bb = E.beam_struct(W.source_at(0.0), W.SETTINGS,
                       engine.aperture(W.square_block(10.0), s), rho8, t8,
                       None)
    bb.d_status = flag8.ptr
    shp = (ctypes.c_int32 * 3)(8, 8, 8)
    org = (ctypes.c_double * 3)(2.0, -102.0, -2.0)
    spc = (ctypes.c_double * 3)(0.5, 0.5, 0.5)
    arr = (_lib.Beam * 1)(bb)
    _lib.check(_lib.lib().ch_terma_noalloc(1, arr, shp, org, spc, None,
                                           s.handle))
    s.sync()
    assert int(flag8.download(stream=s)[0]) == 1
    del rho8, t8
