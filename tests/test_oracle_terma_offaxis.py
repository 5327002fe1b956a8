"""The FP64 oracle's off-axis TERMA terms (oracle/ch_ref.c, ch_ref_terma)
pinned by closed forms away from the central axis, where OAD > 0:

  OAD   = SAD |p_perp| / p_along  (the isocentre-plane image of the voxel,
          conehead/conehead.py:286-314),
  Psi   = sPri (SAD/r)^2
        + [rInner <= OAD zAnn/SAD <= rOuter] sAnn (SAD-zAnn)^2/(r-zAnn)^2
        + sExp/r_e exp(-kExp r_e) (SAD-zExp)^2/(r-zExp)^2,
          r_e = max(OAD, 2) zExp/SAD          (conehead.py:320-360),
  f_s   = 1/(1 - softRatio OAD) if OAD < softLimit else 1,
  f_h   = 1 + hornRatio OAD                  (conehead.py:641-655),
  T     = sum_i w_i Psi exp(-mu_i f_s d_eff) E_i mu_i f_h * blocked
                                             (conehead.py:364-377).

Voxels are placed exactly on the window and threshold boundaries (the
arithmetic there is exact: 5 * 4 / 100 == 0.2 in IEEE double), so a `<`
written for `<=`, a shifted bound, a flipped sign or a dropped clamp fails
one of these tests.  The field is opened wide (Block.set_square(40) opens
the whole 40 cm raster) so the blocked fraction is 1 up to the rounding of
27 additions of 1/27, which test_oracle_terma pins."""
import numpy as np
import pytest

from conehead_amd import engine as eng_mod
from conehead_amd.block import Block
from conehead_amd.source import Source
from conehead_amd.workloads import SETTINGS

SAD = 100.0


def run(oracle, settings, shape, spacing, origin, gantry=0.0,
        collimator=0.0, side=40.0):
    src = Source("varian_clinac_6MV")
    if gantry:
        src.gantry = gantry
    if collimator:
        src.collimator = collimator
    blk = Block()
    blk.set_square(side)
    e, w, mu = eng_mod.spectrum(settings)
    beam = oracle.make_beam(src.position, src.v_y, src.transform, src.sad,
                            settings, e, w, mu)
    dens = np.ones(shape, np.float32)
    out = oracle.terma(beam, shape, origin, spacing, dens, blk.block_values)
    return out, src, (e, w, mu)


def oad_closed(p, src):
    """Distance from the central axis of the voxel's image on the isocentre
    plane: the ray source -> p scaled to the plane at SAD."""
    s = np.asarray(src.position, np.float64)
    axis = -s / np.linalg.norm(s)            # source -> isocentre
    d = np.asarray(p, np.float64) - s
    along = d @ axis
    perp = d - along * axis
    return SAD * np.linalg.norm(perp) / along


def fluence_closed(oad, r, st):
    """Three-source fluence per unit open fraction at OAD and distance r."""
    fp = st["sPri"] * (SAD / r) ** 2
    r_ann = oad * st["zAnn"] / SAD
    on = st["rInner"] <= r_ann <= st["rOuter"]
    fa = st["sAnn"] * (SAD - st["zAnn"]) ** 2 / (r - st["zAnn"]) ** 2 \
        if on else 0.0
    r_e = max(oad, 2.0) * st["zExp"] / SAD
    fe = st["sExp"] / r_e * np.exp(-st["kExp"] * r_e) * \
        (SAD - st["zExp"]) ** 2 / (r - st["zExp"]) ** 2
    return fp + fa + fe, fa


# A grid whose voxel centres sit on x = -10 + 0.5 i, y = 0.5 j, z = -10 +
# 0.5 k: (30, 0, 20) is at x = 5 on the isocentre plane, so its OAD is
# exactly 5 and OAD zAnn / SAD is exactly rInner = 0.2.
GRID = dict(shape=(41, 5, 41), spacing=[0.5] * 3,
            origin=np.array([-10.0, 0.0, -10.0]))


@pytest.fixture(scope="module")
def narrow(oracle):
    """rOuter = 0.4 (OAD 10) and softLimit = 8 cm, so both ends of the
    annular window and the softening threshold fall inside the grid."""
    st = dict(SETTINGS, rOuter=0.4, softLimit=8.0)
    out, src, spec = run(oracle, st, **GRID)
    return dict(out=out, src=src, spec=spec, st=st)


def pos(idx):
    return GRID["origin"] + 0.5 * np.asarray(idx, np.float64)


def test_oad_closed_form_gantry0(narrow):
    out, src = narrow["out"], narrow["src"]
    for idx in [(30, 0, 20), (40, 0, 20), (20, 0, 20), (24, 3, 27),
                (3, 4, 38), (33, 2, 5), (0, 1, 0)]:
        expect = oad_closed(pos(idx), src)
        got = out["oad"][idx]
        assert got == pytest.approx(expect, rel=1e-12, abs=1e-12), idx


@pytest.mark.parametrize("gantry,collimator", [(30.0, 0.0), (217.0, 15.0)])
def test_oad_closed_form_rotated(oracle, gantry, collimator):
    n, s = 21, 0.7
    shape = (n, n, n)
    origin = np.array([-(n - 1) / 2 * s] * 3)
    out, src, _ = run(oracle, SETTINGS, shape, [s] * 3, origin,
                      gantry=gantry, collimator=collimator)
    rng = np.random.default_rng(5)
    for idx in rng.integers(0, n, size=(12, 3)):
        idx = tuple(int(v) for v in idx)
        p = origin + s * np.asarray(idx)
        assert out["oad"][idx] == pytest.approx(oad_closed(p, src),
                                                rel=1e-11, abs=1e-11), idx


def test_exact_boundaries_are_exact(narrow):
    """The boundary voxels really sit on the boundaries (so the window and
    threshold tests below decide <= against <)."""
    out = narrow["out"]
    assert out["oad"][30, 0, 20] == 5.0
    assert out["oad"][40, 0, 20] == 10.0
    assert out["oad"][36, 0, 20] == 8.0
    assert out["oad"][24, 0, 20] == 2.0
    assert 5.0 * 4.0 / 100.0 == 0.2 and 10.0 * 4.0 / 100.0 == 0.4


def test_annular_window(narrow):
    """The annular source is on for rInner <= OAD zAnn/SAD <= rOuter,
    both ends included, and off just outside them."""
    out, src, st = narrow["out"], narrow["src"], narrow["st"]
    cases = {(29, 0, 20): False,    # OAD 4.5
             (30, 0, 20): True,     # OAD 5: r_ann == rInner
             (31, 0, 20): True,     # OAD 5.5
             (40, 0, 20): True,     # OAD 10: r_ann == rOuter
             (35, 0, 28): True,     # OAD 8.5
             (40, 0, 22): False}    # OAD 10.05
    for idx, on in cases.items():
        p = pos(idx)
        r = np.linalg.norm(p - np.asarray(src.position, np.float64))
        flu, fa = fluence_closed(oad_closed(p, src), r, st)
        assert (fa > 0) == on, idx
        got = out["fluence"][idx] / out["blocked"][idx]
        assert got == pytest.approx(flu, rel=1e-12), idx
        # the annular share is 0.3 % of the fluence: far above the tolerance
        assert abs(fa) < 1e-9 * flu or fa > 1e-3 * flu


def test_exponential_source_clamp(narrow):
    """Below OAD 2 cm the exponential source is evaluated at OAD 2."""
    out, src, st = narrow["out"], narrow["src"], narrow["st"]
    vals = []
    for idx in [(20, 0, 20), (21, 0, 20), (22, 0, 21), (24, 0, 20),
                (26, 0, 20), (20, 2, 27)]:
        p = pos(idx)
        r = np.linalg.norm(p - np.asarray(src.position, np.float64))
        flu, _ = fluence_closed(oad_closed(p, src), r, st)
        got = out["fluence"][idx] / out["blocked"][idx]
        assert got == pytest.approx(flu, rel=1e-12), idx
        vals.append(got)
    # the exponential share alone: below OAD 2 it is the OAD-2 value times
    # the inverse square about zExp (it would be infinite at OAD 0 without
    # the clamp), beyond OAD 2 it falls off
    s_pos = np.asarray(src.position, np.float64)
    r_e2 = 2.0 * st["zExp"] / SAD
    for idx, v in zip([(20, 0, 20), (21, 0, 20), (24, 0, 20), (26, 0, 20)],
                      [vals[0], vals[1], vals[3], vals[4]]):
        r = np.linalg.norm(pos(idx) - s_pos)
        fe = v - st["sPri"] * (SAD / r) ** 2
        at2 = st["sExp"] / r_e2 * np.exp(-st["kExp"] * r_e2) * \
            (SAD - st["zExp"]) ** 2 / (r - st["zExp"]) ** 2
        if out["oad"][idx] <= 2.0:
            assert fe == pytest.approx(at2, rel=1e-10), idx
        else:
            assert fe < at2 * (1 - 1e-3), idx


def terma_closed(out, idx, st, spec):
    e, w, mu = spec
    oad = out["oad"][idx]
    fs = 1.0 / (1.0 - st["softRatio"] * oad) if oad < st["softLimit"] \
        else 1.0
    fh = 1.0 + st["hornRatio"] * oad
    flu = out["fluence"][idx]
    d = out["d_eff"][idx]
    return float(np.sum(w * flu * np.exp(-mu * fs * d) * e * mu * fh)) \
        * out["blocked"][idx], fs, fh


def test_softening_and_horn(narrow):
    """T at off-axis voxels on both sides of softLimit (8 cm here), with
    OAD exactly at the limit taking f_s = 1."""
    out, st, spec = narrow["out"], narrow["st"], narrow["spec"]
    seen = set()
    for idx in [(20, 2, 20), (26, 1, 20), (30, 0, 20), (35, 3, 20),
                (36, 0, 20), (37, 0, 20), (40, 4, 33), (2, 2, 2)]:
        expect, fs, fh = terma_closed(out, idx, st, spec)
        assert out["terma"][idx] == pytest.approx(expect, rel=1e-12), idx
        seen.add(fs == 1.0)
        if out["oad"][idx] > 0:
            assert fh > 1.0
    assert seen == {True, False}
    # at OAD == softLimit exactly, softening is off
    expect, fs, _ = terma_closed(out, (36, 0, 20), st, spec)
    assert fs == 1.0


def test_softening_beyond_default_limit(oracle):
    """The example settings' softLimit of 20 cm: voxels whose image lies
    beyond 20 cm of the axis (still inside the open 40 cm raster) are not
    softened, the rest are."""
    shape, s = (33, 3, 33), 1.0
    origin = np.array([-16.0, 0.0, -16.0])
    out, src, spec = run(oracle, SETTINGS, shape, [s] * 3, origin)
    n_beyond = 0
    for idx in [(31, 0, 31), (30, 1, 29), (16, 0, 16), (6, 2, 26),
                (1, 0, 2), (29, 0, 16)]:
        expect, fs, fh = terma_closed(out, idx, SETTINGS, spec)
        assert out["terma"][idx] == pytest.approx(expect, rel=1e-12), idx
        oad = out["oad"][idx]
        assert oad == pytest.approx(
            oad_closed(origin + s * np.asarray(idx), src), rel=1e-12)
        n_beyond += oad >= 20.0
        if oad > 0:
            assert (fs == 1.0) == (oad >= 20.0)
    assert n_beyond >= 2


def test_horn_raises_terma_off_axis(narrow):
    """Two voxels with the same depth and distance to the source (mirror
    images about the axis) have equal T, and T at a larger OAD on the same
    plane includes the horn's rise: T / (Psi exp-sum) grows as 1 + h OAD."""
    out, st, spec = narrow["out"], narrow["st"], narrow["spec"]
    a, b = (28, 2, 20), (12, 2, 20)          # x = +4 and -4
    assert out["terma"][a] == pytest.approx(out["terma"][b], rel=1e-13)
    e, w, mu = spec
    for idx in [(24, 0, 20), (28, 0, 20), (32, 0, 20)]:
        oad = out["oad"][idx]
        fs = 1.0 / (1.0 - st["softRatio"] * oad)
        base = out["fluence"][idx] * out["blocked"][idx] * float(
            np.sum(w * np.exp(-mu * fs * out["d_eff"][idx]) * e * mu))
        assert out["terma"][idx] / base == pytest.approx(
            1.0 + st["hornRatio"] * oad, rel=1e-12)
