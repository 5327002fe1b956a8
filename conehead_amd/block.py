"""Beam aperture (jaws and multi-leaf collimator) as a transmission raster.

Follows conehead/block.py:7-221: a 4000 x 4000 FP32 raster at 0.1 mm pitch
over -20..20 cm in the isocentre plane (row ~ source x, column ~ source z),
looked up by `transmission` with floor(100 p) + 2000 - 1.

The raster is described by a recipe (zeros, squares, a leaf table) and
built on the GPU by the engine (ch_raster_square / ch_raster_mlc), so a
beam's aperture never crosses PCIe.  `block_values` materialises the same
raster on the host for callers of the reference API; once a caller has taken
that array (and may have edited it), the engine uploads it instead.
"""
import itertools

import numpy as np

RASTER = 4000
_UIDS = itertools.count()


def _leaf_area(width, height):
    """Leaf transmission block of conehead/block.py:155-164."""
    area = np.ones((width, height), dtype=np.float32)
    area[0, :] = 0.20
    area[1, :] = 0.50
    area[2, :] = 0.75
    area[-1, :] = 0.20
    area[-2, :] = 0.50
    area[-3, :] = 0.75
    area[:, -15:] *= np.linspace(1.0, 0.0, 15)
    return 1 - area


def mlc_raster(bound, ends_a, ends_b, jaw_x, jaw_y):
    """Host raster of a leaf table in integer tenths of a millimetre.

    Leaf bank of conehead/block.py:130-196, centred on raster row 2000 (the
    reference's bank raster has mlc_width rows starting at the first
    boundary); rows outside the bank are blocked."""
    bound = np.asarray(bound, dtype=np.int64)
    W = int(abs(bound[0] - bound[-1]))
    bank = np.ones((W, RASTER), dtype=np.float32)
    shift = int(abs(bound[0]))
    n_pairs = len(bound) - 1
    for n in range(n_pairs):
        rmin = int(bound[n]) + shift
        w = int(abs(bound[n + 1] - bound[n]))
        cmax = int(2000 + ends_a[n])
        bank[rmin:rmin + w, 0:cmax] = _leaf_area(w, cmax)
    for n in range(n_pairs):
        rmin = int(bound[n]) + shift
        w = int(abs(bound[n + 1] - bound[n]))
        cmin = int(2000 + ends_b[n])
        bank[rmin:rmin + w, cmin:3999] = np.fliplr(np.flipud(
            _leaf_area(w, 3999 - cmin)))
    bank[:, :max(0, int(2000 + jaw_x[0]))] = 0.0
    bank[:, max(0, int(2000 + jaw_x[1])):] = 0.0
    bank[:max(0, int(W / 2 + jaw_y[0])), :] = 0.0
    bank[max(0, int(W / 2 + jaw_y[1])):, :] = 0.0
    raster = np.zeros((RASTER, RASTER), dtype=np.float32)
    r0 = 2000 - W // 2
    raster[r0:r0 + W, :] = bank
    return raster


def square_bounds(length):
    """Row/column range of set_square (conehead/block.py:73-76)."""
    xnum = np.int32(RASTER)
    xres = xnum / (np.float32(20) - np.float32(-20))
    lo = int((xnum / 2) - (length / 2) * xres)
    hi = int((xnum / 2) + (length / 2) * xres)
    return lo, hi


class Block:

    def __init__(self, rotation=np.array([0, 0, 0], dtype=np.float32),
                 plan=None):
        self.rotation = rotation
        self.xmin = np.float32(-20)
        self.xmax = np.float32(20)
        self.xnum = np.int32(RASTER)
        self.xres = self.xnum / (self.xmax - self.xmin)
        self.ymin = np.float32(-20)
        self.ymax = np.float32(20)
        self.ynum = np.int32(RASTER)
        self.yres = self.ynum / (self.ymax - self.ymin)
        self._recipe = [("zeros",)]
        self._uid = next(_UIDS)
        self._values = None
        self._host_owned = False
        self._version = 0
        if plan:
            self._set_from_plan(plan)

    # -- recipe ----------------------------------------------------------
    def recipe(self):
        """Tuple describing how the raster is built, or ("host", version)
        when the host array has been handed out."""
        if self._host_owned:
            return ("host", self._uid, self._version)
        return tuple(self._recipe)

    @property
    def block_values(self):
        if self._values is None:
            self._values = self._materialise()
        self._host_owned = True
        self._version += 1
        return self._values

    @block_values.setter
    def block_values(self, values):
        self._values = np.ascontiguousarray(values, dtype=np.float32)
        self._host_owned = True
        self._version += 1

    def _materialise(self):
        r = np.zeros((RASTER, RASTER), dtype=np.float32)
        for op in self._recipe:
            if op[0] == "zeros":
                r[:] = 0
            elif op[0] == "square":
                lo, hi = square_bounds(op[1])
                r[lo:hi, lo:hi] = np.float32(1)
            elif op[0] == "mlc":
                r = mlc_raster(*op[1:])
        return r

    def set_square(self, length):
        """Open a centred square of side `length` cm
        (conehead/block.py:64-85)."""
        if self._host_owned:
            lo, hi = square_bounds(length)
            self._values[lo:hi, lo:hi] = np.float32(1)
            self._version += 1
        else:
            self._recipe.append(("square", float(length)))
            self._values = None

    def set_mlc(self, boundaries_mm, bank_a_mm, bank_b_mm, jaw_x_mm,
                jaw_y_mm):
        """Leaf bank and jaws from positions in mm (the DICOM units),
        converted to tenths of a mm with floor(10 x) as
        conehead/block.py:117-120 does."""
        bound = np.floor(np.asarray(boundaries_mm, dtype=np.float32) * 10)
        a = np.floor(np.asarray(bank_a_mm, dtype=np.float64) * 10)
        b = np.floor(np.asarray(bank_b_mm, dtype=np.float64) * 10)
        jx = np.floor(np.asarray(jaw_x_mm, dtype=np.float64) * 10)
        jy = np.floor(np.asarray(jaw_y_mm, dtype=np.float64) * 10)
        if len(a) != len(bound) - 1 or len(b) != len(bound) - 1:
            raise ValueError("need one A and one B tip per leaf pair")
        table = (tuple(int(v) for v in bound), tuple(int(v) for v in a),
                 tuple(int(v) for v in b), tuple(int(v) for v in jx),
                 tuple(int(v) for v in jy))
        self._recipe = [("mlc",) + table]
        self._values = None
        self._host_owned = False

    def _set_from_plan(self, plan):
        """Jaws and MLC of a DICOM RT plan (duck-typed pydicom dataset), as
        conehead/block.py:87-108 reads them."""
        mlc_boundaries = jaw_x = jaw_y = mlc_ends = None
        for beam in plan.BeamSequence:
            if beam.BeamType != 'STATIC':
                raise NotImplementedError(
                    "Only beams with type 'STATIC' are currently implemented.")
            for coll in beam.BeamLimitingDeviceSequence:
                if coll.RTBeamLimitingDeviceType == 'MLCX':
                    mlc_boundaries = list(coll.LeafPositionBoundaries)
            cp = beam.ControlPointSequence[0]
            for coll in cp.BeamLimitingDevicePositionSequence:
                if coll.RTBeamLimitingDeviceType in ('X', 'ASYMX'):
                    jaw_x = list(coll.LeafJawPositions)
                if coll.RTBeamLimitingDeviceType in ('Y', 'ASYMY'):
                    jaw_y = list(coll.LeafJawPositions)
                elif coll.RTBeamLimitingDeviceType == 'MLCX':
                    mlc_ends = list(coll.LeafJawPositions)
        half = len(mlc_ends) // 2
        self.set_mlc(mlc_boundaries, mlc_ends[:half], mlc_ends[half:], jaw_x,
                     jaw_y)

    # -- host lookups ----------------------------------------------------
    def transmission(self, position):
        """Raster value at an isocentre-plane position in cm
        (conehead/block.py:36-62)."""
        p = np.floor(np.asarray(position) * np.float32(100)) + np.float32(2000)
        for coord in p:
            if coord < 0 or coord > 3999:
                return np.float32(0)
        if self._values is None:
            self._values = self._materialise()
        return self._values[int(p[0]) - 1, int(p[1]) - 1]

    @property
    def block_locations(self):
        return np.mgrid[self.xmin:self.xmax:self.xnum * 1j,
                        self.ymin:self.ymax:self.ynum * 1j].astype(np.float32)
