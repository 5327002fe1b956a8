"""Minimal DICOM RT Plan reader (no pydicom in this environment).

Reads the data elements that Block._set_from_plan needs (conehead/block.py:
87-108) and the beam geometry of each beam, from Part 10 files (128-byte
preamble, "DICM", explicit VR little endian meta group) or raw data sets
(implicit VR little endian, as the reference's sample RP.3DCRT.dcm is
stored).  Sequences and items of defined or undefined length are parsed;
elements whose tags are not in the keyword table are skipped (sequences
among them included).  Nothing in the file is executed.

The result is a tree of Dataset objects with attribute access by DICOM
keyword (BeamSequence, ControlPointSequence, LeafJawPositions, ...), the
interface pydicom offers, so Block._set_from_plan takes it unchanged.
"""
import struct

import numpy as np

# (group, element) -> (keyword, VR) for the elements used here
TAGS = {
    (0x0008, 0x0060): ("Modality", "CS"),
    (0x300A, 0x0002): ("RTPlanLabel", "SH"),
    (0x300A, 0x00B0): ("BeamSequence", "SQ"),
    (0x300A, 0x00B2): ("TreatmentMachineName", "SH"),
    (0x300A, 0x00B4): ("SourceAxisDistance", "DS"),
    (0x300A, 0x00B6): ("BeamLimitingDeviceSequence", "SQ"),
    (0x300A, 0x00B8): ("RTBeamLimitingDeviceType", "CS"),
    (0x300A, 0x00BC): ("NumberOfLeafJawPairs", "IS"),
    (0x300A, 0x00BE): ("LeafPositionBoundaries", "DS"),
    (0x300A, 0x00C0): ("BeamNumber", "IS"),
    (0x300A, 0x00C2): ("BeamName", "LO"),
    (0x300A, 0x00C4): ("BeamType", "CS"),
    (0x300A, 0x00C6): ("RadiationType", "CS"),
    (0x300A, 0x0110): ("NumberOfControlPoints", "IS"),
    (0x300A, 0x0111): ("ControlPointSequence", "SQ"),
    (0x300A, 0x0112): ("ControlPointIndex", "IS"),
    (0x300A, 0x0114): ("NominalBeamEnergy", "DS"),
    (0x300A, 0x011A): ("BeamLimitingDevicePositionSequence", "SQ"),
    (0x300A, 0x011C): ("LeafJawPositions", "DS"),
    (0x300A, 0x011E): ("GantryAngle", "DS"),
    (0x300A, 0x0120): ("BeamLimitingDeviceAngle", "DS"),
    (0x300A, 0x0122): ("PatientSupportAngle", "DS"),
    (0x300A, 0x012C): ("IsocenterPosition", "DS"),
    (0x300A, 0x0134): ("CumulativeMetersetWeight", "DS"),
}
_KEYWORDS = {kw for kw, _ in TAGS.values()}
_ITEM, _ITEM_END, _SEQ_END = (0xFFFE, 0xE000), (0xFFFE, 0xE00D), \
    (0xFFFE, 0xE0DD)
_LONG_VRS = {b"OB", b"OW", b"OF", b"OD", b"OL", b"SQ", b"UN", b"UT", b"UC",
             b"UR", b"SV", b"UV"}


class Dataset:
    """Attribute access by keyword; missing keywords raise AttributeError."""

    def __init__(self):
        self._items = {}

    def __getattr__(self, name):
        items = self.__dict__.get("_items", {})
        if name in items:
            return items[name]
        raise AttributeError(name)

    def __contains__(self, name):
        return name in self._items

    def keys(self):
        return list(self._items)


def _value(raw, vr):
    text = raw.decode("latin-1").rstrip("\x00 ").strip()
    if vr == "DS":
        vals = [float(v) for v in text.split("\\") if v.strip()]
        return vals[0] if len(vals) == 1 else vals
    if vr == "IS":
        vals = [int(v) for v in text.split("\\") if v.strip()]
        return vals[0] if len(vals) == 1 else vals
    return text


class _Reader:
    def __init__(self, data, explicit):
        self.b = data
        self.explicit = explicit

    def header(self, off):
        g, e = struct.unpack_from("<HH", self.b, off)
        if (g, e) in (_ITEM, _ITEM_END, _SEQ_END) or not self.explicit:
            (n,) = struct.unpack_from("<I", self.b, off + 4)
            return (g, e), None, n, off + 8
        vr = self.b[off + 4:off + 6]
        if vr in _LONG_VRS:
            (n,) = struct.unpack_from("<I", self.b, off + 8)
            return (g, e), vr.decode(), n, off + 12
        (n,) = struct.unpack_from("<H", self.b, off + 6)
        return (g, e), vr.decode(), n, off + 8

    def dataset(self, off, end):
        """Elements from off to end (or to an item delimiter when end is
        None); returns (Dataset, next offset)."""
        ds = Dataset()
        while (end is None or off < end) and off + 8 <= len(self.b):
            tag, vr, n, voff = self.header(off)
            if tag == _ITEM_END:
                return ds, voff
            known = TAGS.get(tag)
            is_seq = (vr == "SQ") or (known is not None and known[1] == "SQ")
            if n == 0xFFFFFFFF or is_seq:
                items, off = self.sequence(voff, n)
                if known:
                    ds._items[known[0]] = items
                continue
            if known:
                ds._items[known[0]] = _value(self.b[voff:voff + n], known[1])
            off = voff + n
        return ds, off

    def sequence(self, off, n):
        end = None if n == 0xFFFFFFFF else off + n
        items = []
        while (end is None or off < end) and off + 8 <= len(self.b):
            tag, _, m, voff = self.header(off)
            if tag == _SEQ_END:
                return items, voff
            if tag != _ITEM:
                raise ValueError("malformed sequence at byte %d" % off)
            if m == 0xFFFFFFFF:
                item, off = self.dataset(voff, None)
            else:
                item, _ = self.dataset(voff, voff + m)
                off = voff + m
            items.append(item)
        return items, off if end is None else end


def read_file(path):
    """Parse a DICOM file; returns the top-level Dataset."""
    with open(path, "rb") as f:
        data = f.read()
    return read_bytes(data)


def read_bytes(data):
    if len(data) >= 132 and data[128:132] == b"DICM":
        # Part 10: explicit VR meta group, then the transfer syntax's body
        r = _Reader(data, explicit=True)
        off = 132
        ts = None
        while off + 8 <= len(data):
            (g, _) = struct.unpack_from("<HH", data, off)
            if g != 0x0002:
                break
            tag, vr, n, voff = r.header(off)
            if tag == (0x0002, 0x0010):
                ts = data[voff:voff + n].decode("latin-1").strip("\x00 ")
            off = voff + n
        explicit = ts != "1.2.840.10008.1.2"   # implicit VR little endian
        if ts not in (None, "1.2.840.10008.1.2", "1.2.840.10008.1.2.1"):
            raise ValueError("unsupported transfer syntax %s" % ts)
        ds, _ = _Reader(data, explicit).dataset(off, len(data))
        return ds
    # raw data set: implicit VR unless the first element carries a VR
    explicit = len(data) >= 6 and data[4:6].isalpha() and \
        data[4:6].isupper()
    ds, _ = _Reader(data, explicit).dataset(0, len(data))
    return ds


def plan_beams(plan):
    """Per static beam of an RT plan: (beam dataset, gantry angle,
    collimator angle, isocentre in mm) from its first control point."""
    out = []
    for beam in plan.BeamSequence:
        cp = beam.ControlPointSequence[0]
        gantry = float(getattr(cp, "GantryAngle", 0.0))
        coll = float(getattr(cp, "BeamLimitingDeviceAngle", 0.0))
        iso = np.asarray(getattr(cp, "IsocenterPosition", [0.0, 0.0, 0.0]),
                         np.float64)
        out.append((beam, gantry, coll, iso))
    return out


def single_beam_plan(plan, index):
    """A view of `plan` holding only beam `index`, for
    Block._set_from_plan, which reads the jaws and leaves of the last beam
    of the plan it is given (conehead/block.py:89-108)."""
    view = Dataset()
    view._items = dict(plan._items)
    view._items["BeamSequence"] = [plan.BeamSequence[index]]
    return view
