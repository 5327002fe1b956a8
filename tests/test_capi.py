"""The C-ABI library loads and exports every function include/*.h declares
(no GPU needed: only symbol lookups and the two pure functions)."""
import glob
import os
import re

import pytest

from conehead_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ch_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared()
    for must in ("ch_dda_3d", "ch_dda_trace_batch", "ch_plan_create",
                 "ch_convolve", "ch_convolve_batch", "ch_terma",
                 "ch_support_box", "ch_raster_mlc"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libconehead_amd.so not built")
    L = _lib.load()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared()) == set(_lib.SIGNATURES)


def test_version_and_error_without_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libconehead_amd.so not built")
    L = _lib.load()
    assert L.ch_version() >= 10000
    assert isinstance(_lib.last_error(), str)


def test_library_is_gfx950_only():
    """The shared library carries device code for gfx950 (checked through
    the offload bundle's target names printed by the compiler driver)."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libconehead_amd.so not built")
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    assert b"gfx942" not in data and b"gfx90a" not in data
