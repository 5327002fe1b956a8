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


def test_struct_mirrors_match_the_header(tmp_path):
    """The ctypes mirrors have the C structs' sizes and field offsets, read
    from the header by the C compiler: a field added to one mirror and not
    the other shifts every array element after the first."""
    import ctypes
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"ch_beam": _lib.Beam, "ch_conv_job": _lib.ConvJob,
               "ch_aperture": _lib.Aperture,
               "ch_conv_group": _lib.ConvGroup}
    lines = ['#include <stdio.h>', '#include <stddef.h>',
             '#include "conehead_amd.h"', 'int main(void) {']
    for cname, py in mirrors.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));'
                     % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));'
                         % (cname, f, cname, f))
    lines += ['return 0;', '}']
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2])
           for line in out if line.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
