"""Build libconehead_amd.so from conehead_amd/csrc with hipcc for gfx950.

Each .hip file is compiled on its own so the geometry files (exact FP64
traversal, positions, plane hits, kernel indices) get -ffp-contract=off
while the gather kernels keep fused multiply-adds.  Packed-FP32 instructions
are disabled for every file.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libconehead_amd.so")
BUILD = os.path.join(HERE, "_build")

ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    # a non-void function that falls off its end is undefined behaviour the
    # device compiler turns into unreachable code: refuse to build it
    "-Werror=return-type",
]
# Files whose arithmetic must round every operation like the FP64 oracle.
EXACT = {"dda.hip", "plan.hip", "terma.hip", "raster.hip", "scatter.hip",
         "resample.hip", "physics.hip"}
SOURCES = ["runtime.hip", "dda.hip", "plan.hip", "ccc.hip", "terma.hip",
           "raster.hip", "scatter.hip", "resample.hip", "physics.hip",
           "spectral.hip"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc",
              shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    flags = list(COMMON)
    flags.append("-ffp-contract=off" if src in EXACT else "-ffp-contract=fast")
    cmd = [hipcc()] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" %
                           (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False, jobs=6):
    """Compile every source and link the shared library (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] \
        + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
