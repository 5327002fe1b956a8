"""Writes conehead_amd/data/edk_6mv.npz from the reference's EDKnrc listings.

    python tests/golden/make_edk.py <conehead checkout>/kernels

Each listing is parsed by KernelMono (conehead/kernel.py:27-53): the
normalised 48 x 24 differential table is stored per energy, with the shared
angle and radius grids.  The listings directory is a required argument."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from conehead_amd.kernel import EdkSet  # noqa: E402

if __name__ == "__main__":
    if len(sys.argv) != 2:
        sys.exit("usage: python tests/golden/make_edk.py <kernels directory>")
    out = os.path.join(ROOT, "conehead_amd", "data", "edk_6mv.npz")
    EdkSet.from_listings(sys.argv[1]).save(out)
    print("wrote", out)
