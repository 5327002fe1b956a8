"""Writes edk_6mv.npz from the reference's EDKnrc listings.

    python conehead_amd/data/make_edk.py /path/to/conehead/kernels

Each listing is parsed by KernelMono (conehead/kernel.py:27-53): the
normalised 48 x 24 differential table is stored per energy, with the shared
angle and radius grids."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__)))))

from conehead_amd.kernel import EdkSet  # noqa: E402

if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/kernels"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "edk_6mv.npz")
    EdkSet.from_listings(root).save(out)
    print("wrote", out)
