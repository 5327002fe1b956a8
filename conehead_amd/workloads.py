"""Seeded synthetic workloads of the benchmark configurations.

  plain_cpu   50^3 water at 4 mm, 10 x 10 cm field, 12 x 8 = 96 cones
  hetero      128^3 at 2.5 mm, lung slab (depth 5-12 cm, rho 0.25), bone slab
              (depth 3-5 cm, rho 1.85), N(0, 0.02^2) density noise (seed
              1234), 10 x 10 cm field, 12 x 12 = 144 cones
  batched     the hetero phantom, 64 gantry angles U[0, 360) (seed 2024),
              60-leaf-pair 5 mm MLC apertures 2-20 cm, 12 x 8 = 96 cones
  uneven      32 phantoms (seed 7): dims x, y in [64, 200], z in [64, 160],
              voxels 2-5 mm, lung and bone ellipsoids, 1-9 beams, 48-144 cones
Depth is the y axis: the source sits at (0, -100, 0) for gantry 0 and the
beam travels along +y (conehead/source.py:20).  Cone sets merge the 48
EDKnrc polar bins into n_phi equal groups (kernel2.merged_kernel).
"""
import numpy as np

from .block import Block
from .cones import ConeSet, default_thetas
from .kernel2 import PolyenergeticKernel, merged_kernel
from .source import Source

# example.py:27-58
SETTINGS = {
    'sPri': 0.90924,
    'sAnn': 2.887e-3,
    'zAnn': 4.0,
    'rInner': 0.2,
    'rOuter': 1.4,
    'zExp': 12.5,
    'sExp': 8.289e-3,
    'kExp': 0.4816,
    'softRatio': 0.0025,
    'softLimit': 20,
    'hornRatio': 0.0065,
    'fluenceResampling': 3,
    'energy_weights': {
        "0.5": 0.08196, "1.0": 0.12385, "1.5": 0.10605, "2.0": 0.08307,
        "2.5": 0.05881, "3.0": 0.03911, "3.5": 0.02131, "4.0": 0.02426,
        "4.5": 0.0, "5.0": 0.00881, "5.5": 0.0, "6.0": 0.00498,
    },
}

_KERNEL = None


def base_kernel():
    global _KERNEL
    if _KERNEL is None:
        _KERNEL = PolyenergeticKernel()
    return _KERNEL


def cone_set(n_theta, n_phi, semantics="convolve_c"):
    k = merged_kernel(base_kernel(), n_phi)
    return ConeSet(default_thetas(n_theta), np.array(k.keys(), np.float64),
                   k, semantics=semantics)


class Case:
    """One phantom on its dose grid plus beams and a cone set."""

    def __init__(self, name, densities, origin, spacing, beams, cones):
        self.name = name
        self.densities = np.ascontiguousarray(densities, np.float32)
        self.shape = self.densities.shape
        self.origin = np.asarray(origin, np.float64)
        self.spacing = np.asarray(spacing, np.float64)
        self.beams = beams          # list of (Source, Block, settings)
        self.cones = cones

    @property
    def n_voxels(self):
        return int(np.prod(self.shape))

    def voxel_cone_updates(self):
        """Dense work count Nv x C per beam."""
        return self.n_voxels * self.cones.n_cones


def source_at(gantry=0.0, collimator=0.0):
    s = Source("varian_clinac_6MV")
    if gantry:
        s.gantry = gantry
    if collimator:
        s.collimator = collimator
    return s


def square_block(side):
    b = Block()
    b.set_square(side)
    return b


def plain_cpu(n=50, spacing=0.4, n_theta=12, n_phi=8):
    """Homogeneous water, isocentre on the surface (SSD 100)."""
    half = (n - 1) / 2.0 * spacing
    origin = (-half, spacing / 2.0, -half)
    dens = np.ones((n, n, n), np.float32)
    beams = [(source_at(0.0), square_block(10), SETTINGS)]
    return Case("plain_cpu", dens, origin, (spacing,) * 3, beams,
                cone_set(n_theta, n_phi))


def hetero_densities(n=128, spacing=0.25, seed=1234):
    depth = spacing / 2.0 + spacing * np.arange(n)      # voxel centre depth
    rho_y = np.ones(n, np.float32)
    rho_y[(depth >= 3.0) & (depth < 5.0)] = 1.85         # bone slab
    rho_y[(depth >= 5.0) & (depth < 12.0)] = 0.25        # lung slab
    dens = np.broadcast_to(rho_y[None, :, None], (n, n, n)).astype(np.float32)
    rng = np.random.default_rng(seed)
    dens = dens + rng.normal(0.0, 0.02, size=dens.shape).astype(np.float32)
    return np.clip(dens, 0.0, None).astype(np.float32)


def hetero(n=128, spacing=0.25, n_theta=12, n_phi=12):
    half = (n - 1) / 2.0 * spacing
    origin = (-half, spacing / 2.0, -half)
    beams = [(source_at(0.0), square_block(10), SETTINGS)]
    return Case("hetero", hetero_densities(n, spacing), origin,
                (spacing,) * 3, beams, cone_set(n_theta, n_phi))


def mlc_block(rng, n_pairs=60, leaf_mm=5.0):
    """Random MLC aperture: openings of 20-200 mm along leaf travel."""
    bound = -n_pairs * leaf_mm / 2.0 + leaf_mm * np.arange(n_pairs + 1)
    width = rng.uniform(20.0, 200.0)
    centre = rng.uniform(-20.0, 20.0)
    wob = rng.uniform(-0.15, 0.15, size=n_pairs) * width
    a = centre - width / 2.0 + wob
    b = centre + width / 2.0 + wob
    jy = rng.uniform(10.0, 100.0)
    blk = Block()
    blk.set_mlc(bound, a, b, (-120.0, 120.0), (-jy, jy))
    return blk


def batched(n_beams=64, n=128, spacing=0.25, n_theta=12, n_phi=8, seed=2024):
    base = hetero(n, spacing)
    rng = np.random.default_rng(seed)
    gantries = rng.uniform(0.0, 360.0, size=n_beams)
    beams = [(source_at(float(g)), mlc_block(rng), SETTINGS)
             for g in gantries]
    return Case("batched", base.densities, base.origin, base.spacing, beams,
                cone_set(n_theta, n_phi))


def _ellipsoid(shape, spacing, rng):
    c = [rng.uniform(0.25, 0.75) * s for s in shape]
    r = [rng.uniform(0.08, 0.25) * s for s in shape]
    ii = np.ogrid[:shape[0], :shape[1], :shape[2]]
    q = sum(((ii[a] - c[a]) / r[a]) ** 2 for a in range(3))
    return q <= 1.0


def uneven(n_phantoms=32, seed=7, max_beams=9, xy=(64, 200), z=(64, 160),
           spacing=(0.2, 0.5), n_theta=12, n_phis=(4, 6, 8, 12)):
    """Phantoms of differing grids, voxel sizes, beam counts and cone sets
    (dims drawn from `xy` and `z`, inclusive)."""
    rng = np.random.default_rng(seed)
    cases = []
    for p in range(n_phantoms):
        shape = (int(rng.integers(xy[0], xy[1] + 1)),
                 int(rng.integers(xy[0], xy[1] + 1)),
                 int(rng.integers(z[0], z[1] + 1)))
        sp = float(rng.uniform(*spacing))
        dens = np.ones(shape, np.float32)
        for _ in range(int(rng.integers(1, 4))):
            dens[_ellipsoid(shape, sp, rng)] = rng.uniform(0.2, 0.35)
        for _ in range(int(rng.integers(1, 4))):
            dens[_ellipsoid(shape, sp, rng)] = rng.uniform(1.6, 1.9)
        origin = [-(s - 1) / 2.0 * sp for s in shape]
        nb = int(rng.integers(1, max_beams + 1))
        beams = [(source_at(float(rng.uniform(0, 360))),
                  square_block(float(rng.uniform(4.0, 15.0))), SETTINGS)
                 for _ in range(nb)]
        n_phi = int(rng.choice(n_phis))
        cases.append(Case("uneven%02d" % p, dens, origin, (sp,) * 3, beams,
                          cone_set(n_theta, n_phi)))
    return cases
