"""conehead_amd -- MI355X collapsed-cone dose engine behind conehead's API.

Drop-in entry points (same names and arguments as samuelpeet/conehead):
  conehead_amd.conehead.Conehead().calculate(source, block, phantom, settings)
  conehead_amd.convolve_c.convolve_c(terma, dose, dim, thetas, phis, kernel)
  conehead_amd.dda_3d_c.dda_3d_c(direction, grid_shape, first_voxel, size)
The compute runs in hand-written HIP kernels (libconehead_amd.so, gfx950)
reached through the C-ABI declared in include/conehead_amd.h.
"""
__version__ = "0.1.0"
