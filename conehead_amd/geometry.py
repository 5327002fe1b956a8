"""Ray / isocentre-plane intersection (host twin of the device code).

Follows conehead/geometry.py:105-134: the plane passes through the
isocentre (the origin) with normal `plane_normal` (the source's v_y).
"""
import numpy as np


def line_block_plane_collision(ray_start, ray_direction, plane_normal,
                               epsilon=1e-6):
    """Point where the line ray_start + s * ray_direction meets the plane."""
    ray_start = np.asarray(ray_start)
    ray_direction = np.asarray(ray_direction)
    plane_normal = np.asarray(plane_normal)
    ndotu = plane_normal.dot(ray_direction)
    if abs(ndotu) < epsilon:
        raise RuntimeError("no intersection or line is within plane")
    si = -plane_normal.dot(ray_start) / ndotu
    return ray_start + si * ray_direction
