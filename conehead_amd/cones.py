"""Cone sets: directions and kernel rows of the collapsed-cone convolution.

Directions are computed on the host with the C library's sin/cos (Python's
math module calls libm, as the reference's Cython does through libc.math),
in the same operation order, so the discretised directions are identical to
the ones convolve_c traces.
"""
import math

import numpy as np

PI_CONVOLVE_C = 3.14159265359  # conehead/convolve_c.pyx:48


def convolve_c_direction(theta, phi):
    """Discretised cone direction (conehead/convolve_c.pyx:95-117).

    z is the polar axis; "normalise" multiplies by 1/|d|^2 and the result is
    rounded up to 1e-5 (np.ceil keeps the sign of -0.0 like C's ceil)."""
    pi = PI_CONVOLVE_C
    theta_rad = float(theta) * pi / 180.0
    phi_rad = float(phi) * pi / 180.0
    c_t = math.cos(theta_rad)
    s_t = math.sin(theta_rad)
    c_p = math.cos(phi_rad)
    s_p = math.sin(phi_rad)
    d = [c_t * s_p, s_t * s_p, c_p]
    n = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
    n_inv = 1.0 / n
    d = [d[0] * n_inv, d[1] * n_inv, d[2] * n_inv]
    return np.array([float(np.ceil(np.float64(v * 100000))) * 0.00001
                     for v in d], dtype=np.float64)


def cuda_dose_direction(theta, phi):
    """Cone direction of cuda_dose (conehead/conehead.py:414-429): y is the
    polar axis, normalised by |d| with no discretisation (FP64 here; the
    reference evaluates it in FP32)."""
    theta_rad = float(theta) * math.pi / 180.0
    phi_rad = float(phi) * math.pi / 180.0
    c_t, s_t = math.cos(theta_rad), math.sin(theta_rad)
    c_p, s_p = math.cos(phi_rad), math.sin(phi_rad)
    d = [c_t * s_p, c_p, s_t * s_p]
    n = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    return np.array([d[0] / n, d[1] / n, d[2] / n], dtype=np.float64)


def kernel_rows(kernel, phis):
    """Rows kernel.cumulative["{:.2f}".format(phi)] (convolve_c.pyx:257-274).

    Raises KeyError for a missing angle (the reference prints the KeyError,
    ignores it and later exits the process)."""
    rows = []
    for p in phis:
        key = "{:.2f}".format(float(p))
        if key not in kernel.cumulative:
            raise KeyError(key)
        rows.append(np.asarray(kernel.cumulative[key], dtype=np.float64))
    n = min(len(r) for r in rows)
    return np.ascontiguousarray(np.stack([r[:n] for r in rows]))


def beam_frame(source):
    """Rows v_x, v_y, v_z of a Source (conehead/source.py:77-111) as FP64:
    v_y is the beam direction, v_x and v_z span the isocentre plane."""
    return np.array([source.v_x, source.v_y, source.v_z], dtype=np.float64)


def beam_aligned_direction(theta, phi, frame):
    """Physics-mode cone direction: the cuda_dose direction taken in the
    beam frame (polar axis = beam direction v_y) and expressed in grid
    coordinates, d = a v_x + b v_y + c v_z.  With the gantry and collimator
    at 0 the frame is the identity and d is the cuda_dose direction bit for
    bit."""
    a, b, c = cuda_dose_direction(theta, phi)
    f = np.asarray(frame, dtype=np.float64)
    return np.array([a * f[0, k] + b * f[1, k] + c * f[2, k]
                     for k in range(3)], dtype=np.float64)


class ConeSet:
    """n_theta x n_phi cones ordered theta-major as convolve_c loops them
    (conehead/convolve_c.pyx:88-89); cone (i, j) uses kernel row j.

    semantics: "convolve_c" (z-polar, discretised directions, kernel dict),
    "cuda_dose" (y-polar, table rows) or "physics" (the cuda_dose cones in
    the beam frame `frame`, see beam_frame, with the kernel looked up at
    the radiological distance)."""

    def __init__(self, thetas, phis, kernel, semantics="convolve_c",
                 frame=None):
        self.thetas = np.asarray(thetas, dtype=np.float64).ravel()
        self.phis = np.asarray(phis, dtype=np.float64).ravel()
        if semantics not in ("convolve_c", "cuda_dose", "physics"):
            raise ValueError(
                "semantics must be convolve_c, cuda_dose or physics")
        if frame is not None and semantics != "physics":
            raise ValueError("a beam frame applies to physics cones only")
        self.semantics = semantics
        self.frame = np.eye(3) if frame is None else \
            np.asarray(frame, dtype=np.float64).reshape(3, 3)
        if semantics == "convolve_c":
            fn = convolve_c_direction
        elif semantics == "cuda_dose":
            fn = cuda_dose_direction
        else:
            def fn(t, p):
                return beam_aligned_direction(t, p, self.frame)
        nt, nph = len(self.thetas), len(self.phis)
        self.dirs = np.zeros((nt * nph, 3), np.float64)
        self.rows = np.zeros(nt * nph, np.int32)
        for i, t in enumerate(self.thetas):
            for j, p in enumerate(self.phis):
                self.dirs[i * nph + j] = fn(t, p)
                self.rows[i * nph + j] = j
        if semantics == "convolve_c":
            self.cum = kernel_rows(kernel, self.phis)
        else:
            cum = np.asarray(kernel, dtype=np.float64)
            self.cum = np.ascontiguousarray(cum)
        self.n_cones = len(self.rows)

    def key(self):
        """Content hash input identifying the cone set."""
        return (self.semantics, self.dirs.tobytes(), self.rows.tobytes(),
                self.cum.tobytes())


def default_thetas(n=12):
    """Azimuths linspace(0, 360, n, endpoint=False)
    (conehead/conehead.py:789)."""
    return np.linspace(0, 360, n, endpoint=False, dtype=np.float64)


def kernel_phis(kernel):
    """All cone keys of a kernel as angles, sorted like the reference's
    commented convolve_c call (conehead/conehead.py:783-786)."""
    return np.array(sorted(p for p in kernel.cumulative if p != "radii"),
                    dtype=np.float64)
