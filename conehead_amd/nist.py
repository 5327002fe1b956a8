"""Mass attenuation coefficient of liquid water (NIST XCOM / Hubbell-Seltzer).

Follows conehead/nist.py:136-187 (mu_water): linear interpolation of the
tabulated mu/rho (cm^2/g) over energy (MeV), stored as float32, with an
error outside the table range like scipy's interp1d (bounds_error=True).
The interpolation is evaluated in float64.
"""
import numpy as np

# (E [MeV], mu/rho [cm^2/g]) -- NIST tabulated values for liquid water.
_WATER = np.array([
    (1.00000E-03, 4.078E+03), (1.50000E-03, 1.376E+03),
    (2.00000E-03, 6.173E+02), (3.00000E-03, 1.929E+02),
    (4.00000E-03, 8.278E+01), (5.00000E-03, 4.258E+01),
    (6.00000E-03, 2.464E+01), (8.00000E-03, 1.037E+01),
    (1.00000E-02, 5.329E+00), (1.50000E-02, 1.673E+00),
    (2.00000E-02, 8.096E-01), (3.00000E-02, 3.756E-01),
    (4.00000E-02, 2.683E-01), (5.00000E-02, 2.269E-01),
    (6.00000E-02, 2.059E-01), (8.00000E-02, 1.837E-01),
    (1.00000E-01, 1.707E-01), (1.50000E-01, 1.505E-01),
    (2.00000E-01, 1.370E-01), (3.00000E-01, 1.186E-01),
    (4.00000E-01, 1.061E-01), (5.00000E-01, 9.687E-02),
    (6.00000E-01, 8.956E-02), (8.00000E-01, 7.865E-02),
    (1.00000E+00, 7.072E-02), (1.25000E+00, 6.323E-02),
    (1.50000E+00, 5.754E-02), (2.00000E+00, 4.942E-02),
    (3.00000E+00, 3.969E-02), (4.00000E+00, 3.403E-02),
    (5.00000E+00, 3.031E-02), (6.00000E+00, 2.770E-02),
    (8.00000E+00, 2.429E-02), (1.00000E+01, 2.219E-02),
    (1.50000E+01, 1.941E-02), (2.00000E+01, 1.813E-02),
], dtype=np.float32)


def mu_water(E):
    """mu/rho of water at energies E (MeV); float64 result."""
    x = _WATER[:, 0].astype(np.float64)
    y = _WATER[:, 1].astype(np.float64)
    e = np.asarray(E, dtype=np.float64)
    if np.any(e < x[0]) or np.any(e > x[-1]):
        raise ValueError("A value in x_new is outside the interpolation "
                         "range.")
    return np.interp(e, x, y)
