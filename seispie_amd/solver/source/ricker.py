import numpy as np


def ricker(t, f0, t0, ang, amp):
    """Ricker wavelet and force direction (seispie/solver/source/ricker.py:7-11).

    s(t) = -amp (1 - 2a (t-t0)^2) exp(-a (t-t0)^2), a = (pi f0)^2; returns
    (s sin(ang), s, -s cos(ang)) for the x, y and z components, with ang in
    degrees.  t may be a scalar or an array (float64 arithmetic either way).
    """
    a = (np.pi * f0) ** 2
    stf = -amp * (1.0 - 2.0 * a * (t - t0) ** 2) * np.exp(-a * (t - t0) ** 2)
    ang_rad = np.deg2rad(ang)
    return stf * np.sin(ang_rad), stf, -stf * np.cos(ang_rad)
