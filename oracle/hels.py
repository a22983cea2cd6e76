"""HELS (Helmholtz equation least squares) restatement.

Reference: ``CUDA/hels.cuh:222-323`` (CPU ``hels`` with GSL), column order
of ``CUDA/hels.cuh:92`` (i -> n = floor(sqrt(i)), m = i - n - n^2, so m runs
-n..n), maths of ``MATLAB/helsV.m:1-41`` / ``helsS.m``:

  Y_n^m(th, ph) = sqrt((2n+1)(n-m)!/(4 pi (n+m)!)) P_n^m(cos th) e^{i m ph},  m >= 0
  P_n^m(x)      = (-1)^m (1-x^2)^{m/2} (2m)!/(2^m m!) C^{(m+1/2)}_{n-m}(x)
  Y_n^{-m}      = (-1)^m conj(Y_n^m)
  M_e = h_n(k r_s) Y,  Q_e = k h_n'(k r_s) Y / (i rho c k),  H = h_n(k r_h) Y,
  h_n = j_n + i y_n  (hels.cuh:109).

Inputs are spherical coordinates (azimuth, elevation, radius) as produced by
``cartesian_2_spherical`` (utils.cuh:414-427); "elevation" is the polar
angle from +z (only its cosine is used).

Pinned by (tests/test_oracle_hels.py): scipy.special spherical_jn/yn (value
and derivative, the GSL Jnu/Ynu equivalents), scipy.special.sph_harm_y
(spherical harmonics with the Condon-Shortley phase), eval_gegenbauer, and
the exact representation of a monopole at the origin by the (0,0) term.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import jv, yv


def column_nm(n_t):
    """CUDA column order (hels.cuh:92): i -> (n, m), m ascending -n..n."""
    i = np.arange(n_t)
    n = np.floor(np.sqrt(i)).astype(int)
    return n, i - n - n * n


def gegenbauer(nmax, lam, x):
    """C^{(lam)}_n(x) for n = 0..nmax by the three-term recurrence of
    MATLAB/gegenbauer.m:5-26 (and hels.cuh:1-16); returns (nmax+1, len(x))."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    G = np.zeros((nmax + 1, x.size))
    G[0] = 1.0
    if nmax >= 1:
        G[1] = 2.0 * lam * x
    for n in range(2, nmax + 1):
        G[n] = (1.0 / n) * (2.0 * (n - 1 + lam) * x * G[n - 1] - (n - 1 + 2.0 * lam - 1.0) * G[n - 2])
    return G


def ynm(N, theta, phi):
    """Y (points x (N+1)^2) in the CUDA column order."""
    theta = np.atleast_1d(np.asarray(theta, np.float64))
    phi = np.atleast_1d(np.asarray(phi, np.float64))
    x = np.cos(theta)
    n_t = (N + 1) ** 2
    Y = np.zeros((theta.size, n_t), dtype=np.complex128)
    for n in range(N + 1):
        for m in range(0, n + 1):
            C = gegenbauer(n - m, m + 0.5, x)[n - m]
            L = ((-1.0) ** m) * (1.0 - x * x) ** (m / 2.0) * (math.factorial(2 * m) / (2 ** m * math.factorial(m))) * C
            norm = math.sqrt((2 * n + 1) * math.factorial(n - m) / (4 * math.pi * math.factorial(n + m)))
            y = norm * L * np.exp(1j * m * phi)
            Y[:, n * n + n + m] = y
            if m > 0:
                Y[:, n * n + n - m] = ((-1.0) ** m) * np.conj(y)
    return Y


def sph_hankel(N, x):
    """h_n(x) = j_n + i y_n and h_n'(x), n = 0..N, from the half-integer
    Bessel functions as hels.cuh:239-251 does with GSL:
    j_n = sqrt(pi/(2x)) J_{n+1/2},  x j_n' = n j_n - x sqrt(pi/(2x)) J_{n+3/2}."""
    x = np.atleast_1d(np.asarray(x, np.float64))
    n = np.arange(N + 1)[:, None]
    s = np.sqrt(np.pi / (2 * x))[None, :]
    J = jv(n + 0.5, x[None, :])
    Yb = yv(n + 0.5, x[None, :])
    J1 = jv(n + 1.5, x[None, :])
    Y1 = yv(n + 1.5, x[None, :])
    h = s * (J + 1j * Yb)
    dh = math.sqrt(math.pi / 2) * x[None, :] ** -1.5 * ((n * J - x[None, :] * J1) + 1j * (n * Yb - x[None, :] * Y1))
    return h, dh  # (N+1, points)


def hels(k, azelr_s, azelr_h, N, rho=1.19, c=343.5):
    """CUDA/hels.cuh:222 ``hels``: returns (M_e (m x n_t), H (l x n_t), Q_e (m x n_t))."""
    s = np.asarray(azelr_s, np.float64)
    h = np.asarray(azelr_h, np.float64)
    n_t = (N + 1) ** 2
    nn, _ = column_nm(n_t)
    Ys = ynm(N, s[:, 1], s[:, 0])
    Yh = ynm(N, h[:, 1], h[:, 0])
    hs, dhs = sph_hankel(N, k * s[:, 2])
    hh, _ = sph_hankel(N, k * h[:, 2])
    M_e = hs[nn].T * Ys
    Q_e = (1.0 / (1j * rho * c * k)) * (k * dhs[nn].T) * Ys
    H = hh[nn].T * Yh
    return M_e, H, Q_e


def cartesian_2_spherical(pts):
    """CUDA/utils.cuh:414-427: (az in [0, 2 pi), el = atan2(sqrt(x^2+y^2), z) - 2 pi if x < 0, r)."""
    p = np.asarray(pts, np.float64)
    xy = p[:, 0] ** 2 + p[:, 1] ** 2
    r = np.sqrt(xy + p[:, 2] ** 2)
    az = np.arctan2(p[:, 1], p[:, 0])
    az = np.where(p[:, 1] < 0, az + 2 * np.pi, az)
    el = np.arctan2(np.sqrt(xy), p[:, 2])
    el = np.where(p[:, 0] < 0, el - 2 * np.pi, el)
    return np.stack([az, el, r], axis=1)
