"""Analytic line source, restated from ``MATLAB/linesource.m:1-65``.

Test infrastructure only.  A line source of length L and width a, vibrating
with velocity U0, lies on the y axis centred at the origin, sampled at n_ls
points y = linspace(-L/2, L/2, n_ls) (linesource.m:29-30).  With
r = |x - (0, y, 0)|:

* pressure (linesource.m:31-34, 54-56):
  p(x) = i (rho c U0 k a / 2) trapz_y e^{ikr} / r;
* node normals (linesource.m:36-53): the cross products (x1 - x3) x (x2 - x3)
  of the node's triangles, summed and normalised;
* normal velocity (linesource.m:57-60):
  v(x) = (U0 a / 2) trapz_y (ikr - 1) e^{ikr} / r^2 * (n . (x - y)) / r.

Pins (tests/test_oracle_linesource.py): the infinite-line limit with a
damped wavenumber against i pi H0^(1)(k rho) and its derivative
-i pi k H1^(1)(k rho) (scipy.special.hankel1); Euler's relation
v = dp/dn / (i k rho c) by central differences of ``pressure``; normals of
planar and hand-built meshes; one sample gives zero (trapz of one point).
"""
from __future__ import annotations

import numpy as np


def source_points(n_ls, L):
    """linesource.m:29-30 (MATLAB linspace ends exactly on +-L/2)."""
    return np.linspace(-L / 2.0, L / 2.0, int(n_ls))


def _trapz(f, y):
    """MATLAB trapz(y, f, 1): sum of interval widths times mean end values."""
    if len(y) < 2:
        return np.zeros(f.shape[1:], dtype=f.dtype)
    dy = np.diff(y)[:, None]
    return np.sum(dy * (f[:-1] + f[1:]), axis=0) / 2.0


def node_normals(r_ms, e_n):
    """Normalised sum over each node's triangles of (x1 - x3) x (x2 - x3)
    (linesource.m:36-53); r_ms (m, 3), e_n (n, 3) 0-based."""
    X = np.asarray(r_ms, dtype=float)
    E = np.asarray(e_n, dtype=int)
    d1 = X[E[:, 0]] - X[E[:, 2]]
    d2 = X[E[:, 1]] - X[E[:, 2]]
    nor = np.cross(d1, d2)
    acc = np.zeros_like(X)
    for a in range(3):
        np.add.at(acc, E[:, a], nor)
    with np.errstate(invalid="ignore", divide="ignore"):
        return acc / np.linalg.norm(acc, axis=1)[:, None]


def pressure(n_ls, pts, rho, c, U0, k, a, L):
    """i (rho c U0 k a / 2) trapz(e^{ikr}/r) at pts (n, 3); k may be complex."""
    y = source_points(n_ls, L)
    P = np.asarray(pts, dtype=float)
    r = np.sqrt(P[None, :, 0] ** 2 + (P[None, :, 1] - y[:, None]) ** 2 + P[None, :, 2] ** 2)
    G = np.exp(1j * k * r) / r
    return 1j * (rho * c * U0 * k * a / 2.0) * _trapz(G, y)


def velocity(n_ls, pts, normals, U0, k, a, L):
    """(U0 a / 2) trapz((ikr - 1) e^{ikr}/r^2 * n.(x - y)/r) at pts (n, 3)."""
    y = source_points(n_ls, L)
    P = np.asarray(pts, dtype=float)
    N = np.asarray(normals, dtype=float)
    dx = P[None, :, 0]
    dy = P[None, :, 1] - y[:, None]
    dz = P[None, :, 2]
    r = np.sqrt(dx ** 2 + dy ** 2 + dz ** 2)
    drn = (N[None, :, 0] * dx + N[None, :, 1] * dy + N[None, :, 2] * dz) / r
    F = (1j * k * r - 1.0) * (np.exp(1j * k * r) / r ** 2) * drn
    return (U0 * a / 2.0) * _trapz(F, y)


def linesource(n_ls, r_ms, e_n, r_lh, rho, c, U0, k, a, L):
    """[xp, xsf, vsf] of linesource.m for one wavenumber k."""
    nv = node_normals(r_ms, e_n)
    xp = pressure(n_ls, r_lh, rho, c, U0, k, a, L) if len(r_lh) else np.zeros(0, complex)
    xsf = pressure(n_ls, r_ms, rho, c, U0, k, a, L)
    vsf = velocity(n_ls, r_ms, nv, U0, k, a, L)
    return xp, xsf, vsf
