"""Radiation-mode inverse solve on elemental radiators -- test
infrastructure (the checker of ``ia_radiation_sweep``).

The reference's radiation-mode path (MATLAB/invrS.m:157-170,
CUDA/invr.cuh:426-433) takes the modes of a radiation-resistance matrix R
(``Q = U_r(:, 1:n)``), maps them to the microphones (``H_r = H Q``) and
solves the regularised problem in mode space.  Here R is the elemental-
radiator form for a baffled plate of np equal patches of area A,
R_ij = (w^2 rho A^2 / (4 pi c)) sin(k r_ij) / (k r_ij) (Elliott & Johnson,
JASA 94 (1993) 2194), so that W = v^H R v for patch velocities v (peak
amplitudes); the microphone transfer is the Rayleigh integral
G_ij = -(i w rho A / (2 pi)) e^{ik R_ij} / R_ij.  Per frequency:
R = Q diag(e) Q^T (LAPACK eigh, the n leading pairs), H_r = G Q, y the
regularised solution of H_r y = b (``pipeline._regularised``, the csvd /
gcvG / tikhonovG chain), W = sum_i e_i |y_i|^2 and v = Q y; modes below
eig_floor e_1 are dropped.

Pinned by (tests/test_oracle_radiation.py): R against its closed form and
symmetric PSD; W from the modes equals v^H R v; with all modes kept and no
noise the modal power equals the power of the true velocities.
"""
from __future__ import annotations

import math

import numpy as np

from . import pipeline


def resistance(pts, area, k, rho, c):
    """R_ij = (w^2 rho A^2/(4 pi c)) sin(k r_ij)/(k r_ij), R_ii = w^2 rho A^2/(4 pi c)."""
    w = k * c
    r = np.linalg.norm(pts[:, None, :] - pts[None, :, :], axis=2)
    return (w * w * rho * area * area / (4 * math.pi * c)) * np.sinc(k * r / math.pi)


def rayleigh(pts, r_h, area, k, rho, c):
    """G_ij = -(i w rho A/(2 pi)) e^{ikR}/R, microphone i, patch j."""
    w = k * c
    R = np.linalg.norm(r_h[:, None, :] - pts[None, :, :], axis=2)
    return (-1j * w * rho * area / (2 * math.pi)) * np.exp(1j * k * R) / R


def modes(pts, area, k, n_modes, rho, c):
    """The n_modes leading eigenpairs of R (descending), each mode's largest
    entry positive."""
    import scipy.linalg as sl

    R = resistance(pts, area, k, rho, c)
    n = R.shape[0]
    e, Q = sl.eigh(R, subset_by_index=[n - n_modes, n - 1])
    e, Q = e[::-1], Q[:, ::-1]
    sgn = np.sign(Q[np.argmax(np.abs(Q), axis=0), np.arange(n_modes)])
    return e, Q * sgn


def radiation_solve(k, pts, area, r_h, b, n_modes, rho, c, filter="tikhonov", select="gcv",
                    n_lambda=50, decades=8.0, driver="gesdd", eig_floor=1e-8, Q=None):
    """The mode-space solve.  Modes with e_i < eig_floor e_1 are dropped
    (zero columns: their shapes are not resolved in FP64, so a fit over them
    would depend on an arbitrary basis of the round-off subspace).  Q, when
    given, replaces the eigh modes (to check the rest of the chain on
    another implementation's modes)."""
    e, Q0 = modes(pts, area, k, n_modes, rho, c)
    Q = (Q0 if Q is None else np.asarray(Q, dtype=np.float64)).copy()
    Q[:, e < eig_floor * e[0]] = 0.0
    H = rayleigh(pts, r_h, area, k, rho, c) @ Q
    r = pipeline._regularised(H, b, filter, select, n_lambda, decades, driver)
    y = r["x"]
    r.update(e=e, Q=Q, y=y, W=float(np.sum(e * np.abs(y) ** 2)), v=Q @ y, H=H)
    return r
