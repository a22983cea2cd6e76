"""ESM (equivalent source method) restatement.

Reference: ``CUDA/esm.cuh:1-36`` (kernels), ``CUDA/esm.cuh:112-174`` (CPU
``esm``), ``MATLAB/esmV.m:20-30`` (vectorized spec, valid for l != m).

Pinned by (tests/test_oracle_esm.py):
* entries against an independent per-entry mpmath evaluation on tiny sets;
* the normal-derivative identity: D_ij equals the finite-difference normal
  derivative of e^{ikR}/(4 pi R) along n_i;
* the solve identity H_e D = M_hq (residual) and M_e D = M_sq;
* the pulsating sphere: with sources retreated inside a sphere, H_e maps the
  uniform surface velocity to the analytic field pressure to discretisation
  accuracy (``CUDA/utils.cuh:384-412``).
"""
from __future__ import annotations

import numpy as np

C_AIR_REF = 343.5   # CUDA/constants.cuh:2
RHO_AIR_REF = 1.19  # CUDA/constants.cuh:5


def esm_assemble(k, r_s, r_h, src, normals=None, rho=RHO_AIR_REF, c=C_AIR_REF):
    """Return (D, M_sq, M_hq) as complex128 (m x m, m x m, l x m).

    D_ij   = ((ikR - 1)/(4 pi R^2)) e^{ikR} cos(theta_ij)        esm.cuh:16, esmV.m:27
    M_sq   = i rho c k e^{ikR}/(4 pi R)                          esm.cuh:17, esmV.m:28
    M_hq   = i rho c k e^{ikR_hq}/(4 pi R_hq)                    esm.cuh:33, esmV.m:29
    with R = |r_s,i - q_j| and cos(theta) = n_i.(r_s,i - q_j)/R; the reference
    uses the radial normal n_i = r_s,i/|r_s,i| (esm.cuh:11-15, esmV.m:26),
    which is the default when ``normals`` is None.
    """
    r_s = np.asarray(r_s, np.float64)
    r_h = np.asarray(r_h, np.float64)
    src = np.asarray(src, np.float64)
    if normals is None:
        normals = r_s / np.linalg.norm(r_s, axis=1, keepdims=True)
    diff = r_s[:, None, :] - src[None, :, :]                 # (m, m, 3)
    R = np.sqrt(np.sum(diff * diff, axis=2))
    cos_t = np.sum(normals[:, None, :] * diff, axis=2) / R
    e = np.exp(1j * k * R)
    D = ((1j * k * R - 1.0) / (4 * np.pi * R * R)) * e * cos_t
    M_sq = (1j * rho * c * k) * (e / (4 * np.pi * R))
    Rh = np.sqrt(np.sum((r_h[:, None, :] - src[None, :, :]) ** 2, axis=2))
    M_hq = (1j * rho * c * k) * (np.exp(1j * k * Rh) / (4 * np.pi * Rh))
    return D, M_sq, M_hq


def esmV(rho, c, k, r_ms, r_lh, r_i):
    """MATLAB/esmV.m:1-30: returns (H_e, M_sq, D) with H_e = M_hq / D."""
    D, M_sq, M_hq = esm_assemble(k, r_ms, r_lh, r_i * np.asarray(r_ms), None, rho, c)
    H_e = np.linalg.solve(D.T, M_hq.T).T
    return H_e, M_sq, D


def esm(k, r_ms, r_lh, r_i, rho=RHO_AIR_REF, c=C_AIR_REF, src=None, normals=None):
    """CUDA/esm.cuh:112-174 (CPU ``esm``): returns (M_e, H_e).

    M_e = M_sq D^-1 (m x m) and H_e = M_hq D^-1 (l x m), both solved from the
    transposed systems D^T X^T = M^T as the reference does with ``zgels``
    (esm.cuh:153-170).  The reference's CUDA layout only works for l == m
    (esm.cuh:34,142,165-168); this restatement follows esmV.m, valid for any l.
    """
    r_ms = np.asarray(r_ms, np.float64)
    if src is None:
        src = r_i * r_ms
    D, M_sq, M_hq = esm_assemble(k, r_ms, r_lh, src, normals, rho, c)
    Xt = np.linalg.solve(D.T, np.concatenate([M_sq.T, M_hq.T], axis=1))
    m = r_ms.shape[0]
    return Xt[:, :m].T.copy(), Xt[:, m:].T.copy()
