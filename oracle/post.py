"""Reconstruction, sound power and the analytic pulsating sphere.

Reference: ``CUDA/utils.cuh:199-289`` (to_pressure, to_velocity, sound_power)
and ``CUDA/utils.cuh:384-412`` (sphere).

Pinned by (tests/test_oracle_post.py): the monopole closed form
W = 2 pi a^2 rho c u^2 (ka)^2/(1+(ka)^2) for a uniformly sampled sphere, the
identity p = i rho c u h_0(kr)/h_0'(ka) with scipy's spherical Bessel
functions, and numpy's vdot for the inner product.
"""
from __future__ import annotations

import numpy as np

from .esm import C_AIR_REF, RHO_AIR_REF


def to_pressure(M, x):
    """p = M x (utils.cuh:199-230, cublasZgemv 'N')."""
    return np.asarray(M) @ np.asarray(x)


def to_velocity(Q, x):
    """v = Q x (utils.cuh:232-263)."""
    return np.asarray(Q) @ np.asarray(x)


def sound_power(p, v, S):
    """W = 1/2 Re(v^H p) S/m, equal nodal areas (utils.cuh:265-289, :279-280)."""
    m = len(p)
    return 0.5 * float(np.real(np.vdot(v, p))) * (S / m)


def power_db(W):
    """L_W = 10 log10(W / 1e-12)."""
    return 10.0 * np.log10(np.asarray(W) / 1e-12)


def sphere(r_ms, r_lh, k, u, a, rho=RHO_AIR_REF, c=C_AIR_REF):
    """Pulsating sphere (utils.cuh:384-412): p = (a/r) rho c u (ika/(ika-1)) e^{ik(r-a)}
    at surface and field points, v_s = u.  Returns (p_s, p_h, v_s)."""
    def p_at(pts):
        r = np.linalg.norm(np.asarray(pts, np.float64), axis=1)
        ika = 1j * k * a
        return (a / r) * rho * c * u * (ika / (ika - 1)) * np.exp(1j * k * (r - a))
    p_s = p_at(r_ms)
    p_h = p_at(r_lh)
    v_s = np.full(len(p_s), u, dtype=np.complex128)
    return p_s, p_h, v_s
