"""Pins of the radiation-mode oracle (oracle/radiation.py)."""
import math

import numpy as np
import pytest

from inverseacoustics_amd import scenarios as sc
from oracle import radiation as orad


def _grid(nx, ny):
    return sc.grid_points(nx, ny, 0.6, 0.5, 0.0), 0.6 * 0.5 / (nx * ny)


def test_resistance_closed_form_and_psd():
    pts, A = _grid(10, 8)
    k, rho, c = 2 * math.pi * 700 / 343.0, 1.21, 343.0
    R = orad.resistance(pts, A, k, rho, c)
    w = k * c
    i, j = 3, 57
    r = np.linalg.norm(pts[i] - pts[j])
    assert R[i, j] == pytest.approx(w * w * rho * A * A / (4 * math.pi * c) * math.sin(k * r) / (k * r),
                                    rel=1e-14)
    assert R[5, 5] == pytest.approx(w * w * rho * A * A / (4 * math.pi * c), rel=1e-15)
    np.testing.assert_array_equal(R, R.T)
    assert np.linalg.eigvalsh(R).min() > -1e-12 * np.abs(R).max()


def test_modal_power_is_quadratic_form():
    pts, A = _grid(12, 10)
    k, rho, c = 2 * math.pi * 1200 / 343.0, 1.21, 343.0
    n = len(pts)
    e, Q = orad.modes(pts, A, k, n, rho, c)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    R = orad.resistance(pts, A, k, rho, c)
    y = Q.T @ v
    assert np.sum(e * np.abs(y) ** 2) == pytest.approx(np.vdot(v, R @ v).real, rel=1e-12)
    # a small patch's power: baffled piston limit W = w^2 rho A^2 |v|^2/(4 pi c)
    assert R[0, 0] == pytest.approx((k * c) ** 2 * rho * A * A / (4 * math.pi * c))


def test_noise_free_reconstruction_recovers_power():
    pts, A = _grid(12, 10)
    rho, c = 1.21, 343.0
    r_h = sc.grid_points(10, 10, 0.6, 0.5, 0.1)
    rng = np.random.default_rng(4)
    v = rng.standard_normal(len(pts)) + 1j * rng.standard_normal(len(pts))
    k = 2 * math.pi * 500 / c
    b = orad.rayleigh(pts, r_h, A, k, rho, c) @ v
    o = orad.radiation_solve(k, pts, A, r_h, b, 40, rho, c, select="gcv")
    W_true = np.vdot(v, orad.resistance(pts, A, k, rho, c) @ v).real
    # the efficient modes carry the power: the mode-space estimate is close
    assert abs(10 * np.log10(o["W"] / W_true)) < 0.5
    np.testing.assert_allclose(o["v"], o["Q"] @ o["y"])
