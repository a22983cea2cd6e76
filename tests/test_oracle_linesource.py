"""Pins of the line-source oracle (MATLAB/linesource.m) against closed forms:
the infinite line (Hankel functions), Euler's relation between pressure and
normal velocity, hand-computed mesh normals and the one-sample case."""
import numpy as np
import pytest
from scipy.special import hankel1

from oracle import linesource as ls
from inverseacoustics_amd.scenarios import icosphere

RHO, C, U0, A = 1.21, 343.0, 0.7, 0.01


def test_infinite_line_pressure_and_velocity_match_hankel():
    # damped wavenumber: the truncated line of length 40 is infinite to e^-40
    k = 10.0 + 2.0j
    rho_pts = np.array([0.3, 0.55, 1.1])
    pts = np.stack([rho_pts * np.cos(0.4), np.zeros(3), rho_pts * np.sin(0.4)], 1)
    normals = pts / rho_pts[:, None]
    p = ls.pressure(4001, pts, RHO, C, U0, k, A, 40.0)
    v = ls.velocity(4001, pts, normals, U0, k, A, 40.0)
    # int e^{ik sqrt(rho^2+y^2)}/sqrt(rho^2+y^2) dy = i pi H0(k rho)
    p_ref = 1j * (RHO * C * U0 * k * A / 2) * (1j * np.pi * hankel1(0, k * rho_pts))
    v_ref = (U0 * A / 2) * (-1j * np.pi * k * hankel1(1, k * rho_pts))
    np.testing.assert_allclose(p, p_ref, rtol=1e-9)
    np.testing.assert_allclose(v, v_ref, rtol=1e-9)


def test_velocity_is_normal_derivative_of_pressure():
    # Euler: v_n = (dp/dn) / (i k rho c), e^{-i w t} convention
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.3, 0.3, (20, 3))
    pts[:, 0] += 0.4  # keep away from the line
    nrm = rng.normal(size=(20, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    for k in (2.0, 17.0):
        v = ls.velocity(101, pts, nrm, U0, k, A, 0.5)
        h = 1e-5
        dp = (ls.pressure(101, pts + h * nrm, RHO, C, U0, k, A, 0.5) -
              ls.pressure(101, pts - h * nrm, RHO, C, U0, k, A, 0.5)) / (2 * h)
        np.testing.assert_allclose(v, dp / (1j * k * RHO * C), rtol=1e-7)


def test_short_line_is_a_monopole():
    k, L = 9.0, 1e-4
    pts = np.array([[0.2, 0.05, -0.1], [0.0, 0.3, 0.4]])
    r = np.linalg.norm(pts, axis=1)
    p = ls.pressure(3, pts, RHO, C, U0, k, A, L)
    ref = 1j * (RHO * C * U0 * k * A / 2) * L * np.exp(1j * k * r) / r
    np.testing.assert_allclose(p, ref, rtol=1e-6)


def test_node_normals_hand_computed():
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 2], [5, 5, 5]], float)
    E = np.array([[0, 1, 2], [0, 3, 1]])
    # (x0-x2)x(x1-x2) = (0,0,1); (x0-x1)x(x3-x1) = (0,2,0)
    n = ls.node_normals(X, E)
    s5 = np.sqrt(5.0)
    np.testing.assert_allclose(n[0], [0, 2 / s5, 1 / s5], atol=1e-15)
    np.testing.assert_allclose(n[1], [0, 2 / s5, 1 / s5], atol=1e-15)
    np.testing.assert_allclose(n[2], [0, 0, 1], atol=1e-15)
    np.testing.assert_allclose(n[3], [0, 1, 0], atol=1e-15)
    assert np.isnan(n[4]).all()  # unused node: 0/0 as in the reference


def test_node_normals_of_sphere_are_radial():
    X, E = icosphere(3, 0.2)
    n = ls.node_normals(X, E)
    cosang = np.sum(n * X / np.linalg.norm(X, axis=1)[:, None], axis=1)
    assert cosang.min() > 0.999


def test_one_sample_gives_zero():
    pts = np.array([[0.3, 0.0, 0.0]])
    assert ls.pressure(1, pts, RHO, C, U0, 5.0, A, 0.2)[0] == 0
    assert ls.velocity(1, pts, np.array([[1.0, 0, 0]]), U0, 5.0, A, 0.2)[0] == 0


def test_trapz_matches_numpy():
    y = ls.source_points(7, 0.3)
    f = np.exp(1j * y)[:, None] * np.arange(1, 4)[None, :]
    assert np.allclose(ls._trapz(f, y), np.trapezoid(f, y, axis=0), rtol=0, atol=1e-15)


@pytest.mark.parametrize("n_ls", [2, 64])
def test_linesource_shapes(n_ls):
    X, E = icosphere(1, 0.2)
    H = 2.0 * X[:10]
    xp, xsf, vsf = ls.linesource(n_ls, X, E, H, RHO, C, U0, 4.0, A, 0.1)
    assert xp.shape == (10,) and xsf.shape == (len(X),) and vsf.shape == (len(X),)
    assert np.isfinite(xp).all() and np.isfinite(vsf).all()
