"""Microphone-array layouts (MATLAB/rectangular.m, arcondoulis.m,
bruelkjaer.m, underbrink.m) against their defining properties: counts,
face planes, inner and outer radii, arm symmetry and the reference's
microphone order on small hand-worked cases."""
import math

import numpy as np
import pytest

from inverseacoustics_amd import arrays


def _face(xyz, n, f):
    return xyz[f * n:(f + 1) * n]


def _check_faces(xyz, n, DOM):
    assert xyz.shape == (6 * n, 3)
    planes = [(2, DOM), (2, -DOM), (1, DOM), (1, -DOM), (0, DOM), (0, -DOM)]
    base = _face(xyz, n, 0)[:, :2]
    for f, (ax, val) in enumerate(planes):
        F = _face(xyz, n, f)
        assert np.all(F[:, ax] == val)
        uv = np.delete(F, ax, axis=1)  # remaining (u, v) in their order
        np.testing.assert_array_equal(uv, base)
    return base


def test_rectangular_order_and_spacing():
    xyz, nfp = arrays.rectangular(0.5, 3, 2.0, 2, 1.0)
    assert nfp == 36
    uv = _check_faces(xyz, 6, 0.5)
    # meshgrid + column-major reshape: y runs fastest
    expect = [(-1, -0.5), (-1, 0.5), (0, -0.5), (0, 0.5), (1, -0.5), (1, 0.5)]
    np.testing.assert_allclose(uv, expect, atol=1e-15)


def test_arcondoulis_spiral_ends():
    r0, N, rmax, ex, ey, phi = 0.05, 24, 0.5, 0.9, 0.9, 5 * math.pi / 2
    xyz, nfp = arrays.arcondoulis(1.0, r0, N, rmax, ex, ey, phi)
    assert nfp == 6 * N
    uv = _check_faces(xyz, N, 1.0)
    np.testing.assert_allclose(uv[0], [r0, 0.0], atol=1e-15)     # a = r0 N/(ex N + 1)
    assert abs(np.hypot(*uv[-1]) - rmax) < 1e-12                # b fixes the last radius
    ang = np.unwrap(np.arctan2(uv[:, 1], uv[:, 0]))
    np.testing.assert_allclose(ang - ang[0], np.arange(N) * phi / (N - 1), atol=1e-12)


def test_arcondoulis_equal_stretch_is_log_spiral():
    # ex = ey: radius ((n + ex N)/N) a e^{b theta}
    r0, N, rmax, e, phi = 0.1, 10, 1.0, 0.5, 3.0
    xyz, _ = arrays.arcondoulis(2.0, r0, N, rmax, e, e, phi)
    uv = xyz[:N, :2]
    n = np.arange(1, N + 1)
    a = r0 * N / (e * N + 1)
    b = math.log(rmax / (a * (1 + e))) / phi
    th = (n - 1) * phi / (N - 1)
    np.testing.assert_allclose(np.hypot(uv[:, 0], uv[:, 1]), (n + e * N) / N * a * np.exp(b * th),
                               rtol=1e-13)


def test_bruelkjaer_arms_and_radii():
    r0, Na, rmax, phi = 0.05, 5, 0.6, math.pi / 3
    d = np.array([0.0, 0.25, 0.5, 1.0])
    xyz, nfp = arrays.bruelkjaer(0.8, r0, Na, rmax, phi, d)
    assert nfp == 6 * Na * len(d)
    uv = _check_faces(xyz, Na * len(d), 0.8)
    rad = np.hypot(uv[:, 0], uv[:, 1]).reshape(Na, len(d))   # arm-major
    np.testing.assert_allclose(rad[:, 0], r0, rtol=1e-14)
    np.testing.assert_allclose(rad[:, -1], rmax, rtol=1e-14)
    # every arm is arm 1 rotated by 2 pi a / Na
    arm0 = uv[:len(d)]
    for a in range(1, Na):
        t = 2 * math.pi * a / Na
        R = np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]])
        np.testing.assert_allclose(uv[a * len(d):(a + 1) * len(d)], arm0 @ R.T, atol=1e-14)
    # theta = asin(...) keeps arm 1 in the half plane u >= 0 (the reference's
    # branch: a chord that crosses the v axis is reflected back)
    assert np.all(arm0[:, 0] >= 0)


def test_bruelkjaer_arm_is_a_straight_chord():
    # l cos(phi) < r0: the chord from (r0, 0) stays in u > 0, where the
    # law-of-sines angle is the geometric one: p - (r0, 0) = l d (-cos phi, sin phi)
    r0, rmax, phi = 0.1, 0.6, 1.45
    d = np.linspace(0, 1, 7)
    xyz, _ = arrays.bruelkjaer(1.0, r0, 1, rmax, phi, d)
    ell = r0 * math.cos(phi) + math.sqrt(rmax ** 2 - r0 ** 2 * math.sin(phi) ** 2)
    expect = np.array([r0, 0.0]) + ell * d[:, None] * np.array([-math.cos(phi), math.sin(phi)])
    np.testing.assert_allclose(xyz[:7, :2], expect, atol=1e-14)


def test_underbrink_rings_and_order():
    r0, Na, Nm, rmax, v = 0.04, 3, 4, 0.5, 5 * math.pi / 16
    xyz, nfp = arrays.underbrink(1.5, r0, Na, Nm, rmax, v)
    assert nfp == 6 * Na * Nm
    uv = _check_faces(xyz, Na * Nm, 1.5)
    rad = np.hypot(uv[:, 0], uv[:, 1]).reshape(Nm, Na)        # ring-major
    j = np.arange(2, Nm + 1)
    ring = np.concatenate([[r0], np.sqrt((2 * j - 3) / (2 * Nm - 3)) * rmax])
    np.testing.assert_allclose(rad, np.repeat(ring[:, None], Na, 1), rtol=1e-14)
    # equal-area rings beyond the first: r_j^2 - r_{j-1}^2 constant for j >= 3
    np.testing.assert_allclose(np.diff(ring[1:] ** 2), 2 * rmax ** 2 / (2 * Nm - 3), rtol=1e-13)
    ang = np.arctan2(uv[:, 1], uv[:, 0]).reshape(Nm, Na)
    lsp = np.log(ring / r0) * math.tan(v)
    for a in range(Na):
        d = np.angle(np.exp(1j * (ang[:, a] - lsp - 2 * math.pi * a / Na)))
        np.testing.assert_allclose(d, 0, atol=1e-13)


@pytest.mark.parametrize("fn,args", [
    (arrays.rectangular, (1.0, 0, 1.0, 2, 1.0)),
    (arrays.arcondoulis, (1.0, 0.1, 1, 1.0, 0.5, 0.5, 1.0)),
    (arrays.underbrink, (1.0, 0.1, 2, 1, 1.0, 1.0)),
    (arrays.bruelkjaer, (1.0, 0.1, 2, 1.0, 1.0, [])),
])
def test_degenerate_sizes_rejected(fn, args):
    with pytest.raises(ValueError):
        fn(*args)
