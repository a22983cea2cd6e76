"""Microphone-array layouts on the six faces of a cube (MATLAB/rectangular.m,
arcondoulis.m, bruelkjaer.m, underbrink.m).

Each function builds one planar array in (u, v) and places a copy on every
face of the cube of half-side DOM centred at the origin, in the reference's
order: (u, v, +DOM), (u, v, -DOM), (u, +DOM, v), (u, -DOM, v), (+DOM, u, v),
(-DOM, u, v).  Returns (xyzFP (Nfp, 3), Nfp) like the MATLAB functions; the
points are float64 rows (x, y, z), in the reference's microphone order.

Host-side geometry: these only generate hologram points for the sweeps.
The planar layouts follow E. Prime and C. Doolan, "A comparison of popular
beamforming arrays", Proc. Acoustics 2013, as the reference does.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["cube_faces", "rectangular", "arcondoulis", "bruelkjaer", "underbrink"]


def cube_faces(u, v, DOM):
    """Six copies of the planar array (u, v) on the faces x, y, z = +-DOM
    (the stacking of rectangular.m:19-24 and the three spiral scripts)."""
    u = np.asarray(u, dtype=np.float64).ravel()
    v = np.asarray(v, dtype=np.float64).ravel()
    if u.shape != v.shape:
        raise ValueError("u and v must have the same length")
    d = np.full_like(u, float(DOM))
    xyz = np.concatenate([
        np.stack([u, v, d], 1), np.stack([u, v, -d], 1),
        np.stack([u, d, v], 1), np.stack([u, -d, v], 1),
        np.stack([d, u, v], 1), np.stack([-d, u, v], 1)])
    return xyz, xyz.shape[0]


def rectangular(DOM, Nx, Lx, Ny, Ly):
    """MATLAB/rectangular.m:14-24: an Nx x Ny grid spanning Lx x Ly, the y
    index running fastest (meshgrid then column-major reshape)."""
    Nx, Ny = int(Nx), int(Ny)
    if Nx < 1 or Ny < 1:
        raise ValueError("Nx and Ny must be >= 1")
    xs = np.linspace(-Lx / 2, Lx / 2, Nx)
    ys = np.linspace(-Ly / 2, Ly / 2, Ny)
    u = np.repeat(xs, Ny)
    v = np.tile(ys, Nx)
    return cube_faces(u, v, DOM)


def arcondoulis(DOM, r0, N, rmax, ex, ey, phi):
    """MATLAB/arcondoulis.m:18-30: N microphones on the Arcondoulis spiral,
    theta_n = (n - 1) phi / (N - 1), u = ((n + ex N)/N) a cos(theta) e^{b theta},
    v = ((n + ey N)/N) a sin(theta) e^{b theta}, with a = r0 N / (ex N + 1) so
    that the first microphone sits at (r0, 0), and b such that the last one
    lies at distance rmax from the centre."""
    N = int(N)
    if N < 2:
        raise ValueError("N must be >= 2")
    n = np.arange(1, N + 1, dtype=np.float64)
    theta = (n - 1) * phi / (N - 1)
    a = r0 * (N / (ex * N + 1))
    b = (1.0 / phi) * math.log(rmax / (a * math.sqrt((1 + ex) ** 2 * math.cos(phi) ** 2 +
                                                      (1 + ey) ** 2 * math.sin(phi) ** 2)))
    grow = a * np.exp(b * theta)
    u = ((n + ex * N) / N) * np.cos(theta) * grow
    v = ((n + ey * N) / N) * np.sin(theta) * grow
    return cube_faces(u, v, DOM)


def bruelkjaer(DOM, r0, Na, rmax, phi, d):
    """MATLAB/bruelkjaer.m:18-34: Na identical arms rotated by 2 pi / Na, each
    with one microphone per entry of d (0 <= d <= 1): a chord of length l d
    leaving the circle of radius r0 at angle phi, l chosen so that d = 1
    reaches rmax.  Arm-major order (all of arm 1, then arm 2, ...)."""
    d = np.asarray(d, dtype=np.float64).ravel()
    Na = int(Na)
    if Na < 1 or d.size == 0:
        raise ValueError("Na >= 1 and a non-empty d are required")
    Nr = d.size
    ell = r0 * math.cos(phi) + math.sqrt(rmax ** 2 - r0 ** 2 * math.sin(phi) ** 2)
    r = np.sqrt(r0 ** 2 + (ell * d) ** 2 - 2 * r0 * ell * d * math.cos(phi))
    arm = np.repeat(np.arange(Na, dtype=np.float64), Nr)
    rr = np.tile(r, Na)
    dd = np.tile(d, Na)
    theta = np.arcsin((ell * dd / rr) * math.sin(phi)) + (arm / Na) * 2 * math.pi
    return cube_faces(rr * np.cos(theta), rr * np.sin(theta), DOM)


def underbrink(DOM, r0, Na, Nm, rmax, v):
    """MATLAB/underbrink.m:18-30: Na log-spiral arms of Nm microphones, radii
    r0 and sqrt((2 j - 3)/(2 Nm - 3)) rmax for j = 2..Nm (equal-area rings),
    angle ln(r / r0) / cot(v) plus 2 pi (arm - 1) / Na.  Ring-major order
    (ring 1 of every arm, then ring 2, ...)."""
    Na, Nm = int(Na), int(Nm)
    if Na < 1 or Nm < 2:
        raise ValueError("Na >= 1 and Nm >= 2 are required")
    j = np.arange(2, Nm + 1, dtype=np.float64)
    radii = np.concatenate([[float(r0)], np.sqrt((2 * j - 3) / (2 * Nm - 3)) * rmax])
    rad = np.repeat(radii, Na)
    arm = np.tile(np.arange(Na, dtype=np.float64), Nm)
    thet = np.log(rad / r0) / (1.0 / math.tan(v)) + (arm / Na) * 2 * math.pi
    return cube_faces(rad * np.cos(thet), rad * np.sin(thet), DOM)
