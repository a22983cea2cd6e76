"""Synthetic measurement scenarios (geometry + microphone pressures).

The reference ships no driver and no data set; its only known-answer source is
the analytic pulsating sphere of ``CUDA/utils.cuh:384-412`` (``sphere()``).
This module builds the seeded inputs every solve in this package is measured
on, so that the GPU path and the CPU checks see byte-identical inputs:

* ``sphere_scenario``  - pulsating + oscillating sphere (monopole amplitude V0,
  dipole amplitude V1 along z) radiating to Fibonacci-lattice microphones,
  equivalent sources on a retreated sphere (r_i * surface points, the
  reference's ESM geometry, ``CUDA/esm.cuh:9``).  Default sizes: 100 surface
  points on r=0.1 m, sources at 0.07 m, 64 microphones on r=0.2 m, 1 kHz.
* ``plate_scenario``   - baffled simply supported steel plate 0.6 x 0.5 x 0.001
  m driven by a unit point force; 32 x 32 microphone grid (0.03 m pitch) at
  z = 0.1 m, 64 x 32 equivalent sources at z = -d, 64 x 32 evaluation points
  at z = 0 with normal +z; pressures from the Rayleigh integral over a finer
  patch grid (avoids the inverse crime).
* ``radiation_scenario`` - the same plate as 64 x 64 elemental radiators for
  the radiation-mode sweep (50-2500 Hz, 20 dB SNR).

Conventions follow the reference: time dependence e^{-i w t}, outgoing
e^{+ikR} (``CUDA/esm.cuh:16-17``), spherical Hankel h_n = j_n + i y_n
(``CUDA/hels.cuh:109``).  Coordinates are returned as (n, 3) float64 arrays;
the C boundary takes them as structure-of-arrays (x block, y block, z block).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C_DEFAULT = 343.0
RHO_DEFAULT = 1.21


def fibonacci_sphere(n: int, radius: float) -> np.ndarray:
    """Fibonacci lattice: z_i = 1 - (2i+1)/n, phi_i = i*pi*(3-sqrt(5))."""
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / n
    rho = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return radius * np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)


def _sph_h_and_dh(n: int, x: np.ndarray):
    """h_n(x) = j_n + i y_n and its derivative, via scipy (input synthesis only)."""
    from scipy.special import spherical_jn, spherical_yn

    h = spherical_jn(n, x) + 1j * spherical_yn(n, x)
    dh = spherical_jn(n, x, derivative=True) + 1j * spherical_yn(n, x, derivative=True)
    return h, dh


def sphere_modal_pressure(k, a, pts, V0=1.0, V1=0.5, rho=RHO_DEFAULT, c=C_DEFAULT):
    """Exterior pressure of a sphere of radius a whose normal velocity is
    V0 + V1*cos(theta):  p = i rho c sum_n V_n P_n(cos th) h_n(kr)/h_n'(ka).

    For V1 = 0 this is exactly the reference's pulsating-sphere formula
    p = rho c u (a/r) (ika/(ika-1)) e^{ik(r-a)}  (``CUDA/utils.cuh:398,409``).
    """
    pts = np.asarray(pts, dtype=np.float64)
    r = np.linalg.norm(pts, axis=1)
    cth = pts[:, 2] / r
    p = np.zeros(len(r), dtype=np.complex128)
    for n, Vn, Pn in ((0, V0, np.ones_like(cth)), (1, V1, cth)):
        if Vn == 0.0:
            continue
        h_r, _ = _sph_h_and_dh(n, k * r)
        _, dh_a = _sph_h_and_dh(n, np.array([k * a]))
        p += 1j * rho * c * Vn * Pn * h_r / dh_a[0]
    return p


def sphere_modal_velocity(pts, V0=1.0, V1=0.5):
    pts = np.asarray(pts, dtype=np.float64)
    cth = pts[:, 2] / np.linalg.norm(pts, axis=1)
    return (V0 + V1 * cth).astype(np.complex128)


def sphere_modal_power(k, a, V0=1.0, V1=0.5, rho=RHO_DEFAULT, c=C_DEFAULT):
    """Closed-form radiated power 1/2 Re sum_n i rho c |V_n|^2 h_n/h_n' 4 pi a^2/(2n+1)."""
    W = 0.0
    for n, Vn in ((0, V0), (1, V1)):
        h, dh = _sph_h_and_dh(n, np.array([k * a]))
        W += 0.5 * np.real(1j * rho * c * abs(Vn) ** 2 * h[0] / dh[0]) * 4 * math.pi * a * a / (2 * n + 1)
    return float(W)


def add_noise(p: np.ndarray, snr_db: float, rng: np.random.Generator) -> np.ndarray:
    """p + sigma (g1 + i g2)/sqrt(2), sigma^2 = mean|p|^2 10^(-SNR/10)."""
    sig = math.sqrt(float(np.mean(np.abs(p) ** 2)) * 10.0 ** (-snr_db / 10.0))
    g = rng.standard_normal((2, p.shape[-1])) if p.ndim == 1 else rng.standard_normal((2,) + p.shape)
    return p + sig * (g[0] + 1j * g[1]) / math.sqrt(2.0)


@dataclass
class Scenario:
    """Inputs of one ESM sweep; ``b`` is (F, l) complex128 microphone pressures."""
    name: str
    freqs: np.ndarray          # (F,) Hz
    k: np.ndarray              # (F,) rad/m
    r_s: np.ndarray            # (m, 3) surface (evaluation) points
    r_h: np.ndarray            # (l, 3) microphones
    src: np.ndarray            # (m, 3) equivalent sources
    normals: np.ndarray | None # (m, 3) unit normals or None = radial (reference)
    r_i: float | None          # reference retreat factor when src = r_i * r_s
    b: np.ndarray              # (F, l) noisy pressures
    area: float                # total surface area S for W = 1/2 Re(v^H p) S/m
    rho: float = RHO_DEFAULT
    c: float = C_DEFAULT
    truth: dict = field(default_factory=dict)

    @property
    def m(self):
        return self.r_s.shape[0]

    @property
    def l(self):
        return self.r_h.shape[0]


def sphere_scenario(freqs=(1000.0,), m=100, l=64, a=0.1, r_mic=0.2, r_i=0.7,
                    V0=1.0, V1=0.5, snr_db=30.0, seed=1001, rho=RHO_DEFAULT, c=C_DEFAULT):
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    k = 2 * math.pi * freqs / c
    r_s = fibonacci_sphere(m, a)
    r_h = fibonacci_sphere(l, r_mic)
    rng = np.random.default_rng(seed)
    b = np.empty((len(freqs), l), dtype=np.complex128)
    p_true = np.empty_like(b)
    ps_true = np.empty((len(freqs), m), dtype=np.complex128)
    W_true = np.empty(len(freqs))
    for f, kf in enumerate(k):
        p_true[f] = sphere_modal_pressure(kf, a, r_h, V0, V1, rho, c)
        ps_true[f] = sphere_modal_pressure(kf, a, r_s, V0, V1, rho, c)
        W_true[f] = sphere_modal_power(kf, a, V0, V1, rho, c)
        b[f] = add_noise(p_true[f], snr_db, rng)
    return Scenario("sphere", freqs, k, r_s, r_h, r_i * r_s, None, r_i, b, 4 * math.pi * a * a,
                    rho, c, truth=dict(p_h=p_true, p_s=ps_true, v_s=sphere_modal_velocity(r_s, V0, V1),
                                       W=W_true))


# ---------------------------------------------------------------- plate
@dataclass
class PlateSpec:
    Lx: float = 0.6
    Ly: float = 0.5
    h: float = 0.001
    E: float = 210e9
    nu: float = 0.3
    rho_s: float = 7850.0
    eta: float = 0.01
    n_modes: int = 40


def plate_modes(spec: PlateSpec):
    """The n_modes lowest (p, q, omega_pq) of a simply supported plate."""
    Dp = spec.E * spec.h ** 3 / (12.0 * (1.0 - spec.nu ** 2))
    cands = []
    for p in range(1, 40):
        for q in range(1, 40):
            w = math.sqrt(Dp / (spec.rho_s * spec.h)) * ((p * math.pi / spec.Lx) ** 2 + (q * math.pi / spec.Ly) ** 2)
            cands.append((w, p, q))
    cands.sort()
    return [(p, q, w) for (w, p, q) in cands[: spec.n_modes]]


def plate_velocity(spec: PlateSpec, omega: float, x0, y0, x, y):
    """Normal velocity of the point-driven plate (unit force at (x0, y0)):
    v = -i w sum_pq phi(x0,y0) phi(x,y) / (M (w_pq^2 (1 + i eta) - w^2)),
    phi = sin(p pi x/Lx) sin(q pi y/Ly), M = rho_s h Lx Ly / 4."""
    M = spec.rho_s * spec.h * spec.Lx * spec.Ly / 4.0
    v = np.zeros(np.shape(x), dtype=np.complex128)
    for p, q, wpq in plate_modes(spec):
        phi0 = math.sin(p * math.pi * x0 / spec.Lx) * math.sin(q * math.pi * y0 / spec.Ly)
        phi = np.sin(p * math.pi * x / spec.Lx) * np.sin(q * math.pi * y / spec.Ly)
        v += phi0 * phi / (M * (wpq ** 2 * (1 + 1j * spec.eta) - omega ** 2))
    return -1j * omega * v


def grid_points(nx, ny, Lx, Ly, z, centred=True):
    """Cell-centred nx x ny grid on [0,Lx]x[0,Ly] at height z, x fastest."""
    xs = (np.arange(nx) + 0.5) * Lx / nx
    ys = (np.arange(ny) + 0.5) * Ly / ny
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1)


def rayleigh_pressure(k, omega, rho, src_pts, v_src, dA, fld_pts, device="cpu"):
    """p(r) = -(i w rho / 2 pi) sum_j v_j e^{ikR}/R dA  (baffled plane, e^{-iwt}).

    The sign matches the reference's pulsating sphere, whose small-ka limit is
    p = -i w rho Q e^{ikr}/(4 pi r) for volume velocity Q (utils.cuh:398)."""
    import torch

    S = torch.as_tensor(src_pts, dtype=torch.float64, device=device)
    Fp = torch.as_tensor(fld_pts, dtype=torch.float64, device=device)
    v = torch.as_tensor(v_src, dtype=torch.complex128, device=device)
    out = torch.empty(len(fld_pts), dtype=torch.complex128, device=device)
    step = 256
    for i0 in range(0, len(fld_pts), step):
        R = torch.cdist(Fp[i0:i0 + step], S)
        G = torch.polar(1.0 / R, k * R)
        out[i0:i0 + step] = G @ v
    return (-1j * omega * rho / (2 * math.pi) * dA * out).cpu().numpy()


def plate_scenario(freqs=None, n_freq=256, fmin=200.0, fmax=1500.0, d=0.02,
                   mic_n=(32, 32), mic_pitch=0.03, mic_z=0.1, src_n=(64, 32),
                   ref_n=(128, 100), snr_db=30.0, seed=1004, spec: PlateSpec | None = None,
                   rho=RHO_DEFAULT, c=C_DEFAULT, device="cpu"):
    spec = spec or PlateSpec()
    if freqs is None:
        freqs = np.linspace(fmin, fmax, n_freq)
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    k = 2 * math.pi * freqs / c
    rng = np.random.default_rng(seed)
    x0 = float(rng.uniform(0.1, 0.9)) * spec.Lx
    y0 = float(rng.uniform(0.1, 0.9)) * spec.Ly
    # microphones: mic_n grid with mic_pitch spacing centred over the plate
    mx = (np.arange(mic_n[0]) - (mic_n[0] - 1) / 2) * mic_pitch + spec.Lx / 2
    my = (np.arange(mic_n[1]) - (mic_n[1] - 1) / 2) * mic_pitch + spec.Ly / 2
    MX, MY = np.meshgrid(mx, my, indexing="xy")
    r_h = np.stack([MX.ravel(), MY.ravel(), np.full(MX.size, mic_z)], axis=1)
    r_s = grid_points(src_n[0], src_n[1], spec.Lx, spec.Ly, 0.0)
    src = r_s.copy()
    src[:, 2] = -d
    normals = np.zeros_like(r_s)
    normals[:, 2] = 1.0
    ref = grid_points(ref_n[0], ref_n[1], spec.Lx, spec.Ly, 0.0)
    dA = spec.Lx * spec.Ly / len(ref)
    b = np.empty((len(freqs), len(r_h)), dtype=np.complex128)
    v_true = np.empty((len(freqs), len(r_s)), dtype=np.complex128)
    for f, kf in enumerate(k):
        w = 2 * math.pi * freqs[f]
        v_ref = plate_velocity(spec, w, x0, y0, ref[:, 0], ref[:, 1])
        p = rayleigh_pressure(kf, w, rho, ref, v_ref, dA, r_h, device)
        b[f] = add_noise(p, snr_db, rng)
        v_true[f] = plate_velocity(spec, w, x0, y0, r_s[:, 0], r_s[:, 1])
    return Scenario("plate", freqs, k, r_s, r_h, src, normals, None, b, spec.Lx * spec.Ly,
                    rho, c, truth=dict(v_s=v_true, force_at=(x0, y0)))


@dataclass
class RadiationScenario:
    """Inputs of one radiation-mode sweep: patch centres (np, 3) of equal
    area, microphones (l, 3), pressures b (F, l)."""
    freqs: np.ndarray
    k: np.ndarray
    r_p: np.ndarray
    r_h: np.ndarray
    b: np.ndarray
    patch_area: float
    rho: float = RHO_DEFAULT
    c: float = C_DEFAULT
    truth: dict = field(default_factory=dict)


def radiation_scenario(freqs=None, n_freq=256, fmin=50.0, fmax=2500.0, patches=(64, 64),
                       mic_n=(32, 32), mic_pitch=0.03, mic_z=0.1, ref_n=(128, 100), snr_db=20.0,
                       seed=1005, spec: PlateSpec | None = None, rho=RHO_DEFAULT, c=C_DEFAULT,
                       device="cpu"):
    """Baffled plate of patches x elemental radiators (the plate of
    ``plate_scenario``), microphones on a grid above it, pressures from the
    Rayleigh integral over a finer grid plus seeded noise at snr_db; the
    patch velocities of the modal model are returned as truth."""
    spec = spec or PlateSpec()
    if freqs is None:
        freqs = np.linspace(fmin, fmax, n_freq)
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    k = 2 * math.pi * freqs / c
    rng = np.random.default_rng(seed)
    x0 = float(rng.uniform(0.1, 0.9)) * spec.Lx
    y0 = float(rng.uniform(0.1, 0.9)) * spec.Ly
    mx = (np.arange(mic_n[0]) - (mic_n[0] - 1) / 2) * mic_pitch + spec.Lx / 2
    my = (np.arange(mic_n[1]) - (mic_n[1] - 1) / 2) * mic_pitch + spec.Ly / 2
    MX, MY = np.meshgrid(mx, my, indexing="xy")
    r_h = np.stack([MX.ravel(), MY.ravel(), np.full(MX.size, mic_z)], axis=1)
    r_p = grid_points(patches[0], patches[1], spec.Lx, spec.Ly, 0.0)
    ref = grid_points(ref_n[0], ref_n[1], spec.Lx, spec.Ly, 0.0)
    dA = spec.Lx * spec.Ly / len(ref)
    b = np.empty((len(freqs), len(r_h)), dtype=np.complex128)
    v_p = np.empty((len(freqs), len(r_p)), dtype=np.complex128)
    for f, kf in enumerate(k):
        w = 2 * math.pi * freqs[f]
        v_ref = plate_velocity(spec, w, x0, y0, ref[:, 0], ref[:, 1])
        p = rayleigh_pressure(kf, w, rho, ref, v_ref, dA, r_h, device)
        b[f] = add_noise(p, snr_db, rng)
        v_p[f] = plate_velocity(spec, w, x0, y0, r_p[:, 0], r_p[:, 1])
    return RadiationScenario(freqs, k, r_p, r_h, b, spec.Lx * spec.Ly / len(r_p), rho, c,
                             truth=dict(v_p=v_p, force_at=(x0, y0)))


def soa(pts: np.ndarray) -> np.ndarray:
    """(n,3) -> contiguous x-block, y-block, z-block (the reference's layout)."""
    return np.ascontiguousarray(np.asarray(pts, dtype=np.float64).T).ravel()


# ---------------------------------------------------------------- HELS
@dataclass
class HelsScenario:
    """Inputs of a HELS sweep: spherical coordinates (az, el, r) as
    ``cartesian_2_spherical`` returns them, per-frequency expansion order."""
    freqs: np.ndarray
    k: np.ndarray
    order: np.ndarray
    azelr_s: np.ndarray   # (m, 3) reconstruction surface
    azelr_h: np.ndarray   # (l, 3) microphones
    b: np.ndarray         # (F, l)
    area: float
    rho: float = RHO_DEFAULT
    c: float = C_DEFAULT
    truth: dict = field(default_factory=dict)


def to_spherical(pts):
    """(az in [0, 2 pi), el = polar angle (minus 2 pi where x < 0), r) -- the
    reference's cartesian_2_spherical (CUDA/utils.cuh:414-427)."""
    p = np.asarray(pts, np.float64)
    xy = p[:, 0] ** 2 + p[:, 1] ** 2
    r = np.sqrt(xy + p[:, 2] ** 2)
    az = np.arctan2(p[:, 1], p[:, 0])
    az = np.where(p[:, 1] < 0, az + 2 * math.pi, az)
    el = np.arctan2(np.sqrt(xy), p[:, 2])
    el = np.where(p[:, 0] < 0, el - 2 * math.pi, el)
    return np.stack([az, el, r], axis=1)


def hels_scenario(freqs=None, n_freq=128, fmin=50.0, fmax=3000.0, a=0.12, n_src=5, l=256,
                  r_mic=0.3, m=256, nmax=16, snr_db=30.0, seed=1003, rho=RHO_DEFAULT, c=C_DEFAULT):
    """n_src seeded monopoles inside r < 0.6 a with CN(0,1) amplitudes,
    l Fibonacci microphones at r_mic, reconstruction on m Fibonacci points at
    r = a, expansion order N_f = min(nmax, ceil(k a) + 4)."""
    if freqs is None:
        freqs = np.linspace(fmin, fmax, n_freq)
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    k = 2 * math.pi * freqs / c
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n_src, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    src = d * (0.6 * a * rng.uniform(0, 1, (n_src, 1)) ** (1 / 3))
    amp = (rng.standard_normal(n_src) + 1j * rng.standard_normal(n_src)) / math.sqrt(2)
    r_h = fibonacci_sphere(l, r_mic)
    r_s = fibonacci_sphere(m, a)
    order = np.minimum(nmax, np.ceil(k * a).astype(int) + 4)
    b = np.empty((len(freqs), l), np.complex128)
    for f, kf in enumerate(k):
        R = np.linalg.norm(r_h[:, None, :] - src[None, :, :], axis=2)
        p = (1j * rho * c * kf * np.exp(1j * kf * R) / (4 * math.pi * R)) @ amp
        b[f] = add_noise(p, snr_db, rng)
    return HelsScenario(freqs, k, order, to_spherical(r_s), to_spherical(r_h), b,
                        4 * math.pi * a * a, rho, c, truth=dict(src=src, amp=amp))


# ----------------------------------------------------------------- IBEM
def icosphere(subdiv: int, radius: float = 1.0):
    """Icosahedron refined `subdiv` times by edge midpoints projected onto the
    sphere: (nodes (m, 3), triangles (n, 3) 0-based), each triangle ordered
    so that (P0 - P2) x (P1 - P2) points outward (the reference's normal,
    ``MATLAB/ibemS.m:39``).  subdiv 3: 642 nodes, 1280 triangles."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [np.array(p, float) / math.sqrt(1 + t * t) for p in
             ((-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
              (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1))]
    tris = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
            (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
            (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        mids: dict = {}

        def mid(i, j):
            key = (i, j) if i < j else (j, i)
            if key not in mids:
                p = verts[i] + verts[j]
                verts.append(p / np.linalg.norm(p))
                mids[key] = len(verts) - 1
            return mids[key]

        nxt = []
        for a, b, c in tris:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nxt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        tris = nxt
    X = radius * np.array(verts)
    E = np.array(tris, dtype=np.int64)
    nrm = np.cross(X[E[:, 0]] - X[E[:, 2]], X[E[:, 1]] - X[E[:, 2]])
    inward = np.einsum("ij,ij->i", nrm, X[E].mean(axis=1)) < 0
    E[inward] = E[inward][:, [1, 0, 2]]
    return X, E


@dataclass
class IbemScenario:
    """Inputs of one IBEM sweep on a closed triangulated surface."""
    freqs: np.ndarray
    k: np.ndarray
    X: np.ndarray              # (m, 3) nodes
    E: np.ndarray              # (n, 3) 0-based triangles, outward (P0-P2)x(P1-P2)
    r_h: np.ndarray            # (l, 3) microphones
    chief: np.ndarray          # (n_ch, 3) interior CHIEF points (may be empty)
    b: np.ndarray              # (F, l)
    area: float
    rho: float
    c: float
    truth: dict = field(default_factory=dict)


def ibem_scenario(freqs=None, n_freq=64, fmin=100.0, fmax=2000.0, subdiv=3, a=0.1, l=128,
                  r_mic=0.25, V0=1.0, V1=0.5, snr_db=30.0, n_chief=0, seed=1002,
                  rho=RHO_DEFAULT, c=C_DEFAULT):
    """Pulsating + oscillating sphere (normal velocity V0 + V1 cos(theta)) on an
    icosphere of radius a, l Fibonacci microphones at r_mic with seeded
    complex Gaussian noise; n_chief seeded interior CHIEF points at r < 0.6 a."""
    if freqs is None:
        freqs = np.linspace(fmin, fmax, n_freq)
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    k = 2 * math.pi * freqs / c
    X, E = icosphere(subdiv, a)
    r_h = fibonacci_sphere(l, r_mic)
    rng = np.random.default_rng(seed)
    chief = np.zeros((0, 3))
    if n_chief:
        d = rng.standard_normal((n_chief, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        chief = d * (0.6 * a * rng.uniform(0.2, 1.0, (n_chief, 1)))
    b = np.empty((len(freqs), l), np.complex128)
    W = np.empty(len(freqs))
    for f, kf in enumerate(k):
        b[f] = add_noise(sphere_modal_pressure(kf, a, r_h, V0, V1, rho, c), snr_db, rng)
        W[f] = sphere_modal_power(kf, a, V0, V1, rho, c)
    # polyhedron area: the sum of the flat triangles
    S = 0.5 * float(np.sum(np.linalg.norm(np.cross(X[E[:, 0]] - X[E[:, 2]],
                                                    X[E[:, 1]] - X[E[:, 2]]), axis=1)))
    return IbemScenario(freqs, k, X, E, r_h, chief, b, S, rho, c,
                        truth=dict(v_s=sphere_modal_velocity(X, V0, V1), W=W))
