"""Indirect BEM (collocation, linear triangles) and radiation-mode restatement.

Reference: ``MATLAB/ibemS.m:1-159`` (the loop-exact specification, with the
page-specific singular shape functions of ``CUDA/ibem.cuh`` rather than the
page-1 shortcut of ibemV.m:38,47) and ``CUDA/ibem.cuh:424-770`` (CPU
``ibem``); radiation modes ``MATLAB/invrS.m:1-170`` with R = real(G Z)
(``invrV.m:112``) and the Gram matrix of ``CUDA/invr.cuh:603-625``.

For element e with nodes (P0, P1, P2), n = (P0 - P2) x (P1 - P2) (not
normalised), quadrature points x_q = xi1 P0 + xi2 P1 + (1-xi1-xi2) P2 and
shape functions N = (xi1, xi2, 1-xi1-xi2), a row point p contributes
  B_a = (1/4pi) sum_q w_q N_a |n| e^{ikR}/R
  A_a = (1/4pi) sum_q w_q N_a F (ikR - 1) e^{ikR}/R,   F = n.(x_q - p)/R^2
  C   = (1/4pi) sum_q w_q (-F/R)                        (collocation rows only)
with the singular page of the collocated vertex when p is a node of e.
A_ss = sum A - diag(1 + C), B_ss = i rho c k sum B (+ CHIEF rows),
Z = A_ss \\ B_ss, H_b = A_hs Z - i rho c k B_hs.

Pinned by (tests/test_oracle_ibem.py): polynomial exactness of the
quadrature rules, the solid-angle identity of C on a closed surface, the
pulsating sphere (Z v and H_b v against the analytic pressure of
``CUDA/utils.cuh:384-412``), Gram-matrix row sums = nodal areas, and
symmetry / positive semi-definiteness of the radiation matrix.
"""
from __future__ import annotations

import math

import numpy as np

INV4PI = 1.0 / (4.0 * math.pi)


# ------------------------------------------------------------ quadrature
def gauss7():
    """7-point degree-5 rule on the reference triangle (Radon's formula,
    published in Stroud, "Approximate calculation of multiple integrals",
    1971, T2:5-1); rows (xi1, xi2, w) with weights summing to 1/2."""
    s = math.sqrt(15.0)
    a1, w1 = (6 - s) / 21, (155 - s) / 1200
    a2, w2 = (6 + s) / 21, (155 + s) / 1200
    pts = [(1 / 3, 1 / 3, 9 / 40)]
    for a, w in ((a1, w1), (a2, w2)):
        pts += [(a, a, w), (1 - 2 * a, a, w), (a, 1 - 2 * a, w)]
    q = np.array(pts)
    q[:, 2] *= 0.5
    return q


def load_ipf(path):
    """Regular table (CUDA/ibem.cuh:268-283): count line, then xi1 xi2 w."""
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    n = int(lines[0])
    return np.array([[float(v) for v in ln.split()] for ln in lines[1:n + 1]])


def load_ipfs(path):
    """Singular pages (CUDA/ibem.cuh:284-312): count line, $PAGE 1..3 sections."""
    lines = [ln for ln in open(path).read().splitlines() if ln.strip()]
    n = int(lines[0])
    pages, cur = [], None
    for ln in lines[1:]:
        if ln.startswith("$PAGE"):
            cur = []
            pages.append(cur)
        else:
            cur.append([float(v) for v in ln.split()])
    assert len(pages) == 3 and all(len(p) == n for p in pages)
    return np.array(pages)  # (3, n, 3)


def duffy_pages(ng=40):
    """Singular rules: Duffy transform of an ng x ng Gauss-Legendre product
    rule collapsing onto vertex v (page v), weights summing to 1/2.  Used
    when no singular table is supplied."""
    x, w = np.polynomial.legendre.leggauss(ng)
    u = 0.5 * (x + 1)
    wu = 0.5 * w
    U, V = np.meshgrid(u, u, indexing="ij")
    W = np.outer(wu, wu) * U
    # barycentric coordinates with the singular vertex at (1, 0, 0) of the page
    lam_s = 1 - U
    lam_a = U * V
    lam_b = U * (1 - V)
    pages = []
    for v in range(3):
        bary = np.zeros((ng * ng, 3))
        bary[:, v] = lam_s.ravel()
        bary[:, (v + 1) % 3] = lam_a.ravel()
        bary[:, (v + 2) % 3] = lam_b.ravel()
        pages.append(np.stack([bary[:, 0], bary[:, 1], W.ravel()], axis=1))
    return np.array(pages)


# ------------------------------------------------------------- geometry
def icosphere(subdiv, radius=1.0):
    """Icosahedron refined `subdiv` times by edge midpoints and projected to
    the sphere; triangles ordered so that (P0-P2) x (P1-P2) points outward.
    subdiv 3 gives 642 nodes and 1280 triangles."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                p = verts[i] + verts[j]
                verts.append(p / np.linalg.norm(p))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    X = radius * np.array(verts)
    E = np.array(f, dtype=np.int64)
    n = np.cross(X[E[:, 0]] - X[E[:, 2]], X[E[:, 1]] - X[E[:, 2]])
    out = np.einsum("ij,ij->i", n, X[E].mean(axis=1)) > 0
    E[~out] = E[~out][:, [1, 0, 2]]
    return X, E


# -------------------------------------------------------------- kernels
def _elem_rows(k, P, pts, rule):
    """Contributions of one element (P: 3x3 node coordinates) to rows at
    `pts` (r x 3) with `rule` (q x 3: xi1, xi2, w).  Returns A (r,3), B (r,3), C (r,)."""
    xi1, xi2, w = rule[:, 0], rule[:, 1], rule[:, 2]
    N = np.stack([xi1, xi2, 1 - xi1 - xi2], axis=1)               # (q,3)
    xq = N @ P                                                     # (q,3)
    n = np.cross(P[0] - P[2], P[1] - P[2])
    an = np.linalg.norm(n)
    d = xq[None, :, :] - pts[:, None, :]                           # (r,q,3)
    R = np.sqrt(np.sum(d * d, axis=2))
    G = np.exp(1j * k * R) / R
    F = (d @ n) / (R * R)
    Bq = INV4PI * G * an
    Aq = INV4PI * F * (1j * k * R - 1) * G
    A = (Aq * w[None, :]) @ N
    B = (Bq * w[None, :]) @ N
    C = INV4PI * ((-F / R) @ w)
    return A, B, C


def ibem_assemble(k, X, E, rule, pages, field=None, chief=None):
    """Return (A_ss, B_ss_unscaled, A_hs, B_hs_unscaled) -- B without the
    i rho c k factor -- following ibemS.m:18-150."""
    m = X.shape[0]
    A = np.zeros((m, m), complex)
    B = np.zeros((m, m), complex)
    Cd = np.zeros(m)
    nodes = np.arange(m)
    for e in range(E.shape[0]):
        en = E[e]
        P = X[en]
        mask = np.ones(m, bool)
        mask[en] = False
        Ae, Be, Ce = _elem_rows(k, P, X[mask], rule)
        rows = nodes[mask]
        A[np.ix_(rows, en)] += Ae
        B[np.ix_(rows, en)] += Be
        Cd[rows] += Ce
        for ikn in range(3):
            As, Bs, Cs = _elem_rows(k, P, X[en[ikn]][None, :], pages[ikn])
            A[en[ikn], en] += As[0]
            B[en[ikn], en] += Bs[0]
            Cd[en[ikn]] += Cs[0]
    A -= np.diag(1.0 + Cd)
    extra = []
    for pts in (chief, field):
        if pts is None or len(pts) == 0:
            extra.append((np.zeros((0, m), complex), np.zeros((0, m), complex)))
            continue
        pts = np.asarray(pts, float)
        Ah = np.zeros((len(pts), m), complex)
        Bh = np.zeros((len(pts), m), complex)
        for e in range(E.shape[0]):
            en = E[e]
            Ae, Be, _ = _elem_rows(k, X[en], pts, rule)
            Ah[:, en] += Ae
            Bh[:, en] += Be
        extra.append((Ah, Bh))
    (Ach, Bch), (Ahs, Bhs) = extra
    return np.vstack([A, Ach]), np.vstack([B, Bch]), Ahs, Bhs, Cd


def ibem(k, X, E, field, rule, pages, chief=None, rho=1.19, c=343.5):
    """CUDA/ibem.cuh:424 ``ibem``: returns (Z (m x m), H_b (l x m))."""
    A, B, Ahs, Bhs, _ = ibem_assemble(k, X, E, rule, pages, field, chief)
    s = 1j * rho * c * k
    B = s * B
    Z = np.linalg.lstsq(A, B, rcond=None)[0] if A.shape[0] > A.shape[1] else np.linalg.solve(A, B)
    H = Ahs @ Z - s * Bhs
    return Z, H


def gram(X, E, rule):
    """Mass matrix G_ab = sum_e sum_q w_q |n| N_a N_b (invrS.m:65, invr.cuh:603-625)."""
    m = X.shape[0]
    G = np.zeros((m, m))
    xi1, xi2, w = rule[:, 0], rule[:, 1], rule[:, 2]
    N = np.stack([xi1, xi2, 1 - xi1 - xi2], axis=1)
    for en in E:
        P = X[en]
        an = np.linalg.norm(np.cross(P[0] - P[2], P[1] - P[2]))
        G[np.ix_(en, en)] += (N * (w * an)[:, None]).T @ N
    return G


def invr(k, X, E, field, rule, pages, n_ma, chief=None, rho=1.19, c=343.5):
    """Radiation modes (invrS.m:157-170): R = real(G Z), U_r = svd(R),
    Q = U_r[:, :n_ma], M_r = Z Q, H_r = H_b Q.  Returns (Q, M_r, H_r, G, R)."""
    Z, H = ibem(k, X, E, field, rule, pages, chief, rho, c)
    G = gram(X, E, rule)
    R = np.real(G @ Z)
    Ur, _, _ = np.linalg.svd(R)
    n_ma = min(n_ma, X.shape[0])
    Q = Ur[:, :n_ma]
    return Q, Z @ Q, H @ Q, G, R


def radiation_resistance(pts, area, k, rho, c):
    """Elemental radiators: R_ij = (w^2 rho A^2/(4 pi c)) sin(k r_ij)/(k r_ij),
    R_ii = w^2 rho A^2/(4 pi c)."""
    w = k * c
    r = np.linalg.norm(pts[:, None, :] - pts[None, :, :], axis=2)
    return (w * w * rho * area * area / (4 * math.pi * c)) * np.sinc(k * r / math.pi)
