"""Pins of the IBEM / radiation-mode oracle (oracle/ibem.py) against what the
mathematics fixes: quadrature exactness, closed-form singular integrals,
exact solid angles, Gauss's law, the analytic pulsating sphere, the closed
form of the linear-triangle mass matrix and the monopole radiation mode.
CPU only."""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import ibem as ib
from oracle import post

GOLD = Path(__file__).parent / "golden"
RHO, C = 1.21, 343.0


def _mono(a, b):
    """Integral of xi1^a xi2^b over the reference triangle."""
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


def _exactness(rule, deg):
    return max(abs(np.sum(rule[:, 2] * rule[:, 0] ** a * rule[:, 1] ** (d - a)) - _mono(a, d - a))
               for d in range(deg + 1) for a in range(d + 1))


def test_gauss7_degree5():
    q = ib.gauss7()
    assert _exactness(q, 5) < 1e-15
    assert _exactness(q, 6) > 1e-6  # exactly degree 5


def test_reference_rule_degree7():
    q = ib.load_ipf(GOLD / "IPF.txt")
    assert q.shape == (13, 3)
    assert _exactness(q, 7) < 1e-15


def _vertex_integral(P, v):
    """Closed form of the integral of 1/|x - P_v| over the triangle P: in
    polar coordinates about P_v the opposite edge lies at distance h and
    the integral is h ln((|B| + t_B)/(|A| + t_A)), t the coordinate along
    the edge from the foot of the perpendicular."""
    A, B = P[(v + 1) % 3] - P[v], P[(v + 2) % 3] - P[v]
    e = (B - A) / np.linalg.norm(B - A)
    foot = A - np.dot(A, e) * e
    h = np.linalg.norm(foot)
    tA, tB = np.dot(A, e), np.dot(B, e)
    return h * math.log((np.linalg.norm(B) + tB) / (np.linalg.norm(A) + tA))


def _page_integral(P, page, v):
    N = np.stack([page[:, 0], page[:, 1], 1 - page[:, 0] - page[:, 1]], axis=1)
    x = N @ P
    an = np.linalg.norm(np.cross(P[0] - P[2], P[1] - P[2]))
    return np.sum(page[:, 2] * an / np.linalg.norm(x - P[v], axis=1))


@pytest.mark.parametrize("src", ["duffy", "reference"])
def test_singular_pages(src):
    pages = ib.duffy_pages(40) if src == "duffy" else ib.load_ipfs(GOLD / "IPFs.txt")
    tol = 1e-8 if src == "duffy" else 1e-11
    rng = np.random.default_rng(3)
    for _ in range(4):
        P = rng.normal(size=(3, 3))
        for v in range(3):
            ref = _vertex_integral(P, v)
            err = abs(_page_integral(P, pages[v], v) - ref)
            assert err < tol * abs(ref)
            # the page of another vertex resolves this singularity far worse
            wrong = abs(_page_integral(P, pages[(v + 1) % 3], v) - ref)
            assert wrong > 100 * err
        for pg in pages:
            assert abs(pg[:, 2].sum() - 0.5) < 1e-13


def _solid_angle(p, P):
    """Van Oosterom-Strackee solid angle of triangle P seen from p."""
    a, b, c = P[0] - p, P[1] - p, P[2] - p
    la, lb, lc = (np.linalg.norm(x) for x in (a, b, c))
    num = np.dot(a, np.cross(b, c))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2 * math.atan2(num, den)


def test_c_term_is_solid_angle():
    X, E = ib.icosphere(1)
    *_, Cd = ib.ibem_assemble(1.0, X, E, ib.gauss7(), ib.duffy_pages(20))
    for o in range(0, len(X), 7):
        ref = -sum(_solid_angle(X[o], X[en]) for en in E if o not in en) / (4 * math.pi)
        assert abs(Cd[o] - ref) < 2e-4
        assert -0.5 < Cd[o] < -0.3


def test_gauss_law_interior():
    X, E = ib.icosphere(2)
    tot = sum(ib._elem_rows(1.0, X[en], np.array([[0.05, -0.02, 0.03]]), ib.gauss7())[2][0]
              for en in E)
    assert abs(tot + 1.0) < 1e-5


def _pulsating_errors(subdiv, pages, rule, k=2.0, a=1.0):
    X, E = ib.icosphere(subdiv, a)
    field = ib.icosphere(0, 2.5)[0]
    Z, H = ib.ibem(k, X, E, field, rule, pages, rho=RHO, c=C)
    u = 1e-3
    p_s, p_h, _ = post.sphere(X, field, k, u, a, rho=RHO, c=C)
    v = np.full(len(X), u, complex)
    es = np.linalg.norm(Z @ v - p_s) / np.linalg.norm(p_s)
    eh = np.linalg.norm(H @ v - p_h) / np.linalg.norm(p_h)
    W = post.sound_power(Z @ v, v, 4 * math.pi * a * a)
    ka = k * a
    W_ref = 2 * math.pi * a * a * RHO * C * u * u * ka * ka / (1 + ka * ka)
    return es, eh, abs(W - W_ref) / W_ref


def test_pulsating_sphere_converges():
    rule, pages = ib.gauss7(), ib.duffy_pages(20)
    e1 = _pulsating_errors(1, pages, rule)
    e2 = _pulsating_errors(2, pages, rule)
    assert max(e2) < 0.02
    assert all(b < a / 3 for a, b in zip(e1, e2))  # O(h^2)


def test_pulsating_sphere_reference_tables():
    rule, pages = ib.load_ipf(GOLD / "IPF.txt"), ib.load_ipfs(GOLD / "IPFs.txt")
    assert max(_pulsating_errors(2, pages, rule)) < 0.02


def test_chief_least_squares_consistent():
    X, E = ib.icosphere(1)
    cp = np.array([[0.1, 0.2, -0.1], [-0.2, 0.0, 0.15]])
    field = ib.icosphere(0, 2.0)[0]
    rule, pages = ib.gauss7(), ib.duffy_pages(16)
    Z0, H0 = ib.ibem(2.0, X, E, field, rule, pages)
    Z1, H1 = ib.ibem(2.0, X, E, field, rule, pages, chief=cp)
    # away from interior resonances the CHIEF rows are nearly consistent
    assert np.linalg.norm(Z1 - Z0) < 0.05 * np.linalg.norm(Z0)


def test_gram_closed_form():
    X, E = ib.icosphere(1, 0.7)
    G = ib.gram(X, E, ib.gauss7())
    ref = np.zeros_like(G)
    M = (np.ones((3, 3)) + np.eye(3)) / 12.0
    for en in E:
        A = 0.5 * np.linalg.norm(np.cross(X[en[0]] - X[en[2]], X[en[1]] - X[en[2]]))
        ref[np.ix_(en, en)] += A * M
    assert np.max(np.abs(G - ref)) < 1e-15
    assert np.allclose(G, G.T)
    areas = np.zeros(len(X))
    for en in E:
        areas[en] += 0.5 * np.linalg.norm(np.cross(X[en[0]] - X[en[2]], X[en[1]] - X[en[2]])) / 3
    assert np.allclose(G.sum(axis=1), areas, rtol=1e-13)


def test_radiation_modes_monopole_leads():
    X, E = ib.icosphere(2)
    field = ib.icosphere(0, 2.0)[0]
    Q, Mr, Hr, G, R = ib.invr(0.5, X, E, field, ib.gauss7(), ib.duffy_pages(16), 6,
                              rho=RHO, c=C)
    assert Q.shape == (len(X), 6) and Hr.shape == (len(field), 6)
    asym = np.linalg.norm(R - R.T) / np.linalg.norm(R)
    assert asym < 0.05
    ev = np.linalg.eigvalsh(0.5 * (R + R.T))
    assert ev.min() > -0.02 * ev.max()
    q0 = Q[:, 0].real
    assert abs(q0 @ np.ones(len(X))) / np.sqrt(len(X)) > 0.99  # uniform at low ka


def test_elemental_radiator_low_frequency():
    rng = np.random.default_rng(0)
    pts = rng.uniform(-0.01, 0.01, size=(20, 3))
    k, A = 0.1, 1e-4
    R = ib.radiation_resistance(pts, A, k, RHO, C)
    scale = (k * C) ** 2 * RHO * A * A / (4 * math.pi * C)
    assert np.allclose(np.diag(R), scale)
    # sources much closer than a wavelength radiate as one monopole: R ~ scale * 1 1^T
    ev = np.linalg.eigvalsh(R)
    assert abs(ev[-1] / (20 * scale) - 1) < 1e-4
    assert np.all(ev[:-1] < 1e-3 * ev[-1])
