"""GPU IBEM and radiation-mode paths against the oracle (ibemS.m / invrS.m
restatement in oracle/ibem.py), with the built-in rules and with the
reference's own quadrature tables."""
from pathlib import Path

import numpy as np
import pytest

from inverseacoustics_amd import scenarios as sc
from oracle import ibem as ib
from oracle import pipeline, post

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).parent / "golden"


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture
def builtin_rules(ia):
    ia.set_quadrature()
    yield ib.gauss7(), ib.duffy_pages(40)
    ia.set_quadrature()


@pytest.fixture
def reference_rules(ia):
    rule, pages = ia.quadrature_load(GOLD / "IPF.txt", GOLD / "IPFs.txt")
    ia.set_quadrature(rule, pages)
    yield rule, pages
    ia.set_quadrature()


@pytest.mark.parametrize("subdiv", [1, 2])
@pytest.mark.parametrize("k", [3.0, 40.0])
def test_ibem_matrices_builtin_rules(ia, builtin_rules, subdiv, k):
    rule, pages = builtin_rules
    X, E = sc.icosphere(subdiv, 0.1)
    field = sc.fibonacci_sphere(24, 0.25)
    Z, H = ia.ibem(X, E, field, k)
    Z0, H0 = ib.ibem(k, X, E, field, rule, pages, rho=1.21, c=343.0)
    assert _rel(Z, Z0) < 1e-10 and _rel(H, H0) < 1e-10


def test_ibem_matrices_reference_tables(ia, reference_rules):
    rule, pages = reference_rules
    X, E = sc.icosphere(2, 0.1)
    field = sc.fibonacci_sphere(17, 0.3)
    Z, H = ia.ibem(X, E, field, 20.0)
    Z0, H0 = ib.ibem(20.0, X, E, field, rule, pages, rho=1.21, c=343.0)
    assert _rel(Z, Z0) < 1e-10 and _rel(H, H0) < 1e-10


def test_ibem_chief_least_squares(ia, builtin_rules):
    rule, pages = builtin_rules
    S = sc.ibem_scenario(n_freq=1, subdiv=2, l=16, n_chief=8)
    Z, H = ia.ibem(S.X, S.E, S.r_h, S.k[0], cp=S.chief)
    Z0, H0 = ib.ibem(S.k[0], S.X, S.E, S.r_h, rule, pages, chief=S.chief, rho=S.rho, c=S.c)
    assert _rel(Z, Z0) < 1e-10 and _rel(H, H0) < 1e-10


def test_ibem_reproduces_pulsating_sphere(ia):
    X, E = sc.icosphere(3, 0.1)
    field = sc.fibonacci_sphere(32, 0.25)
    k = 2 * np.pi * 800 / 343.0
    Z, H = ia.ibem(X, E, field, k)
    p_s, p_h, _ = post.sphere(X, field, k, 1.0, 0.1, rho=1.21, c=343.0)
    v = np.ones(len(X), complex)
    assert np.linalg.norm(Z @ v - p_s) < 5e-3 * np.linalg.norm(p_s)
    assert np.linalg.norm(H @ v - p_h) < 5e-3 * np.linalg.norm(p_h)


def test_ibem_legacy_entry_points(ia):
    import ctypes as C

    from inverseacoustics_amd import _lib

    X, E = sc.icosphere(1, 0.1)
    field = sc.fibonacci_sphere(8, 0.2)
    rs, rh = _lib.soa(X), _lib.soa(field)
    en = np.ascontiguousarray(E.T.astype(np.int32)).ravel()
    me, he = _lib._zp(), _lib._zp()
    ia.set_air(343.5, 1.19)
    _lib.lib().ibem(len(X), len(field), len(E), _lib.dptr(rs), en.ctypes.data_as(_lib._ip),
                    _lib.dptr(rh), C.byref(me), C.byref(he), 10.0, 0, None)
    Z = _lib.take_z(me, len(X) ** 2, (len(X), len(X)))
    H = _lib.take_z(he, len(field) * len(X), (len(field), len(X)))
    Z0, H0 = ib.ibem(10.0, X, E, field, ib.gauss7(), ib.duffy_pages(40))  # reference constants
    assert _rel(Z, Z0) < 1e-10 and _rel(H, H0) < 1e-10


def test_ibem_sweep_matches_oracle(ia, builtin_rules):
    rule, pages = builtin_rules
    S = sc.ibem_scenario(n_freq=4, fmin=300.0, fmax=1800.0, subdiv=2, l=64)
    ia.set_air(S.c, S.rho)
    r = ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area)
    for f in range(len(S.k)):
        o = pipeline.ibem_solve(S.k[f], S.X, S.E, S.r_h, S.b[f], S.area, rule, pages,
                                rho=S.rho, c=S.c)
        assert r.lam_index[f] == o["idx"] and r.k_tsvd[f] == o["k_tsvd"]
        assert np.linalg.norm(r.velocity[f] - o["v"]) <= 1e-10 * np.linalg.norm(o["v"])
        assert np.linalg.norm(r.pressure[f] - o["p"]) <= 1e-10 * np.linalg.norm(o["p"])
        assert r.power[f] == pytest.approx(o["W"], rel=1e-10)
    # the reconstruction tracks the analytic source
    assert np.max(np.abs(post.power_db(r.power) - post.power_db(S.truth["W"]))) < 1.0
    rm = ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area, precision="mixed")
    assert np.max(np.abs(post.power_db(rm.power) - post.power_db(r.power))) < 0.1


def test_ibem_sweep_chunked_and_chief(ia, builtin_rules):
    S = sc.ibem_scenario(n_freq=5, subdiv=1, l=24, n_chief=4)
    ia.set_air(S.c, S.rho)
    a = ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area, cp=S.chief)
    b = ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area, cp=S.chief, chunk=2)
    assert np.array_equal(a.lam_index, b.lam_index)
    assert np.abs(a.velocity - b.velocity).max() <= 1e-12 * np.abs(a.velocity).max()


def test_ibem_sweep_workspace_reuse(ia, builtin_rules):
    """The sweep keeps its workspace between calls: a smaller sweep in
    between leaves a repeated sweep bit for bit equal to the first."""
    S = sc.ibem_scenario(n_freq=5, subdiv=1, l=24)
    ia.set_air(S.c, S.rho)
    a = ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area)
    ia.ibem_sweep(S.k[:2], S.X, S.E, S.r_h, S.b[:2], S.area, precision="mixed")
    c = ia.ibem_sweep(S.k, S.X, S.E, S.r_h, S.b, S.area)
    np.testing.assert_array_equal(c.velocity, a.velocity)
    np.testing.assert_array_equal(c.power, a.power)


def test_ibem_bad_inputs(ia):
    X, E = sc.icosphere(1, 0.1)
    bad = E.copy()
    bad[3, 1] = len(X)  # node out of range
    with pytest.raises(ia.IAError):
        ia.ibem(X, bad, X[:4] * 2, 5.0)
    with pytest.raises(ia.IAError):
        ia.ibem(X, E, X[:4] * 2, -1.0)


def test_invr_matches_oracle(ia, builtin_rules):
    rule, pages = builtin_rules
    X, E = sc.icosphere(2, 0.1)
    field = sc.fibonacci_sphere(20, 0.25)
    k, n_am = 10.0, 8
    Q, Mr, Hr, G = ia.invr(X, E, field, k, n_am)
    Q0, Mr0, Hr0, G0, R0 = ib.invr(k, X, E, field, rule, pages, n_am, rho=1.21, c=343.0)
    assert np.abs(G - G0).max() < 1e-14 * np.abs(G0).max()
    s0 = np.linalg.svd(R0, compute_uv=False)
    # columns are singular vectors up to a unimodular factor; compare where the
    # singular value is separated from its neighbours
    Z0, H0 = ib.ibem(k, X, E, field, rule, pages, rho=1.21, c=343.0)
    for j in range(n_am):
        gap = min(abs(s0[j] - s0[i]) for i in range(len(s0)) if i != j) / s0[0]
        assert abs(np.linalg.norm(Q[:, j]) - 1) < 1e-12
        assert np.linalg.norm(R0 @ (R0.T @ Q[:, j]) - s0[j] ** 2 * Q[:, j]) < 1e-9 * s0[0] ** 2
        if gap > 1e-6:
            assert abs(abs(np.vdot(Q0[:, j], Q[:, j])) - 1) < 1e-8
    assert np.abs(Mr - Z0 @ Q).max() < 1e-10 * np.abs(Mr).max()
    assert np.abs(Hr - H0 @ Q).max() < 1e-10 * np.abs(Hr).max()


def test_radiation_modes_match_eigh(ia):
    pts = np.stack(np.meshgrid(np.linspace(0, 0.6, 12), np.linspace(0, 0.5, 10), [0.0],
                               indexing="ij"), -1).reshape(-1, 3)
    A = (0.6 / 12) * (0.5 / 10)
    k = 2 * np.pi * 900 / 343.0
    ia.set_air(343.0, 1.21)
    ev, modes = ia.radiation_modes(pts, A, k, 10)
    R = ib.radiation_resistance(pts, A, k, 1.21, 343.0)
    w, V = np.linalg.eigh(R)
    w, V = w[::-1], V[:, ::-1]
    assert np.allclose(ev, w[:10], rtol=1e-11, atol=1e-13 * w[0])
    for j in range(10):
        if min(abs(w[j] - w[i]) for i in range(len(w)) if i != j) > 1e-6 * w[0]:
            assert abs(abs(V[:, j] @ modes[:, j]) - 1) < 1e-8
