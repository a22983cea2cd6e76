"""Pins of the regularisation oracle (oracle/regtools.py) against library
routines (LAPACK via numpy/scipy, scipy's bounded Brent) and brute force."""
import numpy as np
import pytest
import scipy.linalg as sl
from scipy.optimize import fminbound

from oracle import regtools as rt


def _rand(m, n, seed=0, decay=None):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    if decay is not None:
        U, s, Vh = np.linalg.svd(A, full_matrices=False)
        s = 10.0 ** (-decay * np.arange(len(s)) / len(s))
        A = (U * s) @ Vh
    b = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    return A, b


@pytest.mark.parametrize("shape", [(30, 20), (20, 30), (25, 25)])
def test_csvd_against_gesvd(shape):
    A, _ = _rand(*shape, seed=3)
    U, s, V = rt.csvd(A)
    U2, s2, Vh2 = sl.svd(A, full_matrices=False, lapack_driver="gesvd")
    p = min(shape)
    assert U.shape == (shape[0], p) and V.shape == (shape[1], p)
    np.testing.assert_allclose(s, s2, rtol=1e-13, atol=1e-14 * s2[0])
    np.testing.assert_allclose((U * s) @ V.conj().T, A, atol=1e-13 * s2[0])
    # phase-invariant: projectors agree
    np.testing.assert_allclose(U @ U.conj().T, U2 @ U2.conj().T, atol=1e-12)


def test_tsvd_full_rank_is_least_squares():
    A, b = _rand(30, 20, seed=4)
    U, s, V = rt.csvd(A)
    x = rt.tsvd(U, s, V, b, 20)
    np.testing.assert_allclose(x, np.linalg.lstsq(A, b, rcond=None)[0], rtol=1e-11)
    assert np.all(rt.tsvd(U, s, V, b, 0) == 0)
    with pytest.raises(ValueError):
        rt.tsvd(U, s, V, b, 21)


@pytest.mark.parametrize("shape", [(30, 20), (20, 30)])
def test_tikhonov_normal_equations(shape):
    A, b = _rand(*shape, seed=5, decay=6)
    U, s, V = rt.csvd(A)
    lam = 1e-2 * s[0]
    x = rt.tikhonov(U, s, V, b, lam)
    n = A.shape[1]
    xr = np.linalg.solve(A.conj().T @ A + lam ** 2 * np.eye(n), A.conj().T @ b)
    np.testing.assert_allclose(x, xr, rtol=1e-9)


def test_dsvd_closed_forms():
    a = np.array([3.0, 1.5, 0.2])
    A = np.diag(a).astype(complex)
    b = np.array([1 + 1j, 2.0, -1j])
    U, s, V = rt.csvd(A)
    np.testing.assert_allclose(rt.dsvd(U, s, V, b, 0.7), b / (a + 0.7), rtol=1e-13)
    A2, b2 = _rand(20, 12, seed=6)
    U, s, V = rt.csvd(A2)
    np.testing.assert_allclose(rt.dsvd(U, s, V, b2, 0.0), np.linalg.lstsq(A2, b2, rcond=None)[0], rtol=1e-11)


def _brute_gcv_tik(A, b, lam):
    m, n = A.shape
    M = A.conj().T @ A + lam ** 2 * np.eye(n)
    x = np.linalg.solve(M, A.conj().T @ b)
    infl = A @ np.linalg.solve(M, A.conj().T)
    return np.linalg.norm(A @ x - b) ** 2 / (m - np.trace(infl).real) ** 2


@pytest.mark.parametrize("shape", [(24, 16), (16, 24)])
def test_gcv_tikhonov_grid_brute_force(shape):
    A, b = _rand(*shape, seed=7, decay=8)
    U, s, V = rt.csvd(A)
    beta = U.conj().T @ b
    rows = A.shape[0]
    delta0 = max(np.linalg.norm(b) ** 2 - np.linalg.norm(beta) ** 2, 0) if rows > len(s) else 0.0
    idx, lam, G = rt.gcv_tikhonov_grid(s, beta, rows, delta0, npoints=12, decades=6)
    Gb = np.array([_brute_gcv_tik(A, b, l) for l in lam])
    np.testing.assert_allclose(G, Gb, rtol=1e-6)
    assert idx == int(np.argmin(Gb))
    assert lam[0] == pytest.approx(s[0] * 1e-6) and lam[-1] == pytest.approx(s[0])


def test_gcv_tsvd_brute_force():
    A, b = _rand(30, 18, seed=8, decay=6)
    U, s, V = rt.csvd(A)
    beta = U.conj().T @ b
    delta0 = np.linalg.norm(b) ** 2 - np.linalg.norm(beta) ** 2
    k, G = rt.gcv_tsvd(beta, 30, 18, delta0)
    for kk in range(1, 18):
        x = rt.tsvd(U, s, V, b, kk)
        Gb = np.linalg.norm(A @ x - b) ** 2 / (30 - kk) ** 2
        assert G[kk - 1] == pytest.approx(Gb, rel=1e-9)
    assert k == int(np.argmin(G)) + 1


@pytest.mark.parametrize("fun,a,b", [
    (lambda x: (x - 0.3) ** 2, -1.0, 2.0),
    (lambda x: np.cos(x), 2.0, 5.0),
    (lambda x: x ** 4 - 3 * x ** 2 + x, -2.5, 0.0),
    (lambda x: abs(x - 1e-3) + 1e-9 * x, 1e-6, 1.0),
])
def test_fminbnd_matches_scipy_fminbound(fun, a, b):
    x = rt.fminbnd(fun, a, b, tol_x=1e-4)
    xs = fminbound(fun, a, b, xtol=1e-4)
    assert x == pytest.approx(xs, abs=2e-4)


def test_lcfun_is_negative_curvature():
    A, b = _rand(30, 20, seed=9, decay=5)
    U, s, V = rt.csvd(A)
    beta = U.conj().T @ b
    xi = beta / s
    beta_ls = np.concatenate([beta, [np.linalg.norm(b - U @ beta)]])
    lam0 = 1e-2 * s[0]

    def logs(l):
        x = rt.tikhonov(U, s, V, b, l)
        return np.log(np.linalg.norm(A @ x - b)), np.log(np.linalg.norm(x))

    h = 1e-4 * lam0
    (x0, y0), (xp, yp), (xm, ym) = logs(lam0), logs(lam0 + h), logs(lam0 - h)
    dx, dy = (xp - xm) / (2 * h), (yp - ym) / (2 * h)
    ddx, ddy = (xp - 2 * x0 + xm) / h ** 2, (yp - 2 * y0 + ym) / h ** 2
    kappa = (dx * ddy - ddx * dy) / (dx ** 2 + dy ** 2) ** 1.5
    g = rt.lcfun(lam0, s, beta_ls, xi, "t")[0]
    assert g == pytest.approx(-kappa, rel=1e-3)


def test_quasiopt_is_half_lambda_dx_dlambda():
    A, b = _rand(20, 20, seed=10, decay=4)
    U, s, V = rt.csvd(A)
    reg_min, grid, Q2 = rt.quasiopt(U, s, b, "t")
    for j in (10, 100, 180):
        l = grid[j]
        h = 1e-6 * l
        dx = (rt.tikhonov(U, s, V, b, l + h) - rt.tikhonov(U, s, V, b, l - h)) / (2 * h)
        assert Q2[j] == pytest.approx(0.5 * np.linalg.norm(l * dx), rel=1e-5)
    assert grid[-1] <= reg_min <= grid[0]


def test_gcv_regtools_literal_grid_and_refinement():
    A, b = _rand(24, 16, seed=11, decay=6)
    U, s, V = rt.csvd(A)
    reg_min, grid, G, gi = rt.gcv_regtools(U, s, b, "d")
    assert len(grid) == 200 and grid[0] == pytest.approx(s[0]) and np.all(np.diff(grid) < 0)
    # the literal gcvG grid uses the damped filter lam/(s+lam) for both methods
    beta = U.conj().T @ b
    d0 = np.linalg.norm(b) ** 2 - np.linalg.norm(beta) ** 2
    j = 57
    f = grid[j] / (s + grid[j])
    assert G[j] == pytest.approx((np.linalg.norm(f * beta) ** 2 + d0) / ((24 - 16) + f.sum()) ** 2, rel=1e-12)
    lo, hi = grid[min(gi + 1, 199)], grid[max(gi - 1, 0)]
    assert lo <= reg_min <= hi


def _brute_gcv_damped(A, b, lam):
    """GCV of the damped SVD from its explicit influence matrix A A#,
    A# = V diag(1/(s + lam)) U^H, with an independent SVD (LAPACK gesvd)."""
    m, n = A.shape
    U, s, Vh = sl.svd(A, full_matrices=False, lapack_driver="gesvd")
    Ash = Vh.conj().T @ np.diag(1.0 / (s + lam)) @ U.conj().T
    x = Ash @ b
    infl = A @ Ash
    return np.linalg.norm(A @ x - b) ** 2 / (m - np.trace(infl).real) ** 2


@pytest.mark.parametrize("shape", [(24, 16), (16, 24), (18, 18)])
def test_gcv_regtools_functions_brute_force(shape):
    """gcvG.m's 'd' function is the damped-SVD GCV of the influence matrix,
    and its 't' function is that same GCV at lam^2 (G_t(lam) = G_d(lam^2)),
    so the literal grid values and both refinement functions are pinned."""
    A, b = _rand(*shape, seed=12, decay=5)
    U, s, V = rt.csvd(A)
    fd = rt.gcv_regtools_fun(U, s, b, "d")
    ft = rt.gcv_regtools_fun(U, s, b, "t")
    reg_min, grid, G, gi = rt.gcv_regtools(U, s, b, "t")
    for j in (0, 17, 60, 123, 199):
        assert G[j] == pytest.approx(_brute_gcv_damped(A, b, grid[j]), rel=1e-9)
    for la in (1e-4 * s[0], 3e-2 * s[0], 0.7 * s[0]):
        assert fd(la) == pytest.approx(_brute_gcv_damped(A, b, la), rel=1e-9)
        assert ft(np.sqrt(la)) == pytest.approx(_brute_gcv_damped(A, b, la), rel=1e-9)
    # refinement (gcvG.m:51-52): the bracket is the grid neighbours of the
    # brute-force argmin of the damped GCV, and the minimiser is the bounded
    # Brent minimum of the brute-force G_d(lam^2) over it (scipy fminbound)
    Gb = np.array([_brute_gcv_damped(A, b, la) for la in grid])
    assert gi == int(np.argmin(Gb))
    lo, hi = grid[min(gi + 1, 199)], grid[max(gi - 1, 0)]
    ref = fminbound(lambda la: _brute_gcv_damped(A, b, la * la), lo, hi, xtol=1e-4)
    assert lo <= reg_min <= hi
    assert reg_min == pytest.approx(ref, abs=2e-4)


@pytest.mark.parametrize("shape", [(24, 16), (16, 24)])
def test_gcv_dsvd_grid_brute_force(shape):
    A, b = _rand(*shape, seed=13, decay=8)
    U, s, V = rt.csvd(A)
    beta = U.conj().T @ b
    rows = A.shape[0]
    delta0 = max(np.linalg.norm(b) ** 2 - np.linalg.norm(beta) ** 2, 0) if rows > len(s) else 0.0
    idx, lam, G = rt.gcv_dsvd_grid(s, beta, rows, delta0, npoints=15, decades=6)
    Gb = np.array([_brute_gcv_damped(A, b, l) for l in lam])
    np.testing.assert_allclose(G, Gb, rtol=1e-8)
    assert idx == int(np.argmin(Gb))


def test_esm_pipeline_dsvd_branch():
    """pipeline.esm_solve with the damped filter: the grid index is the
    argmin of the brute-force damped GCV and x the damped solution
    V diag(1/(s+lam)) U^H b from an independent SVD."""
    from inverseacoustics_amd import scenarios as sc
    from oracle import esm as oesm
    from oracle import pipeline

    S = sc.sphere_scenario()
    o = pipeline.esm_solve(S.k[0], S.r_s, S.r_h, S.src, S.b[0], S.area, None, S.rho, S.c,
                           filter="dsvd")
    D, M_sq, M_hq = oesm.esm_assemble(S.k[0], S.r_s, S.r_h, S.src, None, S.rho, S.c)
    H = np.linalg.solve(D.T, M_hq.T).T
    lam = rt.lambda_grid(np.linalg.norm(H, 2), 50, 8.0)
    Gb = np.array([_brute_gcv_damped(H, S.b[0], l) for l in lam])
    assert o["idx"] == int(np.argmin(Gb))
    assert o["lam"] == pytest.approx(lam[o["idx"]], rel=1e-12)
    U, s, Vh = sl.svd(H, full_matrices=False, lapack_driver="gesvd")
    x = Vh.conj().T @ ((U.conj().T @ S.b[0]) / (s + o["lam"]))
    assert np.linalg.norm(o["x"] - x) <= 1e-9 * np.linalg.norm(x)


def test_gcv_gram_form_equals_svd_gcv():
    """The Gram form the certified sweep decides near ties with: for H (l x m,
    l < m), A = H H^H + lam^2 I, the Tikhonov GCV function equals
    |A^-1 b|^2 / tr(A^-1)^2 and the Tikhonov solution H^H A^-1 b, at every
    grid point (against the SVD form of gcv_tikhonov_grid and tikhonov, from
    LAPACK's SVD), within the Gram form's error model 64 eps cond(A),
    cond(A) = (s_1^2 + lam^2) / lam^2 (a dropped lam^2 factor or a missing
    trace term fails by O(1))."""
    rng = np.random.default_rng(7)
    l, m = 40, 70
    U0 = np.linalg.qr(rng.standard_normal((l, l)) + 1j * rng.standard_normal((l, l)))[0]
    V0 = np.linalg.qr(rng.standard_normal((m, l)) + 1j * rng.standard_normal((m, l)))[0]
    H = (U0 * np.logspace(0, -6, l)) @ V0.conj().T
    b = rng.standard_normal(l) + 1j * rng.standard_normal(l)
    U, s, V = rt.csvd(H)
    beta = U.conj().T @ b
    j, lam, G = rt.gcv_tikhonov_grid(s, beta, l)
    Gg = np.empty_like(G)
    checked = 0
    for i, la in enumerate(lam):
        A = H @ H.conj().T + la * la * np.eye(l)
        X = np.linalg.solve(A, np.column_stack([np.eye(l), b]))
        Gg[i] = np.sum(np.abs(X[:, l]) ** 2) / np.trace(X[:, :l]).real ** 2
        tol = 64 * np.finfo(float).eps * (s[0] ** 2 + la * la) / (la * la)
        if tol < 1e-3:
            checked += 1
            assert abs(Gg[i] / G[i] - 1) <= tol, (i, Gg[i], G[i], tol)
            x = H.conj().T @ X[:, l]
            xt = rt.tikhonov(U, s, V, b, la)
            assert np.linalg.norm(x - xt) <= tol * np.linalg.norm(xt), i
    assert checked >= 25
    assert int(np.argmin(Gg)) == j
