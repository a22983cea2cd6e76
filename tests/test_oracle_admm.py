"""Pins of the ADMM oracle (oracle/admm.py, MATLAB/ADMMG.m): the L2 branch
against its closed form, the L1 branch against an independent LASSO solver
(proximal gradient) at the lambda the continuation ended on, and the
generalised soft thresholding of the Lp branch against brute-force scalar
minimisation, and the start A\\b (MATLAB's basic solution for rectangular A)
against an independent greedy column selection.  CPU only."""
import numpy as np
import pytest

from oracle import admm


def _problem(m, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    x = np.zeros(n, complex)
    idx = rng.choice(n, size=max(1, n // 8), replace=False)
    x[idx] = rng.standard_normal(len(idx)) + 1j * rng.standard_normal(len(idx))
    b = A @ x + 0.01 * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    return A, b


@pytest.mark.parametrize("shape", [(40, 24), (24, 40)])
def test_l2_is_tikhonov(shape):
    A, b = _problem(*shape, 1)
    lam0 = 0.3
    c, it, _ = admm.admm(A, b, lam0, rho=1.0, gamma=1.0, epsilon=1e-12, Nm=20000, nor=2)
    n = A.shape[1]
    ref = np.linalg.solve(A.conj().T @ A + lam0 * np.eye(n), A.conj().T @ b)
    assert it < 20000
    assert np.linalg.norm(c - ref) <= 1e-9 * np.linalg.norm(ref)


def _lasso_pgd(A, b, lam, iters=200000):
    """min |A c - b|^2 + lam |c|_1 by proximal gradient with complex soft
    thresholding (step 1/L, L = 2 |A|_2^2)."""
    L = 2 * np.linalg.norm(A, 2) ** 2
    c = np.zeros(A.shape[1], complex)
    for _ in range(iters):
        g = c - 2 * (A.conj().T @ (A @ c - b)) / L
        c_new = np.maximum(0, np.abs(g) - lam / L) * np.exp(1j * np.angle(g))
        if np.linalg.norm(c_new - c) < 1e-15 * max(1.0, np.linalg.norm(c)):
            c = c_new
            break
        c = c_new
    return c


def test_l1_converges_to_the_lasso_minimiser():
    A, b = _problem(30, 20, 2)
    lam0 = 2.0
    c, it, lam_run = admm.admm(A, b, lam0, rho=2.0, gamma=1.0, epsilon=1e-11, Nm=50000, nor=1)
    # the ADMM x-update is prox of (lam/rho)|.|_1 against (rho/2)|.|^2: the
    # fixed point minimises |A c - b|^2 + lam |c|_1
    ref = _lasso_pgd(A, b, lam_run)
    assert np.linalg.norm(c - ref) <= 1e-6 * np.linalg.norm(ref)
    assert 0 < np.count_nonzero(np.abs(ref) > 1e-9) < A.shape[1]  # a genuinely sparse case


def test_l1_continuation_runs_down_to_lambda0_over_ten():
    A, b = _problem(30, 20, 3)
    # one iteration per stage never converges: the continuation visits every
    # stage 5 lambda_0 * 0.75^k > lambda_0 / 10
    c, it, lam_run = admm.admm(A, b, 1.0, rho=1.0, gamma=1.0, epsilon=0.0, Nm=1, nor=1)
    k = int(np.ceil(np.log(50) / np.log(4 / 3)))
    assert it == k and abs(lam_run - 5 * 0.75 ** (k - 1)) < 1e-12


@pytest.mark.parametrize("p", [0.3, 0.5, 0.8])
@pytest.mark.parametrize("lam_p", [0.05, 0.5, 2.0])
def test_gst_threshold_and_magnitude_minimise_the_scalar_problem(p, lam_p):
    tau = admm.gst_threshold(lam_p, p)
    s_grid = np.linspace(0, 10 * max(tau, 1.0), 400001)
    for a in (0.5 * tau, 0.99 * tau, 1.01 * tau, 1.5 * tau, 4 * tau):
        obj = 0.5 * (s_grid - a) ** 2 + lam_p * s_grid ** p
        s_best = s_grid[np.argmin(obj)]
        if a <= tau:
            assert s_best < 1e-3 * max(tau, 1e-12) or np.min(obj) >= 0.5 * a * a - 1e-12
        else:
            s = admm.gst_magnitude(np.array([a]), lam_p, p, 60)[0].real
            assert abs(s - s_best) <= 2 * (s_grid[1] - s_grid[0]) + 1e-9 * a


def _greedy_basic(A, b, r):
    """Column selection of QR with column pivoting recomputed from scratch:
    step j takes the column whose residual against the span of the columns
    already taken has the largest norm (the first on ties), then x solves
    least squares on the r columns taken and is zero elsewhere."""
    n = A.shape[1]
    taken = []
    for _ in range(r):
        if taken:
            Qt, _ = np.linalg.qr(A[:, taken])
            res = A - Qt @ (Qt.conj().T @ A)
        else:
            res = A
        nrm = np.linalg.norm(res, axis=0)
        nrm[taken] = -1.0
        taken.append(int(np.argmax(nrm)))
    x = np.zeros(n, complex)
    x[taken] = np.linalg.lstsq(A[:, taken], b, rcond=None)[0]
    return x, taken


@pytest.mark.parametrize("shape", [(24, 40), (10, 37), (40, 24), (7, 7)])
def test_start_is_matlab_basic_solution(shape):
    m, n = shape
    A, b = _problem(m, n, 21)
    x = admm.basic_solution(A, b)
    if m == n:
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=0,
                                   atol=1e-12 * np.linalg.norm(x))
        return
    x0, taken = _greedy_basic(A, b, min(m, n))
    assert set(np.flatnonzero(x).tolist()) == set(taken)
    np.testing.assert_allclose(x, x0, rtol=0, atol=1e-10 * np.linalg.norm(x0))
    if m < n:  # consistent: the basic solution fits b exactly with m nonzeros
        assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
    else:      # full column rank: the unique least-squares solution
        np.testing.assert_allclose(x, np.linalg.lstsq(A, b, rcond=None)[0], rtol=0,
                                   atol=1e-10 * np.linalg.norm(x))


def test_start_skips_zero_columns():
    """Zero columns have zero norm, are never pivoted ahead of the others and
    get zero entries; the rank is that of the nonzero columns."""
    A, b = _problem(12, 30, 5)
    A[:, [0, 7, 29]] = 0
    x = admm.basic_solution(A, b)
    assert np.all(x[[0, 7, 29]] == 0)
    assert np.count_nonzero(x) == 12
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
