"""GPU ADMM (ia_admm) against the oracle restatement of MATLAB/ADMMG.m, for
the L1, L2 and Lp branches, batched systems, over- and under-determined."""
import numpy as np
import pytest

from oracle import admm as oadmm

pytestmark = pytest.mark.gpu


def _batch(m, n, B, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((B, m, n)) + 1j * rng.standard_normal((B, m, n))
    X = np.zeros((B, n), complex)
    for i in range(B):
        idx = rng.choice(n, size=max(1, n // 6), replace=False)
        X[i, idx] = rng.standard_normal(len(idx)) + 1j * rng.standard_normal(len(idx))
    b = np.einsum("bij,bj->bi", A, X) + 0.02 * (rng.standard_normal((B, m)) + 1j * rng.standard_normal((B, m)))
    return A, b


@pytest.mark.parametrize("shape", [(40, 24), (24, 40)])
@pytest.mark.parametrize("norm,kw", [(1, {}), (2, {}), (3, {"p": 0.5, "n_gst": 5}), (3, {"p": 0.8, "n_gst": 3})])
def test_admm_matches_oracle(ia, shape, norm, kw):
    A, b = _batch(*shape, 3, 7)
    prm = dict(lambda0=1.5, rho=2.0, gamma=1.0, epsilon=1e-7, max_iter=400, norm=norm)
    c, it = ia.admm(A, b, **prm, **kw)
    for i in range(3):
        c0, it0, _ = oadmm.admm(A[i], b[i], prm["lambda0"], prm["rho"], prm["gamma"], prm["epsilon"],
                                prm["max_iter"], norm, kw.get("p", 0.5), kw.get("n_gst", 5))
        assert it[i] == it0, (i, it[i], it0)
        assert np.linalg.norm(c[i] - c0) <= 1e-9 * max(1.0, np.linalg.norm(c0))


def test_admm_continuation_without_convergence(ia):
    A, b = _batch(30, 20, 2, 3)
    c, it = ia.admm(A, b, lambda0=1.0, rho=1.0, gamma=1.0, epsilon=0.0, max_iter=3, norm=1)
    c0, it0, _ = oadmm.admm(A[0], b[0], 1.0, 1.0, 1.0, 0.0, 3, 1)
    assert it[0] == it0 == 3 * 14
    assert np.linalg.norm(c[0] - c0) <= 1e-10 * np.linalg.norm(c0)


def test_admm_rejects_bad_parameters(ia):
    A, b = _batch(10, 8, 1, 1)
    with pytest.raises(ia.IAError):
        ia.admm(A, b, lambda0=1.0, norm=4)
    with pytest.raises(ia.IAError):
        ia.admm(A, b, lambda0=1.0, norm=3, p=1.5)
    with pytest.raises(ia.IAError):
        ia.admm(A, b, lambda0=1.0, rho=0.0)


@pytest.mark.parametrize("shape", [(24, 40), (40, 24), (10, 37)])
def test_admm_start_is_matlab_basic_solution(ia, shape):
    """One L2 iteration from x_0 gives c_1 = M^-1 (r0 + rho x_0), so the start
    is recovered as x_0 = (M c_1 - r0) / rho: against the oracle's A\\b (the
    basic solution of QR with column pivoting for a rectangular A, at most
    min(m, n) nonzeros in the columns LAPACK's pivoting takes)."""
    A, b = _batch(*shape, 2, 9)
    rho = 2.0
    c, it = ia.admm(A, b, lambda0=1.0, rho=rho, gamma=1.0, epsilon=0.0, max_iter=1, norm=2)
    assert np.all(it == 1)
    for i in range(2):
        M = 2.0 * A[i].conj().T @ A[i] + rho * np.eye(shape[1])
        x0 = (M @ c[i] - 2.0 * A[i].conj().T @ b[i]) / rho
        ref = oadmm.basic_solution(A[i], b[i])
        assert np.linalg.norm(x0 - ref) <= 1e-9 * np.linalg.norm(ref)
        assert np.count_nonzero(ref) == min(shape)
        assert np.abs(x0[ref == 0]).max(initial=0.0) <= 1e-9 * np.linalg.norm(ref)
