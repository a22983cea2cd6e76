"""ADMM sparse / L2 regularised solution, restated from ``MATLAB/ADMMG.m:1-89``.

Test infrastructure only.  For A (m x n) and b, with M = 2 A^H A + rho I and
r0 = 2 A^H b, every iteration solves M c = r0 + rho x - u, shrinks
y = c + u / rho into x, updates the dual u <- u + gamma rho (c - x) and stops
when |c - c_old|, |c - x| and |u - u_old| are all <= epsilon:

* ``nor = 1`` (L1, ADMMG.m:3-30): complex soft threshold
  x = max(0, |y| - lambda/rho) e^{i arg y}, with the continuation
  lambda = 5 lambda_0, 0.75 lambda, ... while lambda > lambda_0 / 10 (the
  loop leaves before running at lambda_0 / 10, ADMMG.m:8,28);
* ``nor = 2`` (L2, ADMMG.m:31-48): x = (rho c + u) / (2 lambda_0 + rho), no
  continuation;
* ``nor = 3`` (Lp, 0 < p < 1, ADMMG.m:49-88): generalised soft thresholding
  (Zuo, Meng, Zhang, Feng, Zhang, "A generalized iterated shrinkage
  algorithm for non-convex sparse coding", ICCV 2013): threshold
  tau = (2 lp (1-p))^(1/(2-p)) + lp p (2 lp (1-p))^((p-1)/(2-p)) with
  lp = 2 lambda / rho, magnitude s from N fixed-point steps
  s <- |y| - lp p s^(p-1), x = sign(y) s where |y| > tau, else 0; the dual
  starts at u = x_0 (ADMMG.m:56).

Start x_0 = A\\b (ADMMG.m:7,35,55) as MATLAB's mldivide computes it:
square A by LU with partial pivoting; rectangular A by Householder QR with
column pivoting (LAPACK zgeqp3), rank r = #{|R_ii| > max(m, n) eps |R_11|},
x[E[:r]] = R[:r, :r]^-1 (Q^H b)[:r] and zero elsewhere -- the basic solution
with at most r nonzeros (for a full-rank overdetermined A the unique
least-squares solution).  Pinned in tests/test_oracle_admm.py against an
independent greedy column selection (largest residual norm against the
span of the columns already chosen, recomputed from scratch at every step)
followed by a least-squares solve on the chosen columns.

Returns (c, iterations, lambda_last) -- ADMMG.m returns c; lambda_last is the
continuation value the last iteration ran at (lambda_0 for L2).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg


def basic_solution(A, b):
    """x = A \\ b as MATLAB's mldivide computes it (ADMMG.m:7,35,55)."""
    A = np.asarray(A, dtype=complex)
    b = np.asarray(b, dtype=complex)
    m, n = A.shape
    if m == n:
        return np.linalg.solve(A, b)
    Q, R, E = scipy.linalg.qr(A, mode="economic", pivoting=True)
    d = np.abs(np.diag(R))
    tol = max(m, n) * np.finfo(float).eps * d[0] if d.size else 0.0
    r = int(np.sum(d > tol))
    x = np.zeros(n, complex)
    if r:
        y = Q[:, :r].conj().T @ b
        x[E[:r]] = scipy.linalg.solve_triangular(R[:r, :r], y)
    return x


def _start(A, b):
    return basic_solution(A, b)


def gst_threshold(lam_p, p):
    """Generalised soft-thresholding threshold of the Lp step (ADMMG.m:61)."""
    t = 2.0 * lam_p * (1.0 - p)
    return t ** (1.0 / (2.0 - p)) + lam_p * p * t ** ((p - 1.0) / (2.0 - p))


def gst_magnitude(ay, lam_p, p, N):
    """N fixed-point steps s <- |y| - lp p s^(p-1) from s = |y| (ADMMG.m:68-71)."""
    s = np.asarray(ay, dtype=complex).copy()
    for _ in range(N):
        with np.errstate(divide="ignore", invalid="ignore"):
            s = ay - lam_p * p * np.power(s, p - 1.0)
    return s


def admm(A, b, lambda0, rho, gamma, epsilon, Nm, nor, p=0.5, N=5):
    A = np.asarray(A, dtype=complex)
    b = np.asarray(b, dtype=complex)
    n = A.shape[1]
    Minv = np.linalg.inv(2.0 * (A.conj().T @ A) + rho * np.eye(n))
    r0 = 2.0 * (A.conj().T @ b)
    x = _start(A, b)
    c = np.zeros(n, complex)
    it = 0

    def converged(c, cold, x, u, uold):
        return (np.linalg.norm(c - cold) <= epsilon and np.linalg.norm(c - x) <= epsilon
                and np.linalg.norm(u - uold) <= epsilon)

    if nor == 2:
        u = np.zeros(n, complex)
        dlr = 2.0 * lambda0 + rho
        for _ in range(Nm):
            cold = c
            c = Minv @ (r0 + rho * x - u)
            x = (rho * c + u) / dlr
            uold = u
            u = u + gamma * rho * (c - x)
            it += 1
            if converged(c, cold, x, u, uold):
                break
        return c, it, lambda0

    lam_e = lambda0 / 10.0
    lam = lambda0 * 5.0
    u = x.copy() if nor == 3 else np.zeros(n, complex)
    lam_run = lam
    while lam > lam_e:
        lam_run = lam
        done = False
        if nor == 1:
            lr = lam / rho
        else:
            lam_p = 2.0 * lam / rho
            tau = gst_threshold(lam_p, p)
        for _ in range(Nm):
            cold = c
            c = Minv @ (r0 + rho * x - u)
            y = c + u / rho
            ay = np.abs(y)
            if nor == 1:
                x = np.maximum(0.0, ay - lr) * np.exp(1j * np.angle(y))
            else:
                s = gst_magnitude(ay, lam_p, p, N)
                with np.errstate(invalid="ignore"):
                    sgn = np.where(ay > 0, y / np.where(ay > 0, ay, 1.0), 0.0)
                x = np.where(ay > tau, sgn * s, 0.0)
            uold = u
            u = u + gamma * rho * (c - x)
            it += 1
            if converged(c, cold, x, u, uold):
                done = True
                break
        if done:
            break
        lam = max(0.75 * lam, lam_e)
    return c, it, lam_run
