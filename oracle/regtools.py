"""Regularisation restatement: compact SVD, TSVD, Tikhonov, damped SVD, GCV,
L-curve, quasi-optimality and the bounded scalar minimiser they use.

Reference: ``CUDA/tsvd.cuh:1-172`` (csvd, tsvd, betak) and
``MATLAB/regtools/*G.m`` (the only implementation of Tikhonov, GCV, L-curve
and quasi-optimality in the reference; adapted from Hansen's Regularization
Tools 4.1).

Pinned by (tests/test_oracle_reg.py):
* ``csvd`` is ``numpy.linalg.svd`` (LAPACK gesdd), checked against
  ``scipy.linalg.svd(lapack_driver='gesvd')`` through phase-invariant products;
* ``tsvd``/``tikhonov``/``dsvd`` against direct normal-equation / pseudo-inverse
  solves (numpy lstsq, (A^H A + lam^2 I)^-1 A^H b);
* ``gcv_*`` against brute-force evaluation of ||A x - b||^2/(trace(I - A A#))^2
  built from explicit influence matrices on tiny problems;
* ``fminbnd`` against ``scipy.optimize.fminbound`` (same published
  Forsythe-Malcolm-Moler / Brent algorithm) with xtol = 1e-4;
* ``lcfun`` against finite-difference curvature of (log rho, log eta).
MATLAB's closed-source ``fminbnd``/``svd`` last-digit behaviour is not pinned.
"""
from __future__ import annotations

import math

import numpy as np

EPS = np.finfo(np.float64).eps


# ------------------------------------------------------------------ SVD
def csvd(A, driver="gesdd"):
    """Compact SVD A = U diag(s) V^H (csvdG.m:17-24, tsvd.cuh:10-120).

    m >= n: U m x n, s n, V n x n.  m < n: U m x m, s m, V n x m (the reference
    works on A^H for this case, tsvd.cuh:93-103).  ``driver`` selects the
    LAPACK routine (gesdd, numpy's default, or gesvd, the routine cuSOLVER and
    the reference's CPU invr use); comparing the two measures how far two
    backward-stable FP64 SVDs of the same matrix can disagree."""
    if driver == "gesdd":
        U, s, Vh = np.linalg.svd(np.asarray(A), full_matrices=False)
    else:
        import scipy.linalg as sl

        U, s, Vh = sl.svd(np.asarray(A), full_matrices=False, lapack_driver=driver)
    return U, s, Vh.conj().T


# -------------------------------------------------------------- filters
def tsvd(U, s, V, b, k):
    """x_k = V[:, :k] (beta[:k] / s[:k]), beta = U^H b (tsvdG.m:16-23,
    tsvd.cuh:151-154 with betak tsvd.cuh:1-8).  0 <= k <= p."""
    p = s.shape[0]
    if k < 0 or k > p:
        raise ValueError("k out of range")
    beta = U[:, :p].conj().T @ b
    xi = beta / s
    return V[:, :k] @ xi[:k]


def tikhonov(U, s, V, b, lam):
    """Standard-form Tikhonov x = V (s beta/(s^2 + lam^2)) (tikhonovG.m:26-30)."""
    if lam < 0:
        raise ValueError("lambda must be >= 0")
    p = s.shape[0]
    beta = U[:, :p].conj().T @ b
    return V[:, :p] @ (s * beta / (s * s + lam * lam))


def dsvd(U, s, V, b, lam):
    """Damped SVD x = V (beta/(s + lam)) (dsvdG.m:20-26)."""
    if lam < 0:
        raise ValueError("lambda must be >= 0")
    p = s.shape[0]
    beta = U[:, :p].conj().T @ b
    return V[:, :p] @ (beta / (s + lam))


# ---------------------------------------------------- bounded minimiser
def fminbnd(fun, ax, bx, tol_x=1e-4, maxfun=500, maxiter=500):
    """Bounded scalar minimisation by golden section + parabolic interpolation
    (Forsythe, Malcolm & Moler, "Computer Methods for Mathematical
    Computations", 1977, routine FMIN; the algorithm MATLAB's ``fminbnd``
    documents, called with its default TolX = 1e-4 at gcvG.m:51,
    l_curveG.m:50, quasioptG.m:40).  Returns the minimiser."""
    if ax > bx:
        raise ValueError("lower bound exceeds upper bound")
    seps = math.sqrt(EPS)
    c = 0.5 * (3.0 - math.sqrt(5.0))
    a, b = float(ax), float(bx)
    v = a + c * (b - a)
    w = xf = v
    d = e = 0.0
    fx = fun(xf)
    fv = fw = fx
    nfev = 1
    it = 0
    xm = 0.5 * (a + b)
    tol1 = seps * abs(xf) + tol_x / 3.0
    tol2 = 2.0 * tol1
    while abs(xf - xm) > (tol2 - 0.5 * (b - a)):
        golden = True
        if abs(e) > tol1:
            golden = False
            r = (xf - w) * (fx - fv)
            q = (xf - v) * (fx - fw)
            p = (xf - v) * q - (xf - w) * r
            q = 2.0 * (q - r)
            if q > 0.0:
                p = -p
            q = abs(q)
            r = e
            e = d
            if abs(p) < abs(0.5 * q * r) and p > q * (a - xf) and p < q * (b - xf):
                d = p / q
                u = xf + d
                if (u - a) < tol2 or (b - u) < tol2:
                    si = 1.0 if (xm - xf) >= 0 else -1.0
                    d = tol1 * si
            else:
                golden = True
        if golden:
            e = (a - xf) if xf >= xm else (b - xf)
            d = c * e
        si = 1.0 if d >= 0 else -1.0
        u = xf + si * max(abs(d), tol1)
        fu = fun(u)
        nfev += 1
        it += 1
        if fu <= fx:
            if u >= xf:
                a = xf
            else:
                b = xf
            v, fv = w, fw
            w, fw = xf, fx
            xf, fx = u, fu
        else:
            if u < xf:
                a = u
            else:
                b = u
            if fu <= fw or w == xf:
                v, fv = w, fw
                w, fw = u, fu
            elif fu <= fv or v == xf or v == w:
                v, fv = u, fu
        xm = 0.5 * (a + b)
        tol1 = seps * abs(xf) + tol_x / 3.0
        tol2 = 2.0 * tol1
        if nfev >= maxfun or it >= maxiter:
            break
    return xf


# ------------------------------------------------------------------ GCV
def reg_grid_regtools(s, npoints=200):
    """gcvG.m:28-30 / l_curveG.m:36-38: lam_j = rpm * ratio^(npoints-j),
    rpm = max(s_p, 16 eps s_1), descending from s_1 to rpm."""
    p = s.shape[0]
    rpm = max(s[p - 1], s[0] * 16 * EPS)
    ratio = (s[0] / rpm) ** (1.0 / (npoints - 1))
    return (ratio ** np.arange(npoints - 1, -1, -1)) * rpm


def gcv_regtools_fun(U, s, b, method="t"):
    """The scalar function gcvG.m minimises (gcvG.m:32-48):
    G(lam) = (||f .* beta||^2 + delta0) / ((m - n) + sum f)^2 with
    'd': f = lam/(s + lam) (also the grid filter of both methods, gcvG.m:38)
    and 't': f = lam^2/(s + lam^2), i.e. G_t(lam) = G_d(lam^2); (m, n) is the
    shape of U and delta0 = ||b||^2 - ||beta||^2 when m > n and positive."""
    m, n = U.shape
    p = s.shape[0]
    beta = U.conj().T @ b
    beta2 = np.linalg.norm(b) ** 2 - np.linalg.norm(beta) ** 2
    delta0 = beta2 if (m > n and beta2 > 0) else 0.0
    bp = beta[:p]
    if method == "d":
        def fun(la):
            ff = la / (s + la)
            return (np.linalg.norm(ff * bp) ** 2 + delta0) / ((m - n) + np.sum(ff)) ** 2
    elif method == "t":
        def fun(la):
            ff = la * la / (s + la * la)
            return (np.linalg.norm(ff * bp) ** 2 + delta0) / ((m - n) + np.sum(ff)) ** 2
    else:
        raise ValueError("method must be 't' or 'd'")
    return fun


def gcv_regtools(U, s, b, method="t"):
    """Literal restatement of gcvG.m:19-52 ('regtools-GPU' mode).

    Grid: f = lam/(s + lam) for BOTH methods (gcvG.m:38).  Refinement by
    fminbnd of 'd': f = lam/(s+lam), 't': f = lam^2/(s + lam^2) (gcvG.m:43-48)
    between the grid neighbours of the argmin (gcvG.m:51-52).  Note this is not
    Hansen's Tikhonov GCV; see ``gcv_tikhonov_grid`` for that.
    Returns (reg_min, reg_param, G, minGi)."""
    p = s.shape[0]
    reg_param = reg_grid_regtools(s)
    fun_d = gcv_regtools_fun(U, s, b, "d")
    G = np.array([fun_d(la) for la in reg_param])
    minGi = int(np.argmin(G))
    fun = gcv_regtools_fun(U, s, b, method)
    npts = len(reg_param)
    lo = reg_param[min(minGi + 1, npts - 1)]
    hi = reg_param[max(minGi - 1, 0)]
    return fminbnd(fun, lo, hi), reg_param, G, minGi


def lambda_grid(s_max, npoints=50, decades=8.0):
    """lam_j = s_max 10^(-decades + decades j/(npoints-1)), j = 0..npoints-1
    (ascending), the configured 50-point grid from 1e-8 to 1 times sigma_max."""
    j = np.arange(npoints, dtype=np.float64)
    return s_max * 10.0 ** (-decades + decades * j / (npoints - 1))


def gcv_tikhonov_grid(s, beta, rows, delta0=0.0, npoints=50, decades=8.0):
    """Hansen's standard-form Tikhonov GCV on the configured grid:
    G(lam) = (sum (lam^2/(s^2+lam^2))^2 |beta|^2 + delta0)
             / ((rows - p) + sum lam^2/(s^2+lam^2))^2
    (regtools gcv.m, Tikhonov branch; the reference's gcvG.m:38 uses the
    damped filter instead).  Returns (index, lam_grid, G); ties -> lowest index."""
    p = s.shape[0]
    lam = lambda_grid(s[0], npoints, decades)
    s2 = (s * s)[:, None]
    f = lam[None, :] ** 2 / (s2 + lam[None, :] ** 2)
    num = np.sum(f * f * (np.abs(beta[:p]) ** 2)[:, None], axis=0) + delta0
    den = ((rows - p) + np.sum(f, axis=0)) ** 2
    G = num / den
    return int(np.argmin(G)), lam, G


def gcv_dsvd_grid(s, beta, rows, delta0=0.0, npoints=50, decades=8.0):
    """GCV of the damped SVD (dsvdG.m:26, filter f = s/(s+lam)) on the
    configured grid: G(lam) = (sum (lam/(s+lam))^2 |beta|^2 + delta0)
    / ((rows - p) + sum lam/(s+lam))^2.  Returns (index, lam_grid, G); ties
    -> lowest index."""
    p = s.shape[0]
    lam = lambda_grid(s[0], npoints, decades)
    f = lam[None, :] / (s[:, None] + lam[None, :])
    num = np.sum(f * f * (np.abs(beta[:p]) ** 2)[:, None], axis=0) + delta0
    den = ((rows - p) + np.sum(f, axis=0)) ** 2
    G = num / den
    return int(np.argmin(G)), lam, G


def gcv_tsvd(beta, rows, p, delta0=0.0):
    """Discrete GCV for TSVD, k = 1..p-1:
    G(k) = (sum_{i>k} |beta_i|^2 + delta0) / (rows - k)^2
    (Hansen, Regularization Tools 4.1, gcv.m 'tsvd' branch; not present in the
    reference, which only exposes ``tsvd`` with a caller-chosen k,
    tsvd.cuh:122).  Returns (k_selected, G) with G[k-1] = G(k)."""
    a2 = np.abs(beta[:p]) ** 2
    tail = np.cumsum(a2[::-1])[::-1]          # tail[i] = sum_{j>=i} |beta_j|^2
    k = np.arange(1, p)
    rho2 = tail[k] + delta0                   # sum_{i>k} (1-based) = tail[k] (0-based)
    G = rho2 / (rows - k) ** 2
    return int(np.argmin(G)) + 1, G


# -------------------------------------------------------------- L-curve
def lcfun(la, s, beta, xi, method="t"):
    """Negative curvature of the L-curve (lcfunG.m:1-49); ``beta`` may carry
    the least-squares residual as an extra last entry (lcfunG.m:6-12)."""
    la = np.atleast_1d(np.asarray(la, dtype=np.float64))
    if beta.shape[0] > s.shape[0]:
        LS = True
        rhoLS2 = np.abs(beta[-1]) ** 2
        beta = beta[:-1]
    else:
        LS = False
    rp = la[None, :]
    if method == "d":
        f = s[:, None] / (s[:, None] + rp)
    elif method == "t":
        f = (s * s)[:, None] / ((s * s)[:, None] + rp * rp)
    else:
        raise ValueError("method must be 't' or 'd'")
    cf = 1 - f
    axi2 = (np.abs(xi) ** 2)[:, None]
    abe2 = (np.abs(beta) ** 2)[:, None]
    eta = np.sqrt(np.sum(f * f * axi2, axis=0))
    rho = np.sqrt(np.sum(cf * cf * abe2, axis=0))
    f1 = -2 * f * cf / rp
    f2 = -f1 * (3 - 4 * f) / rp
    phi = np.sum(f * f1 * axi2, axis=0)
    psi = np.sum(cf * f1 * abe2, axis=0)
    dphi = np.sum((f1 * f1 + f * f2) * axi2, axis=0)
    dpsi = np.sum((-f1 * f1 + cf * f2) * abe2, axis=0)
    if LS:
        rho = np.sqrt(rho * rho + rhoLS2)
    deta = phi / eta
    drho = -psi / rho
    ddeta = dphi / eta - deta * (deta / eta)
    ddrho = -dpsi / rho - drho * (drho / rho)
    dlogeta = deta / eta
    dlogrho = drho / rho
    ddlogeta = ddeta / eta - dlogeta ** 2
    ddlogrho = ddrho / rho - dlogrho ** 2
    g = -(dlogrho * ddlogeta - ddlogrho * dlogeta) / (dlogrho ** 2 + dlogeta ** 2) ** 1.5
    return np.real(g)


def l_curve(U, s, b, method="t"):
    """L-curve corner (l_curveG.m:21-53): 200-point grid of lcfun, argmin,
    then fminbnd between the grid neighbours.  Returns (reg_c, reg_param, g, gi)."""
    m, n = U.shape
    p = s.shape[0]
    beta_full = U.conj().T @ b
    b0 = b - U @ beta_full
    beta = beta_full[:p]
    reg_param = reg_grid_regtools(s)
    with np.errstate(divide="ignore", invalid="ignore"):
        xi = beta / s
    xi[np.isinf(xi)] = 0
    if m > n:
        beta = np.concatenate([beta, [np.linalg.norm(b0)]])
    g = lcfun(reg_param, s, beta, xi, method)
    gi = int(np.argmin(g))
    npts = len(g)
    lo = reg_param[min(gi + 1, npts - 1)]
    hi = reg_param[max(gi - 1, 0)]
    reg_c = fminbnd(lambda la: float(lcfun(la, s, beta, xi, method)[0]), lo, hi)
    return reg_c, reg_param, g, gi


def l_curve_grid(s, beta, rows, method="t", npoints=50, decades=8.0, b_norm2=None):
    """L-curve corner on the configured grid (argmin of lcfun, ties -> lowest
    index).  ``b_norm2`` adds the least-squares residual when rows > p."""
    p = s.shape[0]
    lam = lambda_grid(s[0], npoints, decades)
    bb = beta[:p]
    with np.errstate(divide="ignore", invalid="ignore"):
        xi = bb / s
    xi[np.isinf(xi)] = 0
    if rows > p and b_norm2 is not None:
        r2 = max(b_norm2 - float(np.sum(np.abs(bb) ** 2)), 0.0)
        bb = np.concatenate([bb, [math.sqrt(r2)]])
    g = lcfun(lam, s, bb, xi, method)
    return int(np.nanargmin(g)), lam, g


def quasiopt(U, s, b, method="t"):
    """Quasi-optimality criterion (quasioptG.m:16-41).  Returns (reg_min, reg_param, Q2)."""
    p = s.shape[0]
    xi = (U.conj().T @ b)[:p] / s
    rpm = s[p - 1]
    ratio = (s[0] / s[p - 1]) ** (1.0 / 199)
    reg_param = (ratio ** np.arange(199, -1, -1)) * rpm
    rp = reg_param[None, :]
    if method == "d":
        f = s[:, None] / (s[:, None] + rp)
        def fun(la):
            ff = s / (s + la)
            return np.linalg.norm((1 - ff) * ff * xi)
    elif method == "t":
        f = (s * s)[:, None] / ((s * s)[:, None] + rp * rp)
        def fun(la):
            ff = s * s / (s * s + la * la)
            return np.linalg.norm((1 - ff) * ff * xi)
    else:
        raise ValueError("method must be 't' or 'd'")
    Q2 = np.sqrt(np.sum(np.abs((1 - f) * f * xi[:, None]) ** 2, axis=0))
    mi = int(np.argmin(Q2))
    lo = reg_param[min(mi + 1, 199)]
    hi = reg_param[max(mi - 1, 0)]
    return fminbnd(fun, lo, hi), reg_param, Q2
