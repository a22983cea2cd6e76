"""One-frequency ESM inverse solve, composed from the oracle restatements.

Chain of the reference's call stack (SURVEY call stack A; the reference has
no driver): esm (CUDA/esm.cuh:112) -> csvd (tsvd.cuh:10) -> parameter choice
(regtools gcvG.m / l_curveG.m, or the configured 50-point grid) -> filtered
solution (tsvd.cuh:122, tikhonovG.m) -> to_pressure (utils.cuh:199) with
M_e = M_sq D^-1 -> sound_power (utils.cuh:265).
"""
from __future__ import annotations

import numpy as np

from . import esm as oesm
from . import post
from . import regtools as rt


def esm_solve(k, r_s, r_h, src, b, area, normals=None, rho=1.21, c=343.0, filter="tikhonov",
              select="gcv", n_lambda=50, decades=8.0):
    D, M_sq, M_hq = oesm.esm_assemble(k, r_s, r_h, src, normals, rho, c)
    H_e = np.linalg.solve(D.T, M_hq.T).T
    U, s, V = rt.csvd(H_e)
    beta = U.conj().T @ b
    rows, p = H_e.shape[0], len(s)
    delta0 = max(np.linalg.norm(b) ** 2 - np.linalg.norm(beta) ** 2, 0.0) if rows > p else 0.0
    k_tsvd, _ = rt.gcv_tsvd(beta, rows, p, delta0)
    lam, idx = 0.0, -1
    if filter == "tsvd":
        x = rt.tsvd(U, s, V, b, k_tsvd)
    else:
        if select == "gcv":
            if filter == "tikhonov":
                idx, grid, _ = rt.gcv_tikhonov_grid(s, beta, rows, delta0, n_lambda, decades)
            else:
                idx, grid, _ = rt.gcv_dsvd_grid(s, beta, rows, delta0, n_lambda, decades)
            lam = grid[idx]
        elif select == "lcurve":
            idx, grid, _ = rt.l_curve_grid(s, beta, rows, "t" if filter == "tikhonov" else "d",
                                           n_lambda, decades, np.linalg.norm(b) ** 2)
            lam = grid[idx]
        else:
            raise ValueError(select)
        x = rt.tikhonov(U, s, V, b, lam) if filter == "tikhonov" else rt.dsvd(U, s, V, b, lam)
    q = np.linalg.solve(D, x)
    p_s = M_sq @ q
    W = post.sound_power(p_s, x, area)
    return dict(lam=lam, idx=idx, k_tsvd=k_tsvd, x=x, p_s=p_s, W=W, s=s, beta=beta, H_e=H_e)


def hels_solve(k, N, azelr_s, azelr_h, b, area, rho=1.21, c=343.0, filter="tikhonov",
               select="gcv", n_lambda=50, decades=8.0, driver="gesdd"):
    """HELS chain: hels (hels.cuh:222) -> csvd -> parameter choice -> filtered
    coefficients -> to_velocity / to_pressure (utils.cuh:199,232) -> sound_power."""
    from . import hels as oh

    M_e, H, Q_e = oh.hels(k, azelr_s, azelr_h, N, rho, c)
    r = _regularised(H, b, filter, select, n_lambda, decades, driver)
    v = Q_e @ r["x"]
    p = M_e @ r["x"]
    r.update(v=v, p=p, W=post.sound_power(p, v, area), H=H)
    return r


def _regularised(A, b, filter="tikhonov", select="gcv", n_lambda=50, decades=8.0, driver="gesdd"):
    U, s, V = rt.csvd(A, driver)
    beta = U.conj().T @ b
    rows, p = A.shape[0], len(s)
    delta0 = max(np.linalg.norm(b) ** 2 - np.linalg.norm(beta) ** 2, 0.0) if rows > p else 0.0
    k_tsvd, _ = rt.gcv_tsvd(beta, rows, p, delta0)
    lam, idx = 0.0, -1
    if filter == "tsvd":
        x = rt.tsvd(U, s, V, b, k_tsvd)
    else:
        if select == "gcv" and filter == "tikhonov":
            idx, grid, _ = rt.gcv_tikhonov_grid(s, beta, rows, delta0, n_lambda, decades)
        elif select == "lcurve":
            idx, grid, _ = rt.l_curve_grid(s, beta, rows, "t" if filter == "tikhonov" else "d",
                                           n_lambda, decades, np.linalg.norm(b) ** 2)
        else:
            raise ValueError((filter, select))
        lam = grid[idx]
        x = rt.tikhonov(U, s, V, b, lam) if filter == "tikhonov" else rt.dsvd(U, s, V, b, lam)
    return dict(lam=lam, idx=idx, k_tsvd=k_tsvd, x=x, s=s, beta=beta)


def ibem_solve(k, X, E, field, b, area, rule, pages, chief=None, rho=1.21, c=343.0,
               filter="tikhonov", select="gcv", n_lambda=50, decades=8.0, driver="gesdd"):
    """IBEM chain: ibem (ibem.cuh:424, Z and H_b) -> csvd -> parameter choice ->
    filtered surface velocity v -> p_s = Z v (to_pressure, utils.cuh:199) ->
    sound_power (utils.cuh:265)."""
    from . import ibem as oib

    Z, H = oib.ibem(k, X, E, field, rule, pages, chief, rho, c)
    r = _regularised(H, b, filter, select, n_lambda, decades, driver)
    v = r["x"]
    p = Z @ v
    r.update(v=v, p=p, W=post.sound_power(p, v, area), H=H, Z=Z)
    return r
