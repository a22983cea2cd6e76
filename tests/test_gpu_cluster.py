"""Clustered spectra: matrices with exact 2- to 7-fold multiple singular
values between 1e-1 and 1e-9 s_1 (the multiplets symmetric geometries
produce).  The GPU SVD (ia_csvd) and the batched regularised solve
(ia_regsolve, the FP64 form) against LAPACK.

Singular vectors of a multiplet are defined only up to a unitary mixing, so
the checks are basis-free: orthonormality of U and V, the reconstruction
A = U diag(s) V^H, the projector onto each multiplet, and the filtered
solution x, which weights all vectors of a multiplet alike."""
import numpy as np
import pytest

from oracle import pipeline
from oracle import regtools as rt

pytestmark = pytest.mark.gpu

LEVELS = [(1e-1, 2), (1e-2, 3), (1e-3, 4), (1e-5, 5), (1e-7, 6), (1e-9, 7)]


def _clustered(rows, cols, seed):
    rng = np.random.default_rng(seed)
    p = min(rows, cols)
    U, _ = np.linalg.qr(rng.standard_normal((rows, p)) + 1j * rng.standard_normal((rows, p)))
    V, _ = np.linalg.qr(rng.standard_normal((cols, p)) + 1j * rng.standard_normal((cols, p)))
    s = 10.0 ** (-10.0 * np.arange(p) / (p - 1))
    # replace background values next to each level by an exact multiplet
    for lev, mult in LEVELS:
        i = int(np.argmin(np.abs(np.log10(s) - np.log10(lev))))
        s[i:i + mult] = lev
    s = np.sort(s)[::-1]
    A = (U * s) @ V.conj().T
    x0 = rng.standard_normal(cols) + 1j * rng.standard_normal(cols)
    b = A @ x0
    b = b + 1e-7 * np.linalg.norm(b) / np.sqrt(rows) * (rng.standard_normal(rows)
                                                      + 1j * rng.standard_normal(rows))
    return A, b, s


def _multiplets(s):
    """Index ranges [i0, i1) of the multiplets of the constructed spectrum."""
    out = []
    for lev, mult in LEVELS:
        i0 = int(np.argmin(np.abs(s - lev)))
        while i0 > 0 and s[i0 - 1] == lev:
            i0 -= 1
        out.append((i0, i0 + mult))
    return out


@pytest.mark.parametrize("shape", [(300, 257), (257, 300)])
def test_csvd_clustered_spectrum(ia, shape):
    A, _, s_true = _clustered(*shape, seed=21)
    U, s, V = ia.csvd(A)
    U0, s0, V0 = rt.csvd(A, driver="gesvd")
    p = min(shape)
    np.testing.assert_allclose(s, s0, rtol=0, atol=1e-13 * s0[0])
    np.testing.assert_allclose(U.conj().T @ U, np.eye(p), atol=1e-12)
    np.testing.assert_allclose(V.conj().T @ V, np.eye(p), atol=1e-12)
    np.testing.assert_allclose((U * s) @ V.conj().T, A, atol=1e-13 * np.sqrt(max(shape)) * s0[0])
    for i0, i1 in _multiplets(s_true):
        PU = U[:, i0:i1] @ U[:, i0:i1].conj().T
        PU0 = U0[:, i0:i1] @ U0[:, i0:i1].conj().T
        PV = V[:, i0:i1] @ V[:, i0:i1].conj().T
        PV0 = V0[:, i0:i1] @ V0[:, i0:i1].conj().T
        # the multiplet is separated from its neighbours by >= ~7 % of its value
        gap = 0.07 * s_true[i0]
        tol = max(1e-10, 1e-13 * s0[0] / gap)
        assert np.abs(PU - PU0).max() < tol, (i0, i1)
        assert np.abs(PV - PV0).max() < tol, (i0, i1)


@pytest.mark.parametrize("shape", [(300, 257), (257, 300), (1024, 2048), (2048, 1024)])
def test_regsolve_clustered_spectrum(ia, shape):
    """GCV-selected Tikhonov, TSVD and a fixed small lambda: the same lambda
    index and TSVD k as LAPACK, and x within max(1e-10, 10 x the spread of
    two backward-stable LAPACK SVDs)."""
    A, b, _ = _clustered(*shape, seed=22)
    for filt, sel, lam in (("tikhonov", "gcv", 0.0), ("tsvd", "gcv", 0.0),
                           ("tikhonov", "fixed", 1e-8), ("tikhonov", "fixed", 1e-4)):
        s_max = np.linalg.norm(A, 2)
        x, la, idx, kts, smax = ia.regsolve(A, b, filter=filt, select=sel, lam=lam * s_max)
        if sel == "fixed":
            U, s, V = rt.csvd(A)
            U2, s2, V2 = rt.csvd(A, driver="gesvd")
            xo = rt.tikhonov(U, s, V, b, lam * s_max)
            xo2 = rt.tikhonov(U2, s2, V2, b, lam * s_max)
        else:
            o = pipeline._regularised(A, b, filter=filt)
            o2 = pipeline._regularised(A, b, filter=filt, driver="gesvd")
            xo, xo2 = o["x"], o2["x"]
            assert kts[0] == o["k_tsvd"]
            if filt != "tsvd":
                assert idx[0] == o["idx"]
        spread = np.linalg.norm(xo - xo2) / np.linalg.norm(xo)
        # lambda >= 1e-4 s_1: the solution is well conditioned, the flat bar
        tol = 1e-10 if lam >= 1e-4 else max(1e-10, 10 * spread)
        err = np.linalg.norm(x[0] - xo) / np.linalg.norm(xo)
        print(f"{shape} {filt}/{sel}: err {err:.2e} spread {spread:.2e}")
        assert err <= tol, (filt, sel, err, spread)
