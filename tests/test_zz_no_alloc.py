"""ia_regularize_noalloc (the allocation-free variant of ia_regularize, for
callers under graph capture or in serving loops) against ia_regularize,
through the C ABI with device buffers: two calls in a row, the second with a
smaller problem, for each filter / selection rule."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(m, n, seed):
    from oracle import regtools as rt

    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    U0, s0, Vh0 = np.linalg.svd(A, full_matrices=False)
    s0 = s0 * 10.0 ** (-7 * np.arange(len(s0)) / len(s0))
    A = (U0 * s0) @ Vh0
    b = A @ (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    b = b + 1e-5 * np.linalg.norm(b) / np.sqrt(m) * rng.standard_normal(m)
    U, s, V = rt.csvd(A)
    return U, s, V, b


class _Hip:
    """The caller's device buffers, allocated with the HIP runtime libia.so
    is linked against (already loaded with it)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.bufs.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.mark.parametrize("filt,sel,lam,k", [("tikhonov", "gcv", 0.0, 0), ("tsvd", "gcv", 0.0, 0),
                                            ("tikhonov", "lcurve", 0.0, 0), ("dsvd", "gcv", 0.0, 0),
                                            ("tsvd", "fixed", 0.0, 7), ("tikhonov", "fixed", 1e-3, 0)])
def test_noalloc_matches_regularize(ia, filt, sel, lam, k):
    from inverseacoustics_amd import _lib

    L = _lib.lib()
    hip = _Hip()
    try:
        for (m, n, seed) in ((300, 180, 1), (96, 150, 2)):  # the second call smaller
            U, s, V, b = _problem(m, n, seed)
            p = len(s)
            lam_abs = lam * s[0]
            x0, lam0, idx0 = ia.regularize(U, s, V, b, filt, sel, lam=lam_abs, k=k)
            du = hip.put(np.asfortranarray(U).ravel(order="F"))
            dv = hip.put(np.asfortranarray(V).ravel(order="F"))
            db, ds = hip.put(b.astype(np.complex128)), hip.put(s.astype(np.float64))
            dx = hip.put(np.full(n, np.nan + 1j * np.nan))
            dl, di = hip.put(np.zeros(1)), hip.put(np.full(1, -5, np.int32))
            prm = ia._reg(filt, sel, 50, 8.0, lam_abs, k)
            ia.check(L.ia_regularize_noalloc(m, n, p, du, ds, dv, db, C.byref(prm), dx, dl, di,
                                             None))
            x = hip.get(dx, np.zeros(n, np.complex128))
            idx = int(hip.get(di, np.zeros(1, np.int32))[0])
            lam_g = float(hip.get(dl, np.zeros(1))[0])
            assert idx == idx0, (m, n, idx, idx0)
            assert lam_g == pytest.approx(lam0, rel=1e-14, abs=0)
            # FP64 bar: the two paths sum U^H b in different orders, and the
            # selected lambda amplifies that rounding by up to s_1 / lambda
            assert np.linalg.norm(x - x0) <= 1e-10 * np.linalg.norm(x0), (m, n)
            hip.free()
    finally:
        hip.free()


def test_noalloc_rejects_oversized(ia):
    from inverseacoustics_amd import _lib

    prm = ia._reg("tikhonov", "gcv")
    with pytest.raises(ia.IAError):
        # p beyond the 4096 scratch limit is refused before any pointer is read
        ia.check(_lib.lib().ia_regularize_noalloc(5000, 5000, 5000, 1, 1, 1, 1, C.byref(prm), 1,
                                                  None, None, None))
