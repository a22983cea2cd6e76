"""inverseacoustics_amd -- MI355X-native frequency-domain acoustic inverse solve.

Python mirror of the reference's operator interface (LuisCorralR/
InverseAcoustics ``CUDA/*.cuh``): the same function names, argument meaning
and output shapes, computed by the HIP library ``libia.so`` through its C ABI
(include/ia.h).  Arrays are numpy; complex results are complex128 in column-
major order as the reference returns them.  Importing works without a GPU;
calling any compute entry point without a GPU raises ``IAError``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import IAError, check, lib, soa, zf

__all__ = [
    "IAError", "set_air", "get_air", "device_count", "esm", "esmgpu", "esm_matrices", "csvd",
    "tsvd", "tikhonov", "dsvd", "regularize", "to_pressure", "to_velocity", "sound_power",
    "sphere", "cartesian_2_spherical", "geometry", "esm_sweep", "SweepResult", "zgemm", "cgemm", "zgesv", "hels", "hels_sweep", "regsolve",
    "quadrature_load", "set_quadrature", "ibem", "ibem_sweep", "invr", "radiation_modes", "radiation_sweep", "admm",
    "linesource", "esm_gradient", "write_file_complex", "write_file_double", "write_file_string", "write_matrix",
    "FILTERS", "SELECTS",
]

FILTERS = {"tikhonov": 0, "tsvd": 1, "dsvd": 2}
SELECTS = {"gcv": 0, "lcurve": 1, "fixed": 2, "gcv_regtools": 3, "lcurve_regtools": 4,
           "quasiopt": 5, "rank": 6}


def set_air(c: float, rho: float):
    """Speed of sound and density (replaces C_AIR/RHO_AIR, CUDA/constants.cuh:1-6)."""
    check(lib().ia_set_air(float(c), float(rho)))


def get_air():
    c, r = C.c_double(), C.c_double()
    check(lib().ia_get_air(C.byref(c), C.byref(r)))
    return c.value, r.value


def device_count() -> int:
    return int(lib().ia_device_count())


# ------------------------------------------------------------------ ESM
def esm(m, l, r_ms, r_lh, k, r_i):
    """CUDA/esm.cuh:112 ``esm``: returns (M_e (m x m), H_e (l x m))."""
    rs, rh = soa(r_ms), soa(r_lh)
    me, he = _lib._zp(), _lib._zp()
    lib().esm(int(m), int(l), _lib.dptr(rs), _lib.dptr(rh), C.byref(me), C.byref(he), float(k),
              float(r_i))
    return _lib.take_z(me, m * m, (m, m)), _lib.take_z(he, l * m, (l, m))


def esmgpu(m, l, r_msg, r_lhg, k, r_i):
    """CUDA/esm.cuh:38 ``esmgpu`` (same outputs as ``esm``)."""
    rs, rh = soa(r_msg), soa(r_lhg)
    me, he = _lib._zp(), _lib._zp()
    lib().esmgpu(int(m), int(l), _lib.dptr(rs), _lib.dptr(rh), C.byref(me), C.byref(he),
                 float(k), float(r_i))
    return _lib.take_z(me, m * m, (m, m)), _lib.take_z(he, l * m, (l, m))


def esm_matrices(r_s, r_h, src, k, normals=None, want=("M_e", "H_e", "D", "M_sq", "M_hq")):
    """General ESM (explicit sources and normals): dict of the requested matrices."""
    rs, rh, sq = soa(r_s), soa(r_h), soa(src)
    m, l = len(rs) // 3, len(rh) // 3
    nr = soa(normals) if normals is not None else None
    names = ("M_e", "H_e", "D", "M_sq", "M_hq")
    ptrs = {n: _lib._zp() for n in names}
    args = [C.byref(ptrs[n]) if n in want else None for n in names]
    check(lib().ia_esm(m, l, _lib.dptr(rs), _lib.dptr(rh), _lib.dptr(sq),
                       _lib.dptr(nr) if nr is not None else None, float(k), *args))
    shapes = {"M_e": (m, m), "H_e": (l, m), "D": (m, m), "M_sq": (m, m), "M_hq": (l, m)}
    return {n: _lib.take_z(ptrs[n], shapes[n][0] * shapes[n][1], shapes[n]) for n in want}


# ------------------------------------------------------------ SVD/filters
def csvd(A):
    """CUDA/tsvd.cuh:10 ``csvd``: compact SVD A = U diag(s) V^H."""
    A = zf(A)
    m, n = A.shape
    p = min(m, n)
    u, s, v = _lib._zp(), _lib._dp(), _lib._zp()
    check(lib().ia_csvd(m, n, _lib.zptr(A), C.byref(u), C.byref(s), C.byref(v)))
    if m >= n:
        return _lib.take_z(u, m * n, (m, n)), _lib.take_d(s, n), _lib.take_z(v, n * n, (n, n))
    return _lib.take_z(u, m * m, (m, m)), _lib.take_d(s, m), _lib.take_z(v, n * m, (n, m))


def _svd_args(U, s, V, b):
    U, V = zf(U), zf(V)
    s = np.ascontiguousarray(s, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    return U, s, V, b, U.shape[0], V.shape[0], s.shape[0]


def tsvd(U, s, V, b, k):
    """CUDA/tsvd.cuh:122 ``tsvd``: x = V[:, :k] (U[:, :p]^H b / s)[:k], 1 <= k <= p."""
    U, s, V, b, m, n, p = _svd_args(U, s, V, b)
    x = _lib._zp()
    lib().tsvd(m, n, p, int(k), _lib.zptr(U), _lib.dptr(s), _lib.zptr(V), _lib.zptr(b),
               C.byref(x))
    return _lib.take_z(x, n)


def tikhonov(U, s, V, b, lam):
    """MATLAB/regtools/tikhonovG.m (standard form)."""
    U, s, V, b, m, n, p = _svd_args(U, s, V, b)
    x = _lib._zp()
    check(lib().ia_tikhonov(m, n, p, _lib.zptr(U), _lib.dptr(s), _lib.zptr(V), _lib.zptr(b),
                            float(lam), C.byref(x)))
    return _lib.take_z(x, n)


def dsvd(U, s, V, b, lam):
    """MATLAB/regtools/dsvdG.m."""
    U, s, V, b, m, n, p = _svd_args(U, s, V, b)
    x = _lib._zp()
    check(lib().ia_dsvd(m, n, p, _lib.zptr(U), _lib.dptr(s), _lib.zptr(V), _lib.zptr(b),
                        float(lam), C.byref(x)))
    return _lib.take_z(x, n)


def _reg(filter="tikhonov", select="gcv", n_lambda=50, decades=8.0, lam=0.0, k=0):
    return _lib.ia_reg_params(FILTERS[filter], SELECTS[select], int(n_lambda), float(decades),
                              float(lam), int(k))


def regularize(U, s, V, b, filter="tikhonov", select="gcv", n_lambda=50, decades=8.0, lam=0.0,
               k=0):
    """Select the regularisation parameter and solve: returns (x, lambda, index).

    ``index`` is the grid index (Tikhonov/dsvd with a grid rule) or the TSVD
    truncation index.  ``select`` in {"gcv", "lcurve", "fixed",
    "gcv_regtools", "lcurve_regtools", "quasiopt"}."""
    U, s, V, b, m, n, p = _svd_args(U, s, V, b)
    prm = _reg(filter, select, n_lambda, decades, lam, k)
    x = _lib._zp()
    la, ix = C.c_double(), C.c_int()
    check(lib().ia_regularize(m, n, p, _lib.zptr(U), _lib.dptr(s), _lib.zptr(V), _lib.zptr(b),
                              C.byref(prm), C.byref(x), C.byref(la), C.byref(ix)))
    return _lib.take_z(x, n), la.value, ix.value


# ---------------------------------------------------------- dense kernels
def zgemm(A, B, opa="N", opb="N", alpha=1.0, beta=0.0, Cm=None):
    """C = alpha op(A) op(B) + beta C on the matrix cores (op 'N' or 'C')."""
    A, B = zf(A), zf(B)
    m = A.shape[0] if opa == "N" else A.shape[1]
    k = A.shape[1] if opa == "N" else A.shape[0]
    n = B.shape[1] if opb == "N" else B.shape[0]
    Cm = zf(np.zeros((m, n)) if Cm is None else Cm).copy(order="F")
    al, be = complex(alpha), complex(beta)
    check(lib().ia_zgemm(opa.encode(), opb.encode(), m, n, k, _lib.ia_z(al.real, al.imag),
                         _lib.zptr(A), max(A.shape[0], 1), _lib.zptr(B), max(B.shape[0], 1),
                         _lib.ia_z(be.real, be.imag), _lib.zptr(Cm), max(m, 1)))
    return Cm


def cgemm(A, B, opa="N", opb="N", alpha=1.0, beta=0.0, Cm=None):
    """complex64 product on the FP32 matrix cores (same contract as ``zgemm``)."""
    A = np.asfortranarray(A, dtype=np.complex64)
    B = np.asfortranarray(B, dtype=np.complex64)
    m = A.shape[0] if opa == "N" else A.shape[1]
    k = A.shape[1] if opa == "N" else A.shape[0]
    n = B.shape[1] if opb == "N" else B.shape[0]
    Cm = np.asfortranarray(np.zeros((m, n)) if Cm is None else Cm, dtype=np.complex64).copy(order="F")
    al, be = complex(alpha), complex(beta)
    check(lib().ia_cgemm(opa.encode(), opb.encode(), m, n, k, _lib.ia_c(al.real, al.imag),
                         A.ctypes.data, max(A.shape[0], 1), B.ctypes.data, max(B.shape[0], 1),
                         _lib.ia_c(be.real, be.imag), Cm.ctypes.data, max(m, 1)))
    return Cm


def zgesv(D, B, op="N"):
    """Solve D X = B (op 'N') or D^H X = B (op 'C') with the batched LU."""
    D = zf(D)
    B = zf(B).copy(order="F")
    vec = B.ndim == 1
    if vec:
        B = B.reshape(-1, 1, order="F")
    check(lib().ia_zgesv(op.encode(), D.shape[0], B.shape[1], _lib.zptr(D), _lib.zptr(B)))
    return B[:, 0] if vec else B


# ---------------------------------------------------------------- utils
def to_pressure(M, x):
    """CUDA/utils.cuh:199 ``to_pressure``: p = M x."""
    M = zf(M)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    m, n = M.shape
    p = _lib._zp()
    lib().to_pressure(m, n, _lib.zptr(M), _lib.zptr(x), C.byref(p))
    return _lib.take_z(p, m)


def to_velocity(Q, x):
    """CUDA/utils.cuh:232 ``to_velocity``: v = Q x."""
    Q = zf(Q)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    m, n = Q.shape
    v = _lib._zp()
    lib().to_velocity(m, n, _lib.zptr(Q), _lib.zptr(x), C.byref(v))
    return _lib.take_z(v, m)


def sound_power(p, v, S):
    """CUDA/utils.cuh:265 ``sound_power``: W = 1/2 Re(v^H p) S/m."""
    p = np.ascontiguousarray(p, dtype=np.complex128)
    v = np.ascontiguousarray(v, dtype=np.complex128)
    w = C.c_double(np.nan)
    lib().sound_power(len(p), _lib.zptr(p), _lib.zptr(v), C.byref(w), float(S))
    if not np.isfinite(w.value):
        raise IAError(2, lib().ia_last_error().decode())
    return w.value


def sphere(r_ms, r_lh, k, u, a):
    """CUDA/utils.cuh:384 ``sphere``: pulsating-sphere (p_s, p_h, v_s)."""
    rs, rh = soa(r_ms), soa(r_lh)
    m, l = len(rs) // 3, len(rh) // 3
    ps, ph, vs = _lib._zp(), _lib._zp(), _lib._zp()
    lib().sphere(_lib.dptr(rs), _lib.dptr(rh), C.byref(ps), C.byref(ph), C.byref(vs), float(k),
                 float(u), float(a), m, l)
    return _lib.take_z(ps, m), _lib.take_z(ph, l), _lib.take_z(vs, m)


def cartesian_2_spherical(pts):
    """CUDA/utils.cuh:414: (n,3) cartesian -> (n,3) (azimuth, elevation, radius)."""
    r = soa(pts)
    m = len(r) // 3
    out = _lib._dp()
    lib().cartesian_2_spherical(m, _lib.dptr(r), C.byref(out))
    return _lib.take_d(out, 3 * m).reshape(3, m).T.copy()


def geometry(file, k, t):
    """CUDA/utils.cuh:291 ``geometry``: (r_ms (m,3), r_lh (m,3), e_n (n,3))."""
    rs, rh, en = _lib._dp(), _lib._dp(), _lib._ip()
    m, l, n = C.c_int(), C.c_int(), C.c_int()
    check(lib().ia_geometry(str(file).encode(), float(k), float(t), C.byref(rs), C.byref(rh),
                            C.byref(en), C.byref(m), C.byref(l), C.byref(n)))
    m, l, n = m.value, l.value, n.value
    r_ms = _lib.take_d(rs, 3 * m).reshape(3, m).T.copy()
    r_lh = _lib.take_d(rh, 3 * l).reshape(3, l).T.copy()
    e_n = _lib.take_i(en, max(3 * n, 1))[: 3 * n].reshape(3, n).T.copy() if n else np.zeros((0, 3), int)
    return r_ms, r_lh, e_n


# ----------------------------------------------------------------- HELS
def hels(m, l, r_ms_s, r_lh_s, k, n_sh):
    """CUDA/hels.cuh:222 ``hels``: (M_e (m x n_t), H (l x n_t), Q_e (m x n_t)),
    inputs in spherical (azimuth, elevation, radius) coordinates."""
    rs, rh = soa(r_ms_s), soa(r_lh_s)
    n_t = (n_sh + 1) ** 2
    me, h, qe = _lib._zp(), _lib._zp(), _lib._zp()
    lib().hels(int(m), int(l), _lib.dptr(rs), _lib.dptr(rh), C.byref(me), C.byref(h), C.byref(qe),
               float(k), int(n_sh))
    return (_lib.take_z(me, m * n_t, (m, n_t)), _lib.take_z(h, l * n_t, (l, n_t)),
            _lib.take_z(qe, m * n_t, (m, n_t)))


def hels_sweep(k, order, azelr_s, azelr_h, b, area, precision="fp64", filter="tikhonov",
               select="gcv", n_lambda=50, decades=8.0, lam=0.0, kfix=0):
    """Batched HELS sweep (uneven orders): returns a ``SweepResult``."""
    k = np.ascontiguousarray(k, dtype=np.float64).ravel()
    F = len(k)
    order = np.ascontiguousarray(order, dtype=np.int32).ravel()
    rs, rh = soa(azelr_s), soa(azelr_h)
    m, l = len(rs) // 3, len(rh) // 3
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(F, l))
    desc = _lib.ia_hels_desc(m, l, F, 1 if precision == "mixed" else 0,
                             _reg(filter, select, n_lambda, decades, lam, kfix), float(area))
    out = _sweep_outputs(F, m)
    o = _sweep_struct(out)
    check(lib().ia_hels_sweep(C.byref(desc), order.ctypes.data_as(_lib._ip), _lib.dptr(k),
                              _lib.dptr(rs), _lib.dptr(rh), _lib.zptr(b), C.byref(o)))
    return SweepResult(out["lam"], out["lam_index"], out["k_tsvd"], out["power"], out["sigma_max"],
                       out["velocity"], out["pressure"], out["status"], {})


def regsolve(A, b, filter="tikhonov", select="gcv", n_lambda=50, decades=8.0, lam=0.0, k=0,
             precision="fp64"):
    """Batched regularised solve: A (batch, rows, cols), b (batch, rows) ->
    (x (batch, cols), lambda, index, k_tsvd, sigma_max)."""
    A = np.asarray(A, dtype=np.complex128)
    if A.ndim == 2:
        A = A[None]
    batch, rows, cols = A.shape
    Af = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))  # column-major per item
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(batch, rows))
    x = np.zeros((batch, cols), np.complex128)
    la, ks, sm = np.zeros(batch), np.zeros(batch, np.int32), np.zeros(batch)
    ix = np.zeros(batch, np.int32)
    prm = _reg(filter, select, n_lambda, decades, lam, k)
    check(lib().ia_regsolve(rows, cols, batch, _lib.zptr(Af), _lib.zptr(b), C.byref(prm),
                            1 if precision == "mixed" else 0, _lib.zptr(x), _lib.dptr(la),
                            ix.ctypes.data_as(_lib._ip), ks.ctypes.data_as(_lib._ip), _lib.dptr(sm)))
    return x, la, ix, ks, sm


def admm(A, b, lambda0, rho=1.0, gamma=1.0, epsilon=1e-6, max_iter=100, norm=1, p=0.5, n_gst=5):
    """ADMM regularised solution (MATLAB/ADMMG.m): A (batch, rows, cols) or
    (rows, cols), b (batch, rows) -> (c (batch, cols), iterations (batch,))."""
    A = np.asarray(A, dtype=np.complex128)
    single = A.ndim == 2
    if single:
        A = A[None]
    batch, rows, cols = A.shape
    Af = np.ascontiguousarray(np.transpose(A, (0, 2, 1)))
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(batch, rows))
    c = np.zeros((batch, cols), np.complex128)
    it = np.zeros(batch, np.int32)
    prm = _lib.ia_admm_params(float(lambda0), float(rho), float(gamma), float(epsilon),
                              int(max_iter), int(norm), float(p), int(n_gst))
    check(lib().ia_admm(rows, cols, batch, _lib.zptr(Af), _lib.zptr(b), C.byref(prm),
                        _lib.zptr(c), it.ctypes.data_as(_lib._ip)))
    return (c[0], it[0]) if single else (c, it)


# ----------------------------------------------------------------- IBEM
def quadrature_load(ipf_path, ipfs_path):
    """Read the reference's IPF.txt / IPFs.txt tables (CUDA/ibem.cuh:431-477):
    returns (rule (n_q, 3), pages (3, n_qs, 3)) with rows (xi1, xi2, w)."""
    nq, nqs = C.c_int(), C.c_int()
    ipf, ipfs = _lib._dp(), _lib._dp()
    check(lib().ia_quadrature_load(str(ipf_path).encode(), str(ipfs_path).encode(), C.byref(nq),
                                   C.byref(ipf), C.byref(nqs), C.byref(ipfs)))
    rule = _lib.take_d(ipf, 3 * nq.value).reshape(3, nq.value).T.copy()
    pages = _lib.take_d(ipfs, 9 * nqs.value).reshape(3, 3, nqs.value).transpose(0, 2, 1).copy()
    return rule, pages


def set_quadrature(rule=None, pages=None):
    """Use `rule` (n_q, 3) and singular `pages` (3, n_qs, 3) for the IBEM and
    radiation-mode entry points of this thread; no arguments restores the
    built-in rules (Radon 7-point, Duffy 40x40 Gauss-Legendre)."""
    if rule is None and pages is None:
        check(lib().ia_set_quadrature(0, None, 0, None))
        return
    rule = np.asarray(rule, dtype=np.float64)
    pages = np.asarray(pages, dtype=np.float64)
    r = np.ascontiguousarray(rule.T).ravel()
    p = np.ascontiguousarray(pages.transpose(0, 2, 1)).ravel()
    check(lib().ia_set_quadrature(rule.shape[0], _lib.dptr(r), pages.shape[1], _lib.dptr(p)))


def _mesh_args(r_ms, e_n, r_lh, cp):
    rs = soa(r_ms)
    E = np.asarray(e_n, dtype=np.int32)
    en = np.ascontiguousarray((E.T if E.ndim == 2 and E.shape[1] == 3 else E).ravel())
    rh = soa(r_lh) if r_lh is not None and len(r_lh) else np.zeros(0)
    c = soa(cp) if cp is not None and len(cp) else np.zeros(0)
    return rs, en, rh, c, len(rs) // 3, len(en) // 3, len(rh) // 3, len(c) // 3


def ibem(r_ms, e_n, r_lh, k, cp=None):
    """CUDA/ibem.cuh:424 ``ibem``: (Z (m x m), H_b (l x m)).  r_ms (m, 3),
    e_n (n, 3) 0-based triangles, r_lh (l, 3), cp (n_ch, 3) CHIEF points."""
    rs, en, rh, c, m, n, l, nch = _mesh_args(r_ms, e_n, r_lh, cp)
    z, h = _lib._zp(), _lib._zp()
    check(lib().ia_ibem(m, l, n, _lib.dptr(rs), en.ctypes.data_as(_lib._ip),
                        _lib.dptr(rh) if l else None, float(k), nch, _lib.dptr(c) if nch else None,
                        C.byref(z), C.byref(h) if l else None))
    Z = _lib.take_z(z, m * m, (m, m))
    return Z, (_lib.take_z(h, l * m, (l, m)) if l else np.zeros((0, m), np.complex128))


def ibem_sweep(k, r_ms, e_n, r_lh, b, area, cp=None, precision="fp64", filter="tikhonov",
               select="gcv", n_lambda=50, decades=8.0, lam=0.0, kfix=0, chunk=0):
    """Batched IBEM sweep: assemble Z, H_b per wavenumber, regularised solve
    of H_b v = b[f], p_s = Z v and W.  Returns a ``SweepResult``."""
    k = np.ascontiguousarray(k, dtype=np.float64).ravel()
    F = len(k)
    rs, en, rh, c, m, n, l, nch = _mesh_args(r_ms, e_n, r_lh, cp)
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(F, l))
    desc = _lib.ia_ibem_desc(m, l, n, nch, F, 1 if precision == "mixed" else 0,
                             _reg(filter, select, n_lambda, decades, lam, kfix), float(area),
                             int(chunk))
    out = _sweep_outputs(F, m)
    o = _sweep_struct(out)
    check(lib().ia_ibem_sweep(C.byref(desc), _lib.dptr(k), _lib.dptr(rs),
                              en.ctypes.data_as(_lib._ip), _lib.dptr(rh),
                              _lib.dptr(c) if nch else None, _lib.zptr(b), C.byref(o)))
    return SweepResult(out["lam"], out["lam_index"], out["k_tsvd"], out["power"], out["sigma_max"],
                       out["velocity"], out["pressure"], out["status"], {})


def invr(r_ms, e_n, r_lh, k, n_am, cp=None):
    """Radiation modes (CUDA/invr.cuh:457, MATLAB/invrS.m:157-170):
    returns (Q (m x n_am), M_r = Z Q, H_r = H_b Q, G (m x m))."""
    rs, en, rh, c, m, n, l, nch = _mesh_args(r_ms, e_n, r_lh, cp)
    q, mr, hr, g = _lib._zp(), _lib._zp(), _lib._zp(), _lib._zp()
    check(lib().ia_invr(m, l, n, _lib.dptr(rs), en.ctypes.data_as(_lib._ip),
                        _lib.dptr(rh) if l else None, float(k), nch, _lib.dptr(c) if nch else None,
                        int(n_am), C.byref(q), C.byref(mr), C.byref(hr) if l else None,
                        C.byref(g)))
    na = min(int(n_am), m)
    Q = _lib.take_z(q, m * na, (m, na))
    Mr = _lib.take_z(mr, m * na, (m, na))
    Hr = _lib.take_z(hr, l * na, (l, na)) if l else np.zeros((0, na), np.complex128)
    G = _lib.take_z(g, m * m, (m, m)).real.copy()
    return Q, Mr, Hr, G


def radiation_modes(pts, area, k, n_modes):
    """Elemental-radiator radiation modes: (eigenvalues (n_modes,), modes
    (np, n_modes)) of R_ij = (w^2 rho A^2/(4 pi c)) sin(k r_ij)/(k r_ij)."""
    p = soa(pts)
    npt = len(p) // 3
    ev, md = _lib._dp(), _lib._dp()
    check(lib().ia_radiation_modes(npt, _lib.dptr(p), float(area), float(k), int(n_modes),
                                   C.byref(ev), C.byref(md)))
    nm = min(int(n_modes), npt)
    return _lib.take_d(ev, nm), _lib.take_d(md, npt * nm).reshape(nm, npt).T.copy()


def radiation_sweep(k, r_p, area, r_h, b, n_modes, precision="fp64", filter="tikhonov",
                    select="gcv", n_lambda=50, decades=8.0, lam=0.0, kfix=0, chunk=0,
                    want_modes=False, eig_floor=0.0):
    """Batched radiation-mode sweep (ia_radiation_sweep): patch centres r_p
    (np, 3) of equal area, microphones r_h (l, 3), pressures b (F, l); modes
    with eigenvalues below eig_floor e_1 (0: 1e-8) are dropped.
    Returns a dict with power, lambda, lambda_index, k_tsvd (F,), eigval
    (F, n_modes), amplitude (F, n_modes), velocity (F, np) and, with
    want_modes, modes (F, np, n_modes)."""
    k = np.ascontiguousarray(k, dtype=np.float64).ravel()
    F = len(k)
    rp, rh = soa(r_p), soa(r_h)
    npt, l = len(rp) // 3, len(rh) // 3
    n = int(n_modes)
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(F, l))
    desc = _lib.ia_radiation_desc(npt, l, F, n, 1 if precision == "mixed" else 0,
                                  _reg(filter, select, n_lambda, decades, lam, kfix), float(area),
                                  int(chunk), float(eig_floor))
    out = dict(power=np.zeros(F), lam=np.zeros(F), lambda_index=np.zeros(F, np.int32),
               k_tsvd=np.zeros(F, np.int32), eigval=np.zeros((F, n)),
               modes=np.zeros((F, n, npt)) if want_modes else None,
               amplitude=np.zeros((F, n), np.complex128), velocity=np.zeros((F, npt), np.complex128),
               status=np.zeros(F, np.int32))
    o = _lib.ia_radiation_out(*((out[q].ctypes.data if out[q] is not None else None) for q in
                                ("power", "lam", "lambda_index", "k_tsvd", "eigval", "modes",
                                 "amplitude", "velocity", "status")))
    check(lib().ia_radiation_sweep(C.byref(desc), _lib.dptr(k), _lib.dptr(rp), _lib.dptr(rh),
                                   _lib.zptr(b), C.byref(o)))
    if want_modes:
        out["modes"] = out["modes"].transpose(0, 2, 1).copy()
    out["lambda"] = out.pop("lam")
    return out


# ------------------------------------------------------------ ESM sweep
@dataclass
class SweepResult:
    lam: np.ndarray
    lam_index: np.ndarray
    k_tsvd: np.ndarray
    power: np.ndarray
    sigma_max: np.ndarray
    velocity: np.ndarray   # (F, m)
    pressure: np.ndarray   # (F, m)
    status: np.ndarray
    stats: dict
    selection_gap: np.ndarray | None = None


def _sweep_outputs(F, m):
    return dict(lam=np.zeros(F), lam_index=np.zeros(F, np.int32), k_tsvd=np.zeros(F, np.int32),
                power=np.zeros(F), sigma_max=np.zeros(F), velocity=np.zeros((F, m), np.complex128),
                pressure=np.zeros((F, m), np.complex128), status=np.zeros(F, np.int32),
                gap=np.zeros(F))


def _sweep_struct(out, with_gap=False):
    o = _lib.ia_sweep_out(*(out[n].ctypes.data for n in
                            ("lam", "lam_index", "k_tsvd", "power", "sigma_max", "velocity",
                             "pressure", "status")))
    if with_gap:
        o.selection_gap = out["gap"].ctypes.data
    return o


def esm_sweep(k, r_s, src, r_h, b, area, normals=None, precision="fp64", filter="tikhonov",
              select="gcv", n_lambda=50, decades=8.0, lam=0.0, kfix=0, chunk=0, recheck_gap=0.0):
    """Batched ESM sweep on the GPU: for every wavenumber k[f], assemble, solve
    H_e v = b[f] with the selected regularisation, reconstruct the surface
    pressure and sound power.  Returns a ``SweepResult``.  In the mixed form a
    frequency whose parameter choice is within `recheck_gap` (relative; 0 =
    2e-4, negative = never) of a tie is solved again in FP64."""
    k = np.ascontiguousarray(k, dtype=np.float64).ravel()
    F = len(k)
    rs, sq, rh = soa(r_s), soa(src), soa(r_h)
    m, l = len(rs) // 3, len(rh) // 3
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(F, l))
    nr = soa(normals) if normals is not None else None
    desc = _lib.ia_sweep_desc(m, l, F, 1 if precision == "mixed" else 0,
                              _reg(filter, select, n_lambda, decades, lam, kfix), float(area),
                              int(chunk), float(recheck_gap))
    out = _sweep_outputs(F, m)
    o = _sweep_struct(out, with_gap=True)
    check(lib().ia_esm_sweep(C.byref(desc), _lib.dptr(k), _lib.dptr(rs), _lib.dptr(sq),
                             _lib.dptr(nr) if nr is not None else None, _lib.dptr(rh),
                             _lib.zptr(b), C.byref(o)))
    st = _lib.ia_stats()
    check(lib().ia_last_stats(C.byref(st)))
    stats = {n: getattr(st, n) for n, _ in _lib.ia_stats._fields_}
    return SweepResult(out["lam"], out["lam_index"], out["k_tsvd"], out["power"],
                       out["sigma_max"], out["velocity"], out["pressure"], out["status"], stats,
                       out["gap"])


# ------------------------------------------------------------ gradients
def esm_gradient(k, r_s, normals, d, r_h, b, area, lam=None, n_lambda=50, decades=8.0,
                 select="gcv", chunk=0):
    """Gradients of the ESM solve (ia_esm_gradient): sources at
    r_s - d * normals, Tikhonov solution at lam[f] (or at the lambda selected
    by `select` when lam is None, then held fixed).  Returns a dict with
    lambda, misfit, power, dmisfit_dd, dpower_dd (F,), dmisfit_dmic,
    dpower_dmic (F, l, 3), dmisfit_dloglam, dpower_dloglam (F,) and
    velocity (F, m)."""
    k = np.ascontiguousarray(k, dtype=np.float64).ravel()
    F = len(k)
    rs, nr, rh = soa(r_s), soa(normals), soa(r_h)
    m, l = len(rs) // 3, len(rh) // 3
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(F, l))
    desc = _lib.ia_grad_desc(m, l, F, _reg("tikhonov", select, n_lambda, decades), float(area),
                             int(chunk))
    out = {n: np.zeros(F) for n in ("lambda", "misfit", "power", "dmisfit_dd", "dpower_dd",
                                     "dmisfit_dloglam", "dpower_dloglam")}
    out["dmisfit_dmic"] = np.zeros((F, 3, l))
    out["dpower_dmic"] = np.zeros((F, 3, l))
    out["velocity"] = np.zeros((F, m), np.complex128)
    o = _lib.ia_grad_out(*(out[n].ctypes.data for n in
                           ("lambda", "misfit", "power", "dmisfit_dd", "dpower_dd",
                            "dmisfit_dmic", "dpower_dmic", "dmisfit_dloglam", "dpower_dloglam",
                            "velocity")))
    lam_arr = None if lam is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(lam, dtype=np.float64), (F,)))
    check(lib().ia_esm_gradient(C.byref(desc), _lib.dptr(k), _lib.dptr(rs), _lib.dptr(nr),
                                float(d), _lib.dptr(rh), _lib.zptr(b),
                                _lib.dptr(lam_arr) if lam_arr is not None else None, C.byref(o)))
    out["dmisfit_dmic"] = out["dmisfit_dmic"].transpose(0, 2, 1).copy()
    out["dpower_dmic"] = out["dpower_dmic"].transpose(0, 2, 1).copy()
    return out


# ------------------------------------------------------------ scenarios
def linesource(n_ls, r_ms, e_n, r_lh, rho, c, U0, k, a, L):
    """MATLAB/linesource.m ``[xp, xsf, vsf] = linesource(n_ls, r_ms, e_n,
    r_lh, rho, c, U0, k, a, L)``: pressure at the field points r_lh (l, 3),
    pressure and normal velocity at the surface nodes r_ms (m, 3) of the
    triangles e_n (n, 3, 0-based) for a line source of length L and width a
    on the y axis.  k may be a scalar (outputs (l,), (m,), (m,)) or an array
    of F wavenumbers (outputs (F, l), (F, m), (F, m))."""
    kk = np.ascontiguousarray(np.atleast_1d(np.asarray(k, dtype=np.float64)).ravel())
    F = len(kk)
    rs, en, rh, _, m, n, l, _ = _mesh_args(r_ms, e_n, r_lh, None)
    xp = np.zeros((F, l), np.complex128)
    xsf = np.zeros((F, m), np.complex128)
    vsf = np.zeros((F, m), np.complex128)
    check(lib().ia_linesource(int(n_ls), m, n, _lib.dptr(rs), en.ctypes.data_as(_lib._ip), l,
                              _lib.dptr(rh) if l else None, float(rho), float(c), float(U0),
                              _lib.dptr(kk), F, float(a), float(L),
                              _lib.zptr(xp) if l else None, _lib.zptr(xsf) if m else None,
                              _lib.zptr(vsf) if m else None))
    if np.ndim(k) == 0:
        return xp[0], xsf[0], vsf[0]
    return xp, xsf, vsf


# ---------------------------------------------------------- text output
def write_file_complex(mode, file, v):
    """CUDA/utils.cuh:32 ``write_file_cuDoubleComplex``: one line of
    "%f,%f" entries separated by ';' (mode "w" or "a")."""
    v = np.ascontiguousarray(np.asarray(v, dtype=np.complex128).ravel())
    check(lib().ia_write_complex(str(mode).encode(), str(file).encode(), _lib.zptr(v), len(v)))


def write_file_double(mode, file, v):
    """CUDA/utils.cuh:58 ``write_file_double``: one line of "%f" entries
    separated by ';'."""
    v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).ravel())
    check(lib().ia_write_double(str(mode).encode(), str(file).encode(), _lib.dptr(v), len(v)))


def write_file_string(mode, file, text):
    """CUDA/utils.cuh:50 ``write_file_string``."""
    check(lib().ia_write_string(str(mode).encode(), str(file).encode(), str(text).encode()))


def write_matrix(fmt, file, a):
    """CUDA/utils.cuh:116 ``write_matrix_cuDoubleComplex`` into `file`: rows
    of "%f,%f," (fmt "f") or "%e,%e," (fmt "e") entries, then a blank line."""
    A = zf(np.atleast_2d(a))
    r, cc = A.shape
    check(lib().ia_write_matrix(str(fmt).encode(), str(file).encode(), _lib.zptr(A), r, cc))
