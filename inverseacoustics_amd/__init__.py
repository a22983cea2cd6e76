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
    "sphere", "cartesian_2_spherical", "geometry", "esm_sweep", "SweepResult", "zgemm", "cgemm", "zgesv",
    "FILTERS", "SELECTS",
]

FILTERS = {"tikhonov": 0, "tsvd": 1, "dsvd": 2}
SELECTS = {"gcv": 0, "lcurve": 1, "fixed": 2, "gcv_regtools": 3, "lcurve_regtools": 4,
           "quasiopt": 5}


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


def esm_sweep(k, r_s, src, r_h, b, area, normals=None, precision="fp64", filter="tikhonov",
              select="gcv", n_lambda=50, decades=8.0, lam=0.0, kfix=0, chunk=0):
    """Batched ESM sweep on the GPU: for every wavenumber k[f], assemble, solve
    H_e v = b[f] with the selected regularisation, reconstruct the surface
    pressure and sound power.  Returns a ``SweepResult``."""
    k = np.ascontiguousarray(k, dtype=np.float64).ravel()
    F = len(k)
    rs, sq, rh = soa(r_s), soa(src), soa(r_h)
    m, l = len(rs) // 3, len(rh) // 3
    b = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(F, l))
    nr = soa(normals) if normals is not None else None
    desc = _lib.ia_sweep_desc(m, l, F, 1 if precision == "mixed" else 0,
                              _reg(filter, select, n_lambda, decades, lam, kfix), float(area),
                              int(chunk))
    out = dict(lam=np.zeros(F), lam_index=np.zeros(F, np.int32), k_tsvd=np.zeros(F, np.int32),
               power=np.zeros(F), sigma_max=np.zeros(F), velocity=np.zeros((F, m), np.complex128),
               pressure=np.zeros((F, m), np.complex128), status=np.zeros(F, np.int32))
    o = _lib.ia_sweep_out(*(out[n].ctypes.data for n in
                            ("lam", "lam_index", "k_tsvd", "power", "sigma_max", "velocity",
                             "pressure", "status")))
    check(lib().ia_esm_sweep(C.byref(desc), _lib.dptr(k), _lib.dptr(rs), _lib.dptr(sq),
                             _lib.dptr(nr) if nr is not None else None, _lib.dptr(rh),
                             _lib.zptr(b), C.byref(o)))
    st = _lib.ia_stats()
    check(lib().ia_last_stats(C.byref(st)))
    stats = {n: getattr(st, n) for n, _ in _lib.ia_stats._fields_}
    return SweepResult(out["lam"], out["lam_index"], out["k_tsvd"], out["power"],
                       out["sigma_max"], out["velocity"], out["pressure"], out["status"], stats)
