"""ctypes binding of ``libia.so`` (the C ABI declared in include/ia.h).

The library is built in-tree by ``python -m inverseacoustics_amd.build``;
loading fails loudly when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libia.so"


class ia_z(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class ia_c(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class ia_reg_params(C.Structure):
    _fields_ = [("filter", C.c_int), ("select", C.c_int), ("n_lambda", C.c_int),
                ("decades", C.c_double), ("lambda_", C.c_double), ("k", C.c_int)]


class ia_sweep_desc(C.Structure):
    _fields_ = [("m", C.c_int), ("l", C.c_int), ("n_freq", C.c_int), ("precision", C.c_int),
                ("reg", ia_reg_params), ("area", C.c_double), ("chunk", C.c_int),
                ("recheck_gap", C.c_double)]


class ia_sweep_out(C.Structure):
    _fields_ = [("lambda_", C.c_void_p), ("lambda_index", C.c_void_p), ("k_tsvd", C.c_void_p),
                ("power", C.c_void_p), ("sigma_max", C.c_void_p), ("velocity", C.c_void_p),
                ("pressure", C.c_void_p), ("status", C.c_void_p), ("selection_gap", C.c_void_p)]


class ia_hels_desc(C.Structure):
    _fields_ = [("m", C.c_int), ("l", C.c_int), ("n_freq", C.c_int), ("precision", C.c_int),
                ("reg", ia_reg_params), ("area", C.c_double)]


class ia_ibem_desc(C.Structure):
    _fields_ = [("m", C.c_int), ("l", C.c_int), ("n_elem", C.c_int), ("n_ch", C.c_int),
                ("n_freq", C.c_int), ("precision", C.c_int), ("reg", ia_reg_params),
                ("area", C.c_double), ("chunk", C.c_int)]


class ia_grad_desc(C.Structure):
    _fields_ = [("m", C.c_int), ("l", C.c_int), ("n_freq", C.c_int), ("reg", ia_reg_params),
                ("area", C.c_double), ("chunk", C.c_int)]


class ia_grad_out(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("lambda_", "misfit", "power", "dmisfit_dd", "dpower_dd", "dmisfit_dmic",
                 "dpower_dmic", "dmisfit_dloglam", "dpower_dloglam", "velocity")]


class ia_radiation_desc(C.Structure):
    _fields_ = [("np_", C.c_int), ("l", C.c_int), ("n_freq", C.c_int), ("n_modes", C.c_int),
                ("precision", C.c_int), ("reg", ia_reg_params), ("area", C.c_double),
                ("chunk", C.c_int), ("eig_floor", C.c_double)]


class ia_radiation_out(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("power", "lambda_", "lambda_index", "k_tsvd", "eigval", "modes", "amplitude",
                 "velocity", "status")]


class ia_admm_params(C.Structure):
    _fields_ = [("lambda0", C.c_double), ("rho", C.c_double), ("gamma", C.c_double),
                ("epsilon", C.c_double), ("max_iter", C.c_int), ("norm", C.c_int),
                ("p", C.c_double), ("n_gst", C.c_int)]


class ia_stats(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("assemble", "factor", "solve_he", "bidiag", "bisect", "vectors", "select",
                 "post", "total", "bidiag_kernel_ms", "bidiag_kernel_bytes",
                 "bidiag_kernel_launches", "rechecked", "recheck_ms", "recheck_full")]


_zp = C.POINTER(ia_z)
_zpp = C.POINTER(_zp)
_dp = C.POINTER(C.c_double)
_dpp = C.POINTER(_dp)
_ip = C.POINTER(C.c_int)
_ipp = C.POINTER(_ip)

# name -> (restype, argtypes); every symbol of include/ia.h
SIGNATURES = {
    "ia_last_error": (C.c_char_p, []),
    "ia_version": (C.c_char_p, []),
    "ia_set_air": (C.c_int, [C.c_double, C.c_double]),
    "ia_get_air": (C.c_int, [_dp, _dp]),
    "ia_device_count": (C.c_int, []),
    "ia_release": (None, []),
    "esm": (None, [C.c_int, C.c_int, _dp, _dp, _zpp, _zpp, C.c_double, C.c_double]),
    "esmgpu": (None, [C.c_int, C.c_int, _dp, _dp, _zpp, _zpp, C.c_double, C.c_double]),
    "ia_esm": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double,
                         _zpp, _zpp, _zpp, _zpp, _zpp]),
    "csvd": (None, [C.c_int, C.c_int, _zp, _zpp, _dpp, _zpp]),
    "ia_csvd": (C.c_int, [C.c_int, C.c_int, _zp, _zpp, _dpp, _zpp]),
    "tsvd": (None, [C.c_int, C.c_int, C.c_int, C.c_int, _zp, _dp, _zp, _zp, _zpp]),
    "ia_tikhonov": (C.c_int, [C.c_int, C.c_int, C.c_int, _zp, _dp, _zp, _zp, C.c_double, _zpp]),
    "ia_dsvd": (C.c_int, [C.c_int, C.c_int, C.c_int, _zp, _dp, _zp, _zp, C.c_double, _zpp]),
    "to_pressure": (None, [C.c_int, C.c_int, _zp, _zp, _zpp]),
    "to_velocity": (None, [C.c_int, C.c_int, _zp, _zp, _zpp]),
    "sound_power": (None, [C.c_int, _zp, _zp, _dp, C.c_double]),
    "sphere": (None, [_dp, _dp, _zpp, _zpp, _zpp, C.c_double, C.c_double, C.c_double, C.c_int,
                      C.c_int]),
    "ia_geometry": (C.c_int, [C.c_char_p, C.c_double, C.c_double, _dpp, _dpp, _ipp, _ip, _ip, _ip]),
    "geometry": (None, [_dpp, _dpp, _ipp, C.c_double, C.c_double, _ip, _ip, _ip, C.c_char_p]),
    "cartesian_2_spherical": (None, [C.c_int, _dp, _dpp]),
    "ia_regularize": (C.c_int, [C.c_int, C.c_int, C.c_int, _zp, _dp, _zp, _zp,
                                C.POINTER(ia_reg_params), _zpp, _dp, _ip]),
    "ia_regularize_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.POINTER(ia_reg_params),
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ia_esm_sweep": (C.c_int, [C.POINTER(ia_sweep_desc), _dp, _dp, _dp, _dp, _dp, _zp,
                               C.POINTER(ia_sweep_out)]),
    "ia_esm_sweep_device": (C.c_int, [C.POINTER(ia_sweep_desc), C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(ia_sweep_out), C.c_void_p]),
    "ia_last_stats": (C.c_int, [C.POINTER(ia_stats)]),
    "ia_zgemm": (C.c_int, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, ia_z, _zp, C.c_int, _zp,
                           C.c_int, ia_z, _zp, C.c_int]),
    "ia_zgesv": (C.c_int, [C.c_char, C.c_int, C.c_int, _zp, _zp]),
    "hels": (None, [C.c_int, C.c_int, _dp, _dp, _zpp, _zpp, _zpp, C.c_double, C.c_int]),
    "helsgpu": (None, [C.c_int, C.c_int, _dp, _dp, _zpp, _zpp, _zpp, C.c_double, C.c_int]),
    "ia_hels": (C.c_int, [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, _zpp, _zpp, _zpp]),
    "ia_hels_sweep": (C.c_int, [C.POINTER(ia_hels_desc), _ip, _dp, _dp, _dp, _zp,
                                C.POINTER(ia_sweep_out)]),
    "ia_regsolve": (C.c_int, [C.c_int, C.c_int, C.c_int, _zp, _zp, C.POINTER(ia_reg_params), C.c_int,
                              _zp, _dp, _ip, _ip, _dp]),
    "ia_esm_gradient": (C.c_int, [C.POINTER(ia_grad_desc), _dp, _dp, _dp, C.c_double, _dp, _zp,
                                  _dp, C.POINTER(ia_grad_out)]),
    "ia_radiation_sweep": (C.c_int, [C.POINTER(ia_radiation_desc), _dp, _dp, _dp, _zp,
                                     C.POINTER(ia_radiation_out)]),
    "ia_quadrature_load": (C.c_int, [C.c_char_p, C.c_char_p, _ip, _dpp, _ip, _dpp]),
    "ia_set_quadrature": (C.c_int, [C.c_int, _dp, C.c_int, _dp]),
    "ibem": (None, [C.c_int, C.c_int, C.c_int, _dp, _ip, _dp, _zpp, _zpp, C.c_double, C.c_int,
                    _dp]),
    "ibemgpu": (None, [C.c_int, C.c_int, C.c_int, _dp, _ip, _dp, _zpp, _zpp, C.c_double, C.c_int,
                       _dp]),
    "ia_ibem": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _ip, _dp, C.c_double, C.c_int, _dp,
                          _zpp, _zpp]),
    "ia_ibem_sweep": (C.c_int, [C.POINTER(ia_ibem_desc), _dp, _dp, _ip, _dp, _dp, _zp,
                                C.POINTER(ia_sweep_out)]),
    "invr": (None, [C.c_int, C.c_int, C.c_int, _dp, _ip, _dp, _zpp, _zpp, _zpp, C.c_double,
                    C.c_int, _dp, C.c_int]),
    "invrgpu": (None, [C.c_int, C.c_int, C.c_int, _dp, _ip, _dp, _zpp, _zpp, _zpp, C.c_double,
                       C.c_int, _dp, C.c_int]),
    "ia_invr": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _ip, _dp, C.c_double, C.c_int, _dp,
                          C.c_int, _zpp, _zpp, _zpp, _zpp]),
    "ia_radiation_modes": (C.c_int, [C.c_int, _dp, C.c_double, C.c_double, C.c_int, _dpp,
                                     _dpp]),
    "ia_admm": (C.c_int, [C.c_int, C.c_int, C.c_int, _zp, _zp, C.POINTER(ia_admm_params), _zp,
                          _ip]),
    "ia_cgemm": (C.c_int, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, ia_c, C.c_void_p, C.c_int,
                           C.c_void_p, C.c_int, ia_c, C.c_void_p, C.c_int]),
    "ia_linesource": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _ip, C.c_int, _dp, C.c_double,
                                C.c_double, C.c_double, _dp, C.c_int, C.c_double, C.c_double,
                                _zp, _zp, _zp]),
    "ia_write_complex": (C.c_int, [C.c_char_p, C.c_char_p, _zp, C.c_int]),
    "ia_write_double": (C.c_int, [C.c_char_p, C.c_char_p, _dp, C.c_int]),
    "ia_write_string": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p]),
    "ia_write_matrix": (C.c_int, [C.c_char_p, C.c_char_p, _zp, C.c_int, C.c_int]),
    "write_file_cuDoubleComplex": (None, [C.c_char_p, C.c_char_p, _zp, C.c_int]),
    "write_file_string": (None, [C.c_char_p, C.c_char_p, C.c_char_p]),
    "write_file_double": (None, [C.c_char_p, C.c_char_p, _dp, C.c_int]),
    "write_matrix_cuDoubleComplex": (None, [C.c_char_p, _zp, C.c_int, C.c_int]),
    "print_matrix_cuDoubleComplex": (None, [C.c_char_p, _zp, C.c_int, C.c_int]),
    "print_matrix_double": (None, [C.c_char_p, _dp, C.c_int, C.c_int]),
}

_lib = None
_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]
_libc.free.restype = None


def lib():
    """Load libia.so (once); raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m inverseacoustics_amd.build` "
                "(the HIP library is required; there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class IAError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def check(status):
    if status != 0:
        raise IAError(status, lib().ia_last_error().decode(errors="replace"))


def free(ptr):
    if ptr:
        _libc.free(C.cast(ptr, C.c_void_p))


def zptr(a: np.ndarray):
    return a.ctypes.data_as(_zp)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def take_z(pp, count, shape=None, order="F"):
    """Copy a malloc'd ia_z* of `count` entries into complex128 and free it."""
    if not pp:
        raise IAError(1, lib().ia_last_error().decode(errors="replace") or "NULL output")
    arr = np.ctypeslib.as_array(C.cast(pp, C.POINTER(C.c_double)), shape=(2 * count,)).copy()
    free(pp)
    out = arr.view(np.complex128)
    return out.reshape(shape, order=order) if shape is not None else out


def take_d(pp, count):
    if not pp:
        raise IAError(1, lib().ia_last_error().decode(errors="replace") or "NULL output")
    arr = np.ctypeslib.as_array(pp, shape=(count,)).copy()
    free(pp)
    return arr


def take_i(pp, count):
    if not pp:
        raise IAError(1, "NULL output")
    arr = np.ctypeslib.as_array(pp, shape=(count,)).copy()
    free(pp)
    return arr


def zf(a) -> np.ndarray:
    """complex128 Fortran-ordered copy."""
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


def soa(pts) -> np.ndarray:
    pts = np.asarray(pts, dtype=np.float64)
    if pts.ndim == 2 and pts.shape[1] == 3:
        return np.ascontiguousarray(pts.T).ravel()
    return np.ascontiguousarray(pts.ravel())
