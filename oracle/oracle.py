"""ctypes wrapper of dft_oracle.c plus the reference's error metrics.

TEST INFRASTRUCTURE ONLY (see package docstring).
"""
from __future__ import annotations

import ctypes as ct
import math
import os
import subprocess
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def load():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = ct.CDLL(path)
        i64p = ct.POINTER(ct.c_int64)
        L.oracle_dft_axis.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, i64p, ct.c_int, ct.c_int]
        L.oracle_dft_axis.restype = ct.c_int
        L.oracle_fft_axis.argtypes = [ct.c_void_p, ct.c_int, i64p, ct.c_int, ct.c_int]
        L.oracle_fft_axis.restype = ct.c_int
        L.oracle_scale.argtypes = [ct.c_void_p, ct.c_int64, ct.c_double]
        L.oracle_scale.restype = None
        L.oracle_threads.argtypes = []
        L.oracle_threads.restype = ct.c_int
        L.oracle_uniform.argtypes = [ct.c_void_p, ct.c_int64, ct.c_uint64]
        L.oracle_uniform.restype = None
        L.oracle_mt64_raw.argtypes = [ct.c_void_p, ct.c_int64, ct.c_uint64]
        L.oracle_mt64_raw.restype = None
        L.oracle_gaussian.argtypes = [ct.c_void_p, ct.c_int64, ct.c_uint64]
        L.oracle_gaussian.restype = None
        _LIB = L
    return _LIB


def _dims(shape):
    return (ct.c_int64 * len(shape))(*shape)


def dftn(x: np.ndarray, axes: Sequence[int] | None = None, sign: int = -1) -> np.ndarray:
    """Unnormalised DFT of x over `axes` (direct O(n^2) per line, FP64 result)."""
    L = load()
    a = np.ascontiguousarray(x, dtype=np.complex128)
    axes = range(a.ndim) if axes is None else axes
    d = _dims(a.shape)
    for ax in axes:
        out = np.empty_like(a)
        rc = L.oracle_dft_axis(a.ctypes.data, out.ctypes.data, a.ndim, d, int(ax), int(sign))
        if rc:
            raise RuntimeError(f"oracle_dft_axis rc={rc}")
        a = out
    return a


def fftn_pow2(x: np.ndarray, axes: Sequence[int] | None = None, sign: int = -1) -> np.ndarray:
    """Same transform by the radix-2 FFT (power-of-two axes), in place on a copy."""
    L = load()
    a = np.array(x, dtype=np.complex128, order="C", copy=True)
    axes = range(a.ndim) if axes is None else axes
    d = _dims(a.shape)
    for ax in axes:
        rc = L.oracle_fft_axis(a.ctypes.data, a.ndim, d, int(ax), int(sign))
        if rc:
            raise RuntimeError(f"oracle_fft_axis rc={rc}")
    return a


def threads() -> int:
    return int(load().oracle_threads())


def uniform_input(shape, seed: int, dtype=np.complex128) -> np.ndarray:
    """std::mt19937_64(seed); per element re then im; value 2*((x>>11)*2^-53)-1.

    complex64 inputs are the FP64 draws rounded to float."""
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.complex128)
    load().oracle_uniform(out.ctypes.data, n, seed)
    return out.reshape(shape).astype(dtype, copy=False)


def gaussian_input(shape, seed: int, dtype=np.complex128) -> np.ndarray:
    """std::mt19937_64(seed), Box-Muller on consecutive draws per element:
    u1 = ((x0>>11)+1)*2^-53 in (0,1], u2 = (x1>>11)*2^-53, r = sqrt(-2 ln u1),
    value r*(cos 2 pi u2 + i sin 2 pi u2).  complex64 inputs are the FP64
    draws rounded to float."""
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.complex128)
    load().oracle_gaussian(out.ctypes.data, n, seed)
    return out.reshape(shape).astype(dtype, copy=False)


def mt64_raw(count: int, seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    load().oracle_mt64_raw(out.ctypes.data, count, seed)
    return out


def norm_factor_nd(dims, exponent: int, prec: int) -> float:
    """(1/sqrt(prod dims))^e in long double, rounded once to the working precision
    (src/cpp/shafft_internal.cpp:136,329).  numpy longdouble is x87 80-bit here."""
    den = np.sqrt(np.prod(np.array(dims, dtype=np.longdouble)))
    f = np.longdouble(1.0) / den ** exponent
    return float(np.float32(f)) if prec == 0 else float(np.float64(f))


def norm_factor_1d(n: int, exponent: int, prec: int) -> float:
    """fft1d_method.hip:663-668: float path sqrt in float, pow in double."""
    if prec == 0:
        s = np.sqrt(np.float32(n))  # std::sqrt(float)
        return float(np.float32(1.0 / np.float64(s) ** exponent))
    return 1.0 / math.pow(math.sqrt(float(n)), exponent)


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def check_rel_error(result, expected, tol: float) -> bool:
    """tests/include/test_utils.hpp:84-122: relative L-inf and L2 both within
    tol * sqrt(ln N)."""
    r = np.asarray(result, dtype=np.complex128).ravel()
    e = np.asarray(expected, dtype=np.complex128).ravel()
    n = r.size
    scale = math.sqrt(math.log(max(n, 2)))
    linf = np.max(np.abs(r - e)) / max(np.max(np.abs(e)), 1e-300)
    l2 = rel_l2(r, e)
    return bool(linf <= tol * scale and l2 <= tol * scale)
