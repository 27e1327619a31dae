"""CPU oracle for the binary-sparse 2-D R2C FFT -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker.  The product
(``cuspfft_amd``) never imports it and fails loudly without its HIP library.

Python wrappers (ctypes + numpy) around ``oracle/spfft_oracle.cpp``; see that
file's header for the reference file:line each function restates and how each
one is pinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libspfft_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_next_3_smooth.argtypes = [ctypes.c_int]
        L.orc_next_7_smooth.argtypes = [ctypes.c_int]
        L.orc_factor_3_smooth.argtypes = [ctypes.c_int, _i32p, ctypes.c_int]
        L.orc_can_pack_u16.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_compact_csc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                      _i32p, _i32p, _i32p]
        L.orc_bluestein_chirp.argtypes = [ctypes.c_int, _f32p]
        L.orc_bluestein_b.argtypes = [ctypes.c_int, ctypes.c_int, _f32p]
        L.orc_fft64.argtypes = [ctypes.c_int, _f64p, ctypes.c_int]
        L.orc_sparse_rfft2_f64.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                           _i32p, ctypes.c_int, _f64p, ctypes.c_int]
        L.orc_sparse_rfft2_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                           ctypes.c_int, _i32p, ctypes.c_int, _f32p, ctypes.c_int]
        _lib = L
    return _lib


def next_3_smooth(n: int) -> int:
    return lib().orc_next_3_smooth(int(n))


def next_7_smooth(n: int) -> int:
    return lib().orc_next_7_smooth(int(n))


def factor_3_smooth(n: int) -> list[int]:
    buf = np.zeros(128, np.int32)
    k = lib().orc_factor_3_smooth(int(n), buf, 128)
    if k < 0:
        raise RuntimeError("factor_3_smooth: not 3-smooth")
    return buf[:k].tolist()


def can_pack_u16(rows: int, cols: int) -> bool:
    return bool(lib().orc_can_pack_u16(int(rows), int(cols)))


def _coo(row, col):
    return (np.ascontiguousarray(row, dtype=np.int32), np.ascontiguousarray(col, dtype=np.int32))


def compact_csc(rows: int, cols: int, row, col):
    """Returns (sparse_cols, col_ptr, row_idx) as int32 arrays."""
    r, c = _coo(row, col)
    nnz = len(r)
    sc = np.zeros(max(cols, 1), np.int32)
    cp = np.zeros(max(cols, 1) + 1, np.int32)
    ri = np.zeros(max(nnz, 1), np.int32)
    n_sp = lib().orc_compact_csc(rows, cols, nnz, r, c, sc, cp, ri)
    return sc[:n_sp].copy(), cp[:n_sp + 1].copy(), ri[:nnz].copy()


def bluestein_chirp(n: int) -> np.ndarray:
    out = np.zeros(2 * n, np.float32)
    lib().orc_bluestein_chirp(n, out)
    return out.view(np.complex64)


def bluestein_b(n: int, M: int) -> np.ndarray:
    out = np.zeros(2 * M, np.float32)
    lib().orc_bluestein_b(n, M, out)
    return out.view(np.complex64)


def fft64(x, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, np.complex128)).copy()
    buf = a.view(np.float64)
    lib().orc_fft64(len(a), buf, int(sign))
    return buf.view(np.complex128)


def sparse_rfft2_f64(rows: int, cols: int, row, col, u_list=None, nthreads: int = 0) -> np.ndarray:
    """FP64 X[u][v] for u in u_list (default: all rows), v <= cols//2."""
    r, c = _coo(row, col)
    u = np.arange(rows, dtype=np.int32) if u_list is None else np.ascontiguousarray(u_list, np.int32)
    V = cols // 2 + 1
    out = np.zeros(len(u) * V * 2, np.float64)
    rc = lib().orc_sparse_rfft2_f64(rows, cols, len(r), r, c, u, len(u), out, nthreads)
    if rc != 0:
        raise ValueError("orc_sparse_rfft2_f64 failed")
    return out.view(np.complex128).reshape(len(u), V)


def sparse_rfft2_f32(rows: int, cols: int, row, col, smooth: int = 3, u_list=None,
                     nthreads: int = 0) -> np.ndarray:
    """FP32 restatement of the reference's plain algorithm, rows u_list."""
    r, c = _coo(row, col)
    u = np.arange(rows, dtype=np.int32) if u_list is None else np.ascontiguousarray(u_list, np.int32)
    V = cols // 2 + 1
    out = np.zeros(len(u) * V * 2, np.float32)
    rc = lib().orc_sparse_rfft2_f32(rows, cols, len(r), r, c, int(smooth), u, len(u), out, nthreads)
    if rc != 0:
        raise ValueError("orc_sparse_rfft2_f32 failed")
    return out.view(np.complex64).reshape(len(u), V)


def rel_l2(a, b) -> float:
    """||a - b||_2 / ||b||_2 in FP64."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def write_reference(path: str, rows: int, cols: int, row, col, nthreads: int = 0) -> np.ndarray:
    """Writes the FP64 half spectrum of a binary COO matrix as a reference file for the
    benchmark driver's --check-ref: the 8 bytes "SPFFTREF", int32 rows, int32 cols, then
    rows x (cols//2 + 1) complex128 values, row-major.  Returns the spectrum."""
    X = sparse_rfft2_f64(rows, cols, row, col, nthreads=nthreads)
    with open(path, "wb") as f:
        f.write(b"SPFFTREF")
        f.write(np.asarray([rows, cols], np.int32).tobytes())
        f.write(np.ascontiguousarray(X, np.complex128).tobytes())
    return X
