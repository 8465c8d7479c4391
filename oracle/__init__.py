"""CPU oracle for C = A*B -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, as the checker or as the timed CPU baseline.
The library ``rtspmspm_amd`` never uses it.

It wraps ``liboracle.so`` (``oracle.c``), a C restatement of the reference's
product semantics: one accumulate ``C(i,j) += A(i,k)*B(k,j)`` per ray/sphere
hit (optixSpMSpM.cu:143-150,233-259; Sphere.cpp:72-75), with the structure of
C the union pattern of the products, cancellation zeros kept and columns
sorted (the structure of the two-phase comparator, cuSparse/src/main.cpp:464-539).

How each function is pinned (tests/test_oracle.py):
  spgemm_f64 / count   scipy.sparse values on the 9 shipped fixtures
                        (golden files under tests/golden/), pattern product
                        |A|*|B| for structure, dense brute force on n <= 300,
                        identity / permutation / diagonal-scaling closed forms,
                        an explicit cancellation case, empty rows, rectangular.
  spgemm_f32            bit-exact against integer-valued products (all-ones
                        inputs), within the fp32 summation bound of spgemm_f64
                        elsewhere, and against a pure-Python loop on tiny cases
                        that follows the same A-row-then-B-row order.
  stream_f32            nnz and checksum equal those of spgemm_f32's result.
  rows_f64              equals the matching rows of spgemm_f64.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i64 = ctypes.c_int64
        L.orc_count.argtypes = [i64, i64, _i64p, _i32p, _i64p, _i32p, _i64p]
        L.orc_fill_f64.argtypes = [i64, i64, _i64p, _i32p, _f32p, _i64p, _i32p, _f32p, _i64p,
                                   _i32p, _f64p, _f64p]
        L.orc_fill_f32.argtypes = [i64, i64, _i64p, _i32p, _f32p, _i64p, _i32p, _f32p, _i64p,
                                   _i32p, _f32p, ctypes.c_int]
        L.orc_stream_f32.argtypes = [i64, i64, i64, _i64p, _i32p, _f32p, _i64p, _i32p, _f32p,
                                     ctypes.c_int, _f64p]
        L.orc_rows_f64.argtypes = [i64, _i64p, i64, _i64p, _i32p, _f32p, _i64p, _i32p, _f32p,
                                   _i64p, _i32p, _f64p, _f64p]
        for f in (L.orc_count, L.orc_fill_f64, L.orc_fill_f32, L.orc_stream_f32, L.orc_rows_f64):
            f.restype = None
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def _ops(A, B):
    arp = np.ascontiguousarray(A.row_ptr, dtype=np.int64)
    acol = np.ascontiguousarray(A.col, dtype=np.int32)
    aval = np.ascontiguousarray(A.val, dtype=np.float32)
    brp = np.ascontiguousarray(B.row_ptr, dtype=np.int64)
    bcol = np.ascontiguousarray(B.col, dtype=np.int32)
    bval = np.ascontiguousarray(B.val, dtype=np.float32)
    if A.shape[1] != B.shape[0]:
        raise ValueError("shape mismatch")
    return arp, acol, aval, brp, bcol, bval


@dataclass
class OracleResult:
    shape: tuple
    row_ptr: np.ndarray   # int64
    col: np.ndarray       # int32
    val: np.ndarray       # float64 (spgemm_f64) or float32 (spgemm_f32)
    absum: np.ndarray | None = None  # sum |a*b| per entry (spgemm_f64 only)


def count(A, B) -> np.ndarray:
    arp, acol, aval, brp, bcol, bval = _ops(A, B)
    nnz = np.zeros(A.shape[0], dtype=np.int64)
    _lib().orc_count(A.shape[0], B.shape[1], _p(arp, _i64p), _p(acol, _i32p), _p(brp, _i64p),
                     _p(bcol, _i32p), _p(nnz, _i64p))
    return nnz


def _rowptr(nnz):
    rp = np.zeros(nnz.size + 1, dtype=np.int64)
    np.cumsum(nnz, out=rp[1:])
    return rp


def spgemm_f64(A, B) -> OracleResult:
    arp, acol, aval, brp, bcol, bval = _ops(A, B)
    rp = _rowptr(count(A, B))
    n = int(rp[-1])
    col = np.empty(max(n, 1), dtype=np.int32)
    val = np.empty(max(n, 1), dtype=np.float64)
    ab = np.empty(max(n, 1), dtype=np.float64)
    _lib().orc_fill_f64(A.shape[0], B.shape[1], _p(arp, _i64p), _p(acol, _i32p), _p(aval, _f32p),
                        _p(brp, _i64p), _p(bcol, _i32p), _p(bval, _f32p), _p(rp, _i64p),
                        _p(col, _i32p), _p(val, _f64p), _p(ab, _f64p))
    return OracleResult((A.shape[0], B.shape[1]), rp, col[:n], val[:n], ab[:n])


def spgemm_f32(A, B, threads: int = 0) -> OracleResult:
    arp, acol, aval, brp, bcol, bval = _ops(A, B)
    rp = _rowptr(count(A, B))
    n = int(rp[-1])
    col = np.empty(max(n, 1), dtype=np.int32)
    val = np.empty(max(n, 1), dtype=np.float32)
    _lib().orc_fill_f32(A.shape[0], B.shape[1], _p(arp, _i64p), _p(acol, _i32p), _p(aval, _f32p),
                        _p(brp, _i64p), _p(bcol, _i32p), _p(bval, _f32p), _p(rp, _i64p),
                        _p(col, _i32p), _p(val, _f32p), int(threads))
    return OracleResult((A.shape[0], B.shape[1]), rp, col[:n], val[:n])


def stream_f32(A, B, r0: int, r1: int, threads: int = 0):
    """(nnz, checksum) of rows [r0, r1) computed and discarded."""
    arp, acol, aval, brp, bcol, bval = _ops(A, B)
    out = np.zeros(2, dtype=np.float64)
    _lib().orc_stream_f32(int(r0), int(r1), B.shape[1], _p(arp, _i64p), _p(acol, _i32p),
                          _p(aval, _f32p), _p(brp, _i64p), _p(bcol, _i32p), _p(bval, _f32p),
                          int(threads), _p(out, _f64p))
    return int(out[0]), float(out[1])


def rows_f64(A, B, rows) -> OracleResult:
    """FP64 result restricted to the listed rows (row_ptr indexes the list)."""
    arp, acol, aval, brp, bcol, bval = _ops(A, B)
    sel = np.ascontiguousarray(rows, dtype=np.int64)
    off = np.zeros(sel.size + 1, dtype=np.int64)
    L = _lib()
    L.orc_rows_f64(sel.size, _p(sel, _i64p), B.shape[1], _p(arp, _i64p), _p(acol, _i32p),
                   _p(aval, _f32p), _p(brp, _i64p), _p(bcol, _i32p), _p(bval, _f32p),
                   _p(off, _i64p), None, None, None)
    np.cumsum(off[1:], out=off[1:])
    n = int(off[-1])
    col = np.empty(max(n, 1), dtype=np.int32)
    val = np.empty(max(n, 1), dtype=np.float64)
    ab = np.empty(max(n, 1), dtype=np.float64)
    L.orc_rows_f64(sel.size, _p(sel, _i64p), B.shape[1], _p(arp, _i64p), _p(acol, _i32p),
                   _p(aval, _f32p), _p(brp, _i64p), _p(bcol, _i32p), _p(bval, _f32p),
                   _p(off, _i64p), _p(col, _i32p), _p(val, _f64p), _p(ab, _f64p))
    return OracleResult((sel.size, B.shape[1]), off, col[:n], val[:n], ab[:n])
