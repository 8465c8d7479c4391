"""Python host side of the MI355X-native SpGEMM (C = A*B, CSR in, CSR out).

A thin ctypes layer over ``librtspmspm.so`` (C ABI in ``include/rt_spmspm.h``).
The reference exposes its method only as the ``optixSpMSpM`` program
(optixSpMSpM/src/optixSpMSpM/optixSpMSpM.cpp:496-660); the same program is
built here as ``rtspmspm_amd/optixSpMSpM``, and this module offers the same
operation to Python callers and tests:

    C, info = spmspm(A, B, mode="plain")        # host CSR in, host CSR out
    dA = to_device(A); dC, info = spmspm(dA, dB) # device-resident operands

Errors raise ``RtError`` carrying the library's code and message (the
reference turns every CUDA/OptiX failure into exit status 1,
optixSpMSpM.cpp:654-658).  There is no fallback: if the native library is
missing the import fails.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtspmspm.so")

RT_OK, RT_E_INVALID, RT_E_SHAPE, RT_E_NOMEM, RT_E_HIP, RT_E_IO, RT_E_LIMIT, RT_E_NODEVICE = \
    0, -1, -2, -3, -4, -5, -6, -7
MEM_HOST, MEM_DEVICE = 0, 1
MODES = {"plain": 0, "ray": 1, "learned": 2}
LEARNED_LONG = {None: 0, "learned": 1, "exact": 2}
FLAG_TRUSTED = 1
ABLATIONS = {"full": 0, "no_atomic": 1, "traverse": 2, "no_hit": 3, "empty": 4}
RAY_OUTPUTS = {"auto": 0, "dense": 1, "slot": 2}
NFEAT = 8


class RtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class rt_csr(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("row_ptr", ctypes.c_void_p), ("col", ctypes.c_void_p), ("val", ctypes.c_void_p),
                ("mem", ctypes.c_int32), ("owner", ctypes.c_int32)]


class rt_opts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("ablation", ctypes.c_int32),
                ("ray_output", ctypes.c_int32), ("profile", ctypes.c_int32),
                ("use_tuned", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("learned_long", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("ms_total", ctypes.c_double), ("products", ctypes.c_int64),
                ("overflow_rows", ctypes.c_int64), ("predicted_rows", ctypes.c_int64)]


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                      "(make -C rtspmspm_amd/csrc)")
_L = ctypes.CDLL(LIB_PATH)

_P = ctypes.POINTER
_c = ctypes
_sig = {
    "rt_spmspm": ([_P(rt_csr), _P(rt_csr), _P(rt_csr)], _c.c_int),
    "rt_spmspm_ex": ([_P(rt_csr), _P(rt_csr), _P(rt_csr), _P(rt_opts), _c.c_void_p], _c.c_int),
    "rt_spmspm_batched": ([_c.c_int32, _P(_c.c_int64), _P(_c.c_int64), _P(_c.c_int64),
                           _P(rt_csr), _P(rt_csr), _P(rt_csr), _P(rt_opts), _c.c_void_p], _c.c_int),
    "rt_spmspm_noalloc": ([_P(rt_csr), _P(rt_csr), _c.c_void_p, _c.c_void_p, _c.c_void_p,
                           _c.c_int64, _P(_c.c_int64), _c.c_int32, _c.c_void_p], _c.c_int),
    "rt_plan_create": ([_P(rt_csr), _P(rt_csr), _P(_c.c_void_p), _P(rt_csr), _c.c_void_p], _c.c_int),
    "rt_plan_execute": ([_c.c_void_p, _c.c_void_p, _c.c_void_p, _P(rt_csr), _P(_c.c_double)], _c.c_int),
    "rt_plan_destroy": ([_c.c_void_p], None),
    "rt_csr_free": ([_P(rt_csr)], None),
    "rt_last_error": ([], _c.c_char_p),
    "rt_csr_to_device": ([_P(rt_csr), _P(rt_csr)], _c.c_int),
    "rt_csr_to_host": ([_P(rt_csr), _P(rt_csr)], _c.c_int),
    "rt_memcpy_d2h": ([_c.c_void_p, _c.c_void_p, _c.c_int64], _c.c_int),
    "rt_device_copy": ([_c.c_void_p, _c.c_void_p, _c.c_int64, _P(_c.c_double)], _c.c_int),
    "rt_device_alloc": ([_P(_c.c_void_p), _c.c_int64], _c.c_int),
    "rt_device_free": ([_c.c_void_p], _c.c_int),
    "rt_device_synchronize": ([], _c.c_int),
    "rt_pool_trim": ([], _c.c_int),
    "rt_pool_stats": ([_P(_c.c_int64), _P(_c.c_int64), _P(_c.c_int64)], _c.c_int),
    "rt_debug_fail_alloc": ([_c.c_int64], _c.c_int),
    "rt_stream_create": ([_P(_c.c_void_p)], _c.c_int),
    "rt_stream_destroy": ([_c.c_void_p], _c.c_int),
    "rt_device_info": ([_c.c_char_p, _c.c_int64, _P(_c.c_int64), _P(_c.c_int32)], _c.c_int),
    "rt_autotune": ([_P(rt_csr), _P(rt_csr), _c.c_int32, _c.c_int32, _c.c_char_p, _c.c_int64],
                    _c.c_int),
    "rt_autotune_batched": ([_c.c_int32, _P(_c.c_int64), _P(_c.c_int64), _P(_c.c_int64),
                             _P(rt_csr), _P(rt_csr), _c.c_int32, _c.c_int32, _c.c_char_p,
                             _c.c_int64], _c.c_int),
    "rt_tune_cache_check": ([_c.c_char_p, _P(_c.c_int32), _P(_c.c_int32)], _c.c_int),
    "rt_prof_reset": ([], None),
    "rt_prof_count": ([], _c.c_int32),
    "rt_prof_get": ([_c.c_int32, _c.c_char_p, _c.c_int64, _P(_c.c_int64), _P(_c.c_double)], _c.c_int),
    "rt_row_features": ([_P(rt_csr), _P(rt_csr), _P(_c.c_float), _P(_c.c_int64)], _c.c_int),
    "rt_mlp_param_count": ([], _c.c_int64),
    "rt_mlp_set_weights": ([_P(_c.c_float), _c.c_int64, _c.c_float], _c.c_int),
    "rt_mlp_get_weights": ([_P(_c.c_float), _c.c_int64, _P(_c.c_float)], _c.c_int),
    "rt_mlp_train": ([_P(_c.c_float), _P(_c.c_float), _c.c_int64, _c.c_int32, _c.c_int32,
                      _c.c_float, _c.c_uint64, _P(_c.c_float)], _c.c_int),
    "rt_mlp_loss_grad": ([_P(_c.c_float), _P(_c.c_float), _P(_c.c_float), _c.c_int64,
                          _P(_c.c_float), _P(_c.c_float)], _c.c_int),
    "rt_mlp_predict": ([_P(_c.c_float), _P(_c.c_float), _c.c_int64, _P(_c.c_float)], _c.c_int),
    "rt_gen_uniform_k": ([_c.c_int64, _c.c_int64, _c.c_uint64, _P(rt_csr)], _c.c_int),
    "rt_gen_rmat": ([_c.c_int32, _c.c_int32, _c.c_double, _c.c_double, _c.c_double, _c.c_uint64,
                     _P(rt_csr)], _c.c_int),
    "rt_gen_stencil27": ([_c.c_int64, _c.c_uint64, _P(rt_csr)], _c.c_int),
    "rt_gen_aggregation": ([_c.c_int64, _P(rt_csr), _P(rt_csr)], _c.c_int),
    "rt_gen_lognormal_rows": ([_c.c_int64, _c.c_int64, _c.c_double, _c.c_double, _c.c_int64,
                               _c.c_int64, _c.c_uint64, _c.c_uint64, _P(rt_csr)], _c.c_int),
    "rt_gen_fixed_normal": ([_c.c_int64, _c.c_int64, _c.c_int64, _c.c_uint64, _c.c_uint64,
                             _P(rt_csr)], _c.c_int),
    "rt_gen_poisson": ([_c.c_int64, _c.c_double, _c.c_uint64, _c.c_uint64, _P(rt_csr)], _c.c_int),
    "rt_gen_batch_poisson": ([_c.c_int32, _c.c_int64, _c.c_int64, _c.c_double, _c.c_double,
                              _c.c_uint64, _P(_c.c_int64), _P(_c.c_int64), _P(_c.c_int64),
                              _P(rt_csr), _P(rt_csr)], _c.c_int),
    "rt_mtx_read": ([_c.c_char_p, _P(rt_csr)], _c.c_int),
    "rt_mtx_write": ([_c.c_char_p, _P(rt_csr), _c.c_int32], _c.c_int),
}
for _name, (_args, _res) in _sig.items():
    _f = getattr(_L, _name)
    _f.argtypes = _args
    _f.restype = _res

EXPORTED = tuple(_sig.keys())


def _check(rc: int) -> None:
    if rc != RT_OK:
        raise RtError(rc, _L.rt_last_error().decode(errors="replace"))


# ---------------------------------------------------------------------------
@dataclass
class CSR:
    """Host CSR: int64 row_ptr, int32 col (sorted per row), float32 val."""
    shape: tuple
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1]) if self.row_ptr.size else 0

    @staticmethod
    def from_scipy(M) -> "CSR":
        M = M.tocsr()
        M.sort_indices()
        M.sum_duplicates()
        return CSR((M.shape[0], M.shape[1]), M.indptr.astype(np.int64), M.indices.astype(np.int32),
                   M.data.astype(np.float32))

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.row_ptr), shape=self.shape)

    def _struct(self) -> rt_csr:
        self.row_ptr = np.ascontiguousarray(self.row_ptr, dtype=np.int64)
        self.col = np.ascontiguousarray(self.col, dtype=np.int32)
        self.val = np.ascontiguousarray(self.val, dtype=np.float32)
        return rt_csr(self.shape[0], self.shape[1], self.nnz, self.row_ptr.ctypes.data,
                      self.col.ctypes.data, self.val.ctypes.data, MEM_HOST, 0)


def _take_host(s: rt_csr) -> CSR:
    """Copy a library-owned host rt_csr into numpy arrays and free it."""
    rows, nnz = s.rows, s.nnz
    rp = np.ctypeslib.as_array(ctypes.cast(s.row_ptr, _P(_c.c_int64)), (rows + 1,)).copy()
    if nnz:
        col = np.ctypeslib.as_array(ctypes.cast(s.col, _P(_c.c_int32)), (nnz,)).copy()
        val = np.ctypeslib.as_array(ctypes.cast(s.val, _P(_c.c_float)), (nnz,)).copy()
    else:
        col = np.zeros(0, np.int32)
        val = np.zeros(0, np.float32)
    out = CSR((s.rows, s.cols), rp, col, val)
    _L.rt_csr_free(ctypes.byref(s))
    return out


class DeviceCSR:
    """A library-owned CSR in device memory (freed when garbage-collected)."""

    def __init__(self, s: rt_csr):
        self._s = s

    @property
    def shape(self):
        return (self._s.rows, self._s.cols)

    @property
    def nnz(self) -> int:
        return int(self._s.nnz)

    def row_ptr_host(self) -> np.ndarray:
        """Only the row pointer, copied to the host."""
        rp = np.zeros(self._s.rows + 1, np.int64)
        _check(_L.rt_memcpy_d2h(rp.ctypes.data, self._s.row_ptr, rp.nbytes))
        return rp

    def segment(self, e0: int, e1: int):
        """Entries [e0, e1) of col and val, copied to the host."""
        n = int(e1) - int(e0)
        col = np.zeros(max(n, 0), np.int32)
        val = np.zeros(max(n, 0), np.float32)
        if n > 0:
            _check(_L.rt_memcpy_d2h(col.ctypes.data, self._s.col + 4 * int(e0), 4 * n))
            _check(_L.rt_memcpy_d2h(val.ctypes.data, self._s.val + 4 * int(e0), 4 * n))
        return col, val

    def to_host(self) -> CSR:
        h = rt_csr()
        _check(_L.rt_csr_to_host(ctypes.byref(self._s), ctypes.byref(h)))
        return _take_host(h)

    def free(self) -> None:
        if self._s.owner:
            _L.rt_csr_free(ctypes.byref(self._s))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_device(A: CSR) -> DeviceCSR:
    d = rt_csr()
    h = A._struct()
    _check(_L.rt_csr_to_device(ctypes.byref(h), ctypes.byref(d)))
    return DeviceCSR(d)


GRAPH = {None: 0, True: 1, False: 2}


class Stream:
    """A HIP stream of the library's device (rt_stream_create), for running
    products concurrently from several host threads."""

    def __init__(self):
        self.handle = _c.c_void_p()
        _check(_L.rt_stream_create(ctypes.byref(self.handle)))

    def close(self) -> None:
        if self.handle:
            _check(_L.rt_stream_destroy(self.handle))
            self.handle = _c.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream(stream):
    return None if stream is None else stream.handle


def _opts(mode, ablation, ray_output, profile, use_tuned, graph=None, learned_long=None,
          trusted=False) -> rt_opts:
    o = rt_opts()
    o.mode = MODES[mode]
    o.ablation = ABLATIONS[ablation]
    o.ray_output = RAY_OUTPUTS[ray_output]
    o.profile = 1 if profile else 0
    o.use_tuned = 1 if use_tuned else 0
    o.graph = GRAPH[graph]
    o.learned_long = LEARNED_LONG[learned_long]
    o.flags = FLAG_TRUSTED if trusted else 0
    return o


def _info(o: rt_opts) -> dict:
    return {"ms": o.ms_total, "products": o.products, "overflow_rows": o.overflow_rows,
            "predicted_rows": o.predicted_rows}


def spmspm(A, B, mode: str = "plain", ablation: str = "full", ray_output: str = "auto",
           profile: bool = False, use_tuned: bool = False, graph=None, learned_long=None,
           trusted: bool = False, stream: Stream | None = None):
    """C = A*B.  Host CSR operands give a host CSR; DeviceCSR operands give a
    DeviceCSR.  Returns (C, info) with info = {ms, products, overflow_rows,
    predicted_rows}; ms is the device time from inputs resident to C complete.
    learned_long: None (tuned choice), "learned" or "exact" handling of the
    learned method's long rows; trusted: skip the device check of device
    operands; stream: run on this Stream instead of the library's."""
    o = _opts(mode, ablation, ray_output, profile, use_tuned, graph, learned_long, trusted)
    C = rt_csr()
    if isinstance(A, DeviceCSR):
        if not isinstance(B, DeviceCSR):
            raise TypeError("A and B must both be DeviceCSR or both host CSR")
        _check(_L.rt_spmspm_ex(ctypes.byref(A._s), ctypes.byref(B._s), ctypes.byref(C),
                               ctypes.byref(o), _stream(stream)))
        return DeviceCSR(C), _info(o)
    a, b = A._struct(), B._struct()
    _check(_L.rt_spmspm_ex(ctypes.byref(a), ctypes.byref(b), ctypes.byref(C), ctypes.byref(o),
                           _stream(stream)))
    return _take_host(C), _info(o)


def _pairs(a_rows, b_rows, b_cols):
    a_rows = np.ascontiguousarray(a_rows, np.int64)
    b_rows = np.ascontiguousarray(b_rows, np.int64)
    b_cols = np.ascontiguousarray(b_cols, np.int64)
    if not (a_rows.size == b_rows.size == b_cols.size):
        raise ValueError("a_rows, b_rows and b_cols must have one entry per pair")
    return a_rows, b_rows, b_cols


def spmspm_batched(a_rows, b_rows, b_cols, A_stack: DeviceCSR, B_stack: DeviceCSR,
                   mode: str = "plain", profile: bool = False, use_tuned: bool = False,
                   graph=None, learned_long=None, trusted: bool = False,
                   stream: Stream | None = None):
    """Independent products of stacked device operands (see rt_spmspm_batched).
    graph: None = library default (hipGraph stages for batched calls),
    True / False = force on / off."""
    a_rows, b_rows, b_cols = _pairs(a_rows, b_rows, b_cols)
    o = _opts(mode, "full", "auto", profile, use_tuned, graph, learned_long, trusted)
    C = rt_csr()
    _check(_L.rt_spmspm_batched(int(a_rows.size), a_rows.ctypes.data_as(_P(_c.c_int64)),
                                b_rows.ctypes.data_as(_P(_c.c_int64)),
                                b_cols.ctypes.data_as(_P(_c.c_int64)), ctypes.byref(A_stack._s),
                                ctypes.byref(B_stack._s), ctypes.byref(C), ctypes.byref(o),
                                _stream(stream)))
    return DeviceCSR(C), _info(o)


class Plan:
    """Fixed-structure product (rt_plan_*): the symbolic pass runs once at
    construction; execute() recomputes C's values for new A/B values."""

    def __init__(self, A: DeviceCSR, B: DeviceCSR):
        self._p = _c.c_void_p()
        C = rt_csr()
        _check(_L.rt_plan_create(ctypes.byref(A._s), ctypes.byref(B._s), ctypes.byref(self._p),
                                 ctypes.byref(C), None))
        self.C = DeviceCSR(C)
        self._operands = (A, B)  # the plan borrows their structure arrays

    def execute(self, A_values: DeviceCSR, B_values: DeviceCSR) -> float:
        """New values are taken from the val arrays of two device CSRs with
        the plan's structure; returns the device time in ms."""
        A, B = self._operands
        for new, old, name in ((A_values, A, "A"), (B_values, B, "B")):
            if not isinstance(new, DeviceCSR):
                raise TypeError(f"{name}_values must be a DeviceCSR")
            if new.shape != old.shape or new.nnz != old.nnz:
                raise ValueError(f"{name}_values has shape {new.shape} and {new.nnz} entries; "
                                 f"the plan's {name} has {old.shape} and {old.nnz}")
        ms = _c.c_double()
        _check(_L.rt_plan_execute(self._p, A_values._s.val, B_values._s.val,
                                  ctypes.byref(self.C._s), ctypes.byref(ms)))
        return ms.value

    def close(self) -> None:
        if self._p:
            _L.rt_plan_destroy(self._p)
            self._p = _c.c_void_p()

    def free(self) -> None:
        """Destroy the plan and release its C."""
        self.close()
        self.C.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_copy_probe(nbytes: int) -> float:
    """Time the library's streaming copy kernel over nbytes (device to device);
    returns ms.  Used to calibrate hardware byte counters."""
    src, dst = _c.c_void_p(), _c.c_void_p()
    _check(_L.rt_device_alloc(ctypes.byref(src), nbytes))
    _check(_L.rt_device_alloc(ctypes.byref(dst), nbytes))
    try:
        ms = _c.c_double()
        _check(_L.rt_device_copy(dst, src, nbytes, ctypes.byref(ms)))
        return ms.value
    finally:
        _L.rt_device_free(src)
        _L.rt_device_free(dst)


def autotune(A: DeviceCSR, B: DeviceCSR, mode: str = "plain", reps: int = 2) -> str:
    """Time the candidate row-bin settings on these operands and cache the
    fastest for their matrix class (used with use_tuned=True).  Returns the
    tuner's one-line report."""
    buf = ctypes.create_string_buffer(2048)
    _check(_L.rt_autotune(ctypes.byref(A._s), ctypes.byref(B._s), MODES[mode], reps, buf, 2048))
    return buf.value.decode()


def autotune_batched(a_rows, b_rows, b_cols, A_stack: DeviceCSR, B_stack: DeviceCSR,
                     mode: str = "plain", reps: int = 2) -> str:
    """As autotune() for batched calls with these stacked operands."""
    a_rows, b_rows, b_cols = _pairs(a_rows, b_rows, b_cols)
    buf = ctypes.create_string_buffer(2048)
    _check(_L.rt_autotune_batched(int(a_rows.size), a_rows.ctypes.data_as(_P(_c.c_int64)),
                                  b_rows.ctypes.data_as(_P(_c.c_int64)),
                                  b_cols.ctypes.data_as(_P(_c.c_int64)), ctypes.byref(A_stack._s),
                                  ctypes.byref(B_stack._s), MODES[mode], reps, buf, 2048))
    return buf.value.decode()


def tune_cache_check(path: str):
    """(valid, rejected) line counts of a tuning-cache file as the library reads it."""
    v, r = _c.c_int32(), _c.c_int32()
    _check(_L.rt_tune_cache_check(os.fsencode(path), ctypes.byref(v), ctypes.byref(r)))
    return v.value, r.value


def pool_stats() -> dict:
    lb, lk, cb = _c.c_int64(), _c.c_int64(), _c.c_int64()
    _check(_L.rt_pool_stats(ctypes.byref(lb), ctypes.byref(lk), ctypes.byref(cb)))
    return {"live_bytes": lb.value, "live_blocks": lk.value, "cached_bytes": cb.value}


def debug_fail_alloc(n: int) -> None:
    """Make the n-th pool allocation from now fail (0 disarms); test hook."""
    _check(_L.rt_debug_fail_alloc(int(n)))


def synchronize() -> None:
    _check(_L.rt_device_synchronize())


def pool_trim() -> None:
    _check(_L.rt_pool_trim())


def device_info() -> dict:
    name = ctypes.create_string_buffer(128)
    mem = _c.c_int64()
    cus = _c.c_int32()
    _check(_L.rt_device_info(name, 128, ctypes.byref(mem), ctypes.byref(cus)))
    return {"name": name.value.decode(), "total_mem": mem.value, "cus": cus.value}


def prof_reset() -> None:
    _L.rt_prof_reset()


def prof_stats() -> dict:
    """{kernel name: (launches, total ms)} since the last prof_reset()."""
    out = {}
    name = ctypes.create_string_buffer(128)
    n = _L.rt_prof_count()
    for i in range(n):
        l, ms = _c.c_int64(), _c.c_double()
        _check(_L.rt_prof_get(i, name, 128, ctypes.byref(l), ctypes.byref(ms)))
        out[name.value.decode()] = (l.value, ms.value)
    return out


# ---- generators ----------------------------------------------------------
def _gen(fn, *args) -> CSR:
    s = rt_csr()
    _check(fn(*args, ctypes.byref(s)))
    return _take_host(s)


def gen_uniform_k(n: int, k: int, seed: int) -> CSR:
    return _gen(_L.rt_gen_uniform_k, n, k, seed)


def gen_rmat(scale: int, edge_factor: int = 16, a=0.57, b=0.19, c=0.19, seed: int = 1) -> CSR:
    return _gen(_L.rt_gen_rmat, scale, edge_factor, a, b, c, seed)


def gen_stencil27(n: int, seed: int = 2) -> CSR:
    return _gen(_L.rt_gen_stencil27, n, seed)


def gen_aggregation(n: int):
    P, R = rt_csr(), rt_csr()
    _check(_L.rt_gen_aggregation(n, ctypes.byref(P), ctypes.byref(R)))
    return _take_host(P), _take_host(R)


def gen_lognormal_rows(rows, cols, mu, sigma, lo, hi, seed, stream=0) -> CSR:
    return _gen(_L.rt_gen_lognormal_rows, rows, cols, mu, sigma, lo, hi, seed, stream)


def gen_fixed_normal(rows, cols, k, seed, stream=1) -> CSR:
    return _gen(_L.rt_gen_fixed_normal, rows, cols, k, seed, stream)


def gen_poisson(n, lam, seed, stream=0) -> CSR:
    return _gen(_L.rt_gen_poisson, n, lam, seed, stream)


def gen_batch_poisson(count, nmin, nmax, lmin, lmax, seed):
    """Returns (a_rows, b_rows, b_cols, A_stack, B_stack) as host data."""
    ar = np.zeros(count, np.int64)
    br = np.zeros(count, np.int64)
    bc = np.zeros(count, np.int64)
    A, B = rt_csr(), rt_csr()
    _check(_L.rt_gen_batch_poisson(count, nmin, nmax, lmin, lmax, seed,
                                   ar.ctypes.data_as(_P(_c.c_int64)),
                                   br.ctypes.data_as(_P(_c.c_int64)),
                                   bc.ctypes.data_as(_P(_c.c_int64)), ctypes.byref(A),
                                   ctypes.byref(B)))
    return ar, br, bc, _take_host(A), _take_host(B)


# ---- Matrix Market ------------------------------------------------------
def mtx_read(path: str) -> CSR:
    s = rt_csr()
    _check(_L.rt_mtx_read(os.fsencode(path), ctypes.byref(s)))
    return _take_host(s)


def mtx_write(path: str, C: CSR, ref_compat=False) -> None:
    """ref_compat: False (every entry, 9 digits), True / 1 (the reference's
    format with the true count) or 2 (also its in-place header rewrite)."""
    h = C._struct()
    _check(_L.rt_mtx_write(os.fsencode(path), ctypes.byref(h), int(ref_compat)))


# ---- learned predictor ----------------------------------------------------
def _fp(a):
    return a.ctypes.data_as(_P(_c.c_float))


def row_features(A: CSR, B: CSR):
    """(features[rows, NFEAT] float32, exact row nnz int64), computed on the GPU."""
    feat = np.zeros((A.shape[0], NFEAT), np.float32)
    nnz = np.zeros(A.shape[0], np.int64)
    a, b = A._struct(), B._struct()
    _check(_L.rt_row_features(ctypes.byref(a), ctypes.byref(b), _fp(feat),
                              nnz.ctypes.data_as(_P(_c.c_int64))))
    return feat, nnz


def mlp_param_count() -> int:
    return int(_L.rt_mlp_param_count())


def mlp_set_weights(params: np.ndarray, margin: float) -> None:
    p = np.ascontiguousarray(params, np.float32)
    _check(_L.rt_mlp_set_weights(_fp(p), p.size, float(margin)))


def mlp_get_weights():
    p = np.zeros(mlp_param_count(), np.float32)
    m = _c.c_float()
    _check(_L.rt_mlp_get_weights(_fp(p), p.size, ctypes.byref(m)))
    return p, float(m.value)


def mlp_train(feat, target, epochs=20, batch=4096, lr=2e-3, seed=0) -> float:
    f = np.ascontiguousarray(feat, np.float32)
    t = np.ascontiguousarray(target, np.float32)
    loss = _c.c_float()
    _check(_L.rt_mlp_train(_fp(f), _fp(t), t.size, epochs, batch, lr, seed, ctypes.byref(loss)))
    return float(loss.value)


def mlp_loss_grad(params, feat, target):
    p = np.ascontiguousarray(params, np.float32)
    f = np.ascontiguousarray(feat, np.float32)
    t = np.ascontiguousarray(target, np.float32)
    g = np.zeros(p.size, np.float32)
    loss = _c.c_float()
    _check(_L.rt_mlp_loss_grad(_fp(p), _fp(f), _fp(t), t.size, ctypes.byref(loss), _fp(g)))
    return float(loss.value), g


def mlp_predict(params, feat):
    p = np.ascontiguousarray(params, np.float32)
    f = np.ascontiguousarray(feat, np.float32)
    out = np.zeros(f.shape[0], np.float32)
    _check(_L.rt_mlp_predict(_fp(p), _fp(f), f.shape[0], _fp(out)))
    return out
