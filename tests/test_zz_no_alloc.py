"""rt_spmspm_noalloc: the exact product into caller-provided arrays with its
scratch in a file-scope device array.  Called through the C ABI and checked
against rt_spmspm_ex (same structure; both within the oracle's bound), on
two calls in a row -- the second with a smaller input, so that the first
call's leftovers in the shared scratch must not matter -- plus the size
query, the scratch limit and the pool staying untouched."""
import ctypes

import numpy as np
import pytest

import oracle
import rtspmspm_amd as rt
from tests.helpers import check_parity

pytestmark = pytest.mark.gpu


def _dev_array(nbytes):
    p = ctypes.c_void_p()
    rt._check(rt._L.rt_device_alloc(ctypes.byref(p), max(int(nbytes), 1)))
    return p


def _to_host(ptr, count, dtype):
    out = np.zeros(count, dtype)
    if count:
        rt._check(rt._L.rt_memcpy_d2h(out.ctypes.data, ptr, out.nbytes))
    return out


def noalloc(dA, dB, capacity, rp, col, val):
    n = ctypes.c_int64()
    rc = rt._L.rt_spmspm_noalloc(ctypes.byref(dA._s), ctypes.byref(dB._s), rp, col, val,
                                 capacity, ctypes.byref(n), 0, None)
    return rc, n.value


def run_case(A, B):
    dA, dB = rt.to_device(A), rt.to_device(B)
    m = A.shape[0]
    rp = _dev_array(8 * (m + 1))
    try:
        # size query: capacity 0 -> RT_E_LIMIT, nnz(C) and C's row pointer
        before = rt.pool_stats()
        rc, need = noalloc(dA, dB, 0, rp, None, None)
        assert rc == rt.RT_E_LIMIT and need >= 0
        assert rt.pool_stats()["live_bytes"] == before["live_bytes"], "the query took pool memory"
        rp_q = _to_host(rp, m + 1, np.int64)
        col, val = _dev_array(4 * need), _dev_array(4 * need)
        try:
            before = rt.pool_stats()
            rc, n = noalloc(dA, dB, need, rp, col, val)
            assert rc == rt.RT_OK, rt._L.rt_last_error()
            assert n == need
            after = rt.pool_stats()
            assert after["live_bytes"] == before["live_bytes"], "the call took pool memory"
            got = rt.CSR((m, B.shape[1]), _to_host(rp, m + 1, np.int64),
                         _to_host(col, n, np.int32), _to_host(val, n, np.float32))
        finally:
            rt._L.rt_device_free(col)
            rt._L.rt_device_free(val)
        assert np.array_equal(got.row_ptr, rp_q)
        ref, _ = rt.spmspm(dA, dB)  # the existing entry point
        ref = ref.to_host()
        assert np.array_equal(got.row_ptr, ref.row_ptr) and np.array_equal(got.col, ref.col)
        err = check_parity(got, oracle.spgemm_f64(A, B))
        print(f"noalloc {A.shape}x{B.shape}: nnz {n}, max err/sum|ab| {err:.2e}")
    finally:
        rt._L.rt_device_free(rp)


def test_noalloc_matches_existing_entry_point_twice():
    # R-MAT scale 13: hash bins of every size and long rows (bitmap symbolic,
    # dense-tile numeric); then a smaller uniform product.
    A1 = rt.gen_rmat(13, 16, seed=101)
    run_case(A1, A1)
    A2 = rt.gen_uniform_k(3000, 20, 102)
    B2 = rt.gen_uniform_k(3000, 12, 103)
    run_case(A2, B2)


def test_noalloc_rejects_products_beyond_its_scratch():
    m = 5_000_000  # per-row scratch alone exceeds 64 MiB
    A = rt.CSR((m, 8), np.zeros(m + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32))
    B = rt.CSR((8, 8), np.zeros(9, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32))
    dA, dB = rt.to_device(A), rt.to_device(B)
    rp = _dev_array(8 * (m + 1))
    try:
        rc, n = noalloc(dA, dB, 0, rp, None, None)
        assert rc == rt.RT_E_LIMIT and n == -1
        assert b"MiB of scratch" in rt._L.rt_last_error()
    finally:
        rt._L.rt_device_free(rp)
