"""Parity at the benchmark's own sizes.

* R-MAT scale 20 (edge factor 16, a/b/c/d .57/.19/.19/.05, seed 1), C = A*A,
  plain and learned: nnz(C) = 9,703,855,703 (past 2^31 and 2^32, so C's row
  offsets need 64 bits), every row's nnz bounded by its product count, and
  the values and columns of a deterministic row sample checked against the
  FP64 oracle: every row with more than 2^20 products, 2,000 seeded uniform
  rows, and the rows whose C offsets straddle each multiple of 2^31.
* Galerkin R*(A*P) on the 128^3 grid: both products checked whole.
* Rectangular 500,000 x 2,000,000 * 2,000,000 x 500,000 (seed 3): whole.
* Batch of 2,048 pairs (seed 4), batched plain and learned: equal row
  pointers and columns, and 64 seeded pairs checked against the oracle.

The reference semantics are one accumulate C(i,j) += A(i,k)*B(k,j) per
product (optixSpMSpM.cu:245-259); the bar is a bit-exact structure and
|c - c64| <= 1e-5 * sum|a*b| per entry.  Each test prints the largest
|c - c64| / sum|a*b| it saw.
"""
import numpy as np
import pytest

import oracle
import rtspmspm_amd as rt
from rtspmspm_amd import workloads as W
from tests.helpers import check_parity

pytestmark = pytest.mark.gpu

RMAT20_NNZ_C = 9_703_855_703


def rows_of(dC, rp, rows, ncols):
    """Host CSR of the listed rows of a device result (row_ptr indexes the list)."""
    out_rp = np.zeros(len(rows) + 1, np.int64)
    np.cumsum(rp[np.asarray(rows) + 1] - rp[np.asarray(rows)], out=out_rp[1:])
    col = np.empty(int(out_rp[-1]), np.int32)
    val = np.empty(int(out_rp[-1]), np.float32)
    for q, i in enumerate(rows):
        c, v = dC.segment(rp[i], rp[i + 1])
        col[out_rp[q]:out_rp[q + 1]] = c
        val[out_rp[q]:out_rp[q + 1]] = v
    return rt.CSR((len(rows), ncols), out_rp, col, val)


@pytest.fixture(scope="module")
def rmat20():
    A = W.rmat(20)
    P = W.row_products(A, A)
    dA = rt.to_device(A)
    return {"A": A, "P": P, "dA": dA, "ref": {}, "rp": {}}


def _sample(rm, rp):
    """Hub rows, 2,000 seeded rows and the rows straddling k * 2^31."""
    P = rm["P"]
    hubs = np.nonzero(P > 2 ** 20)[0]
    rng = np.random.default_rng(2020)
    uni = rng.choice(P.size, 2000, replace=False)
    straddle = [int(np.searchsorted(rp, k * 2 ** 31, side="right") - 1)
                for k in range(1, int(rp[-1] // 2 ** 31) + 1)]
    return np.unique(np.concatenate([hubs, uni, straddle, [0, P.size - 1]])), hubs.size, straddle


@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_rmat20_full_size(rmat20, mode):
    A, P, dA = rmat20["A"], rmat20["P"], rmat20["dA"]
    rt.pool_trim()
    dC, info = rt.spmspm(dA, dA, mode=mode)
    try:
        assert info["products"] == int(P.sum())
        rp = dC.row_ptr_host()
        assert dC.nnz == RMAT20_NNZ_C and rp[-1] == RMAT20_NNZ_C
        nnz = np.diff(rp)
        assert np.all(nnz <= np.minimum(P, A.shape[1])) and np.array_equal(nnz > 0, P > 0)
        for other, orp in rmat20["rp"].items():
            assert np.array_equal(rp, orp), f"{mode} and {other} row pointers differ"
        rmat20["rp"][mode] = rp
        rows, nhub, straddle = _sample(rmat20, rp)
        for k, i in enumerate(straddle, 1):
            assert rp[i] < k * 2 ** 31 <= rp[i + 1]
        if "rows" not in rmat20["ref"]:
            rmat20["ref"]["rows"] = (rows, oracle.rows_f64(A, A, rows))
        srows, ref = rmat20["ref"]["rows"]
        assert np.array_equal(srows, rows)
        err = check_parity(rows_of(dC, rp, rows, A.shape[1]), ref)
        print(f"PARITY rmat20 {mode}: {rows.size} rows ({nhub} hub rows, "
              f"{int(ref.row_ptr[-1])} entries), max |c-c64|/sum|ab| = {err:.3e}")
    finally:
        dC.free()


@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_galerkin128_full_size(mode):
    A, Pm, R = W.galerkin(128)
    dA, dP, dR = rt.to_device(A), rt.to_device(Pm), rt.to_device(R)
    dAP, _ = rt.spmspm(dA, dP, mode=mode)
    AP = dAP.to_host()
    assert AP.nnz == 254 ** 3
    e1 = check_parity(AP, oracle.spgemm_f64(A, Pm))
    dC, _ = rt.spmspm(dR, dAP, mode=mode)
    C = dC.to_host()
    assert C.shape == (64 ** 3, 64 ** 3) and C.nnz == 190 ** 3
    e2 = check_parity(C, oracle.spgemm_f64(R, AP))
    print(f"PARITY galerkin128 {mode}: A*P max {e1:.3e}, R*(AP) max {e2:.3e}")


@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_rect_full_size(mode):
    A, B = W.rect()
    dA, dB = rt.to_device(A), rt.to_device(B)
    dC, info = rt.spmspm(dA, dB, mode=mode)
    assert info["products"] == int(W.row_products(A, B).sum())
    err = check_parity(dC.to_host(), oracle.spgemm_f64(A, B))
    print(f"PARITY rect {mode}: nnz {dC.nnz}, overflow rows {info['overflow_rows']}, "
          f"max {err:.3e}")


def test_batch2048_sample():
    ar, br, bc, SA, SB = W.batch()
    dSA, dSB = rt.to_device(SA), rt.to_device(SB)
    rt.pool_trim()
    dCp, ip = rt.spmspm_batched(ar, br, bc, dSA, dSB, mode="plain")
    rp_p = dCp.row_ptr_host()
    dCp.free()
    rt.pool_trim()
    dC, il = rt.spmspm_batched(ar, br, bc, dSA, dSB, mode="learned")
    try:
        rp = dC.row_ptr_host()
        assert np.array_equal(rp, rp_p) and il["products"] == ip["products"]
        a_start = np.concatenate([[0], np.cumsum(ar)])
        b_start = np.concatenate([[0], np.cumsum(br)])
        pick = np.sort(np.random.default_rng(404).choice(ar.size, 64, replace=False))
        worst = 0.0
        for p in pick:
            a0, n = int(a_start[p]), int(ar[p])
            b0, nb = int(b_start[p]), int(br[p])
            sa, ea = SA.row_ptr[a0], SA.row_ptr[a0 + n]
            sb, eb = SB.row_ptr[b0], SB.row_ptr[b0 + nb]
            Ap = rt.CSR((n, nb), SA.row_ptr[a0:a0 + n + 1] - sa, SA.col[sa:ea], SA.val[sa:ea])
            Bp = rt.CSR((nb, int(bc[p])), SB.row_ptr[b0:b0 + nb + 1] - sb, SB.col[sb:eb],
                        SB.val[sb:eb])
            col, val = dC.segment(rp[a0], rp[a0 + n])
            Cp = rt.CSR((n, int(bc[p])), rp[a0:a0 + n + 1] - rp[a0], col, val)
            worst = max(worst, check_parity(Cp, oracle.spgemm_f64(Ap, Bp)))
        print(f"PARITY batch2048 learned: nnz {dC.nnz}, products {il['products']}, "
              f"64 pairs max {worst:.3e}")
    finally:
        dC.free()
