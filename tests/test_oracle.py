"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each check is chosen so that a plausible mistake -- a transposed operand, a
dropped product, a wrong accumulation order, a dropped cancellation zero --
fails at least one of them (see test_oracle_detects_planted_mistakes)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from rtspmspm_amd import CSR
from tests.helpers import (FIXTURES, check_against_golden, csr_from_coo, csr_from_dense, golden,
                           load_fixture)


def as_csr64(res):
    return CSR(res.shape, res.row_ptr, res.col, res.val)


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_f64_matches_golden_fixtures(name):
    A = load_fixture(name)
    g = golden()[name]
    res = oracle.spgemm_f64(A, A)
    check_against_golden(as_csr64(res), g)
    # values: scipy's FP64 product, entry by entry (scipy drops exact zeros)
    S = sp.csr_matrix((A.val.astype(np.float64), A.col, A.row_ptr), shape=A.shape)
    ref = (S @ S).toarray() if A.shape[0] <= 12000 else None
    if ref is not None:
        D = np.zeros(A.shape)
        rows = np.repeat(np.arange(A.shape[0]), np.diff(res.row_ptr))
        D[rows, res.col] = res.val
        assert np.allclose(D, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", [n for n in FIXTURES if n != "email-Enron"])
def test_oracle_f32_bit_exact_on_integer_products(name):
    """All-ones inputs give integer sums < 2^24, exact in fp32 in any order."""
    A = load_fixture(name)
    if not golden()[name]["all_ones"]:
        pytest.skip("fixture has non-integer values")
    r64 = oracle.spgemm_f64(A, A)
    r32 = oracle.spgemm_f32(A, A)
    assert np.array_equal(r32.row_ptr, r64.row_ptr) and np.array_equal(r32.col, r64.col)
    assert np.array_equal(r32.val.astype(np.float64), r64.val)


def dense_product(A: CSR, B: CSR):
    Da = A.to_scipy().toarray().astype(np.float64)
    Db = B.to_scipy().toarray().astype(np.float64)
    return Da @ Db, (np.abs(Da) > 0).astype(np.int64) @ (np.abs(Db) > 0).astype(np.int64)


def pattern(A: CSR):
    M = np.zeros(A.shape, bool)
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.row_ptr))
    M[rows, A.col] = True
    return M


def rand_csr(rng, m, n, density, explicit_zeros=0.0):
    mask = rng.random((m, n)) < density
    r, c = np.nonzero(mask)
    v = rng.uniform(-1, 1, r.size).astype(np.float32)
    if explicit_zeros:
        v[rng.random(r.size) < explicit_zeros] = 0.0
    return csr_from_coo(m, n, r, c, v)


@pytest.mark.parametrize("seed,m,k,n,dens", [(0, 50, 60, 70, 0.1), (1, 200, 150, 180, 0.03),
                                             (2, 300, 300, 300, 0.02), (3, 1, 5, 1, 0.9),
                                             (4, 120, 1, 90, 0.5)])
def test_oracle_brute_force(seed, m, k, n, dens):
    rng = np.random.default_rng(seed)
    A = rand_csr(rng, m, k, dens, explicit_zeros=0.05)
    B = rand_csr(rng, k, n, dens, explicit_zeros=0.05)
    res = oracle.spgemm_f64(A, B)
    D, _ = dense_product(A, B)
    # structure = union pattern of stored entries (explicit zeros included)
    Pa = pattern(A).astype(np.int64)
    Pb = pattern(B).astype(np.int64)
    S = (Pa @ Pb) > 0
    got = np.zeros((m, n), bool)
    rows = np.repeat(np.arange(m), np.diff(res.row_ptr))
    got[rows, res.col] = True
    assert np.array_equal(got, S)
    for i in range(m):  # sorted, unique columns
        seg = res.col[res.row_ptr[i]:res.row_ptr[i + 1]]
        assert np.all(np.diff(seg) > 0)
    V = np.zeros((m, n))
    V[rows, res.col] = res.val
    assert np.allclose(V, D, rtol=1e-12, atol=1e-13)
    Ab = np.abs(A.to_scipy().toarray().astype(np.float64)) @ np.abs(B.to_scipy().toarray().astype(np.float64))
    AB = np.zeros((m, n))
    AB[rows, res.col] = res.absum
    assert np.allclose(AB, Ab, rtol=1e-12, atol=1e-13)


def test_oracle_identity_and_permutation():
    rng = np.random.default_rng(7)
    B = rand_csr(rng, 64, 40, 0.2)
    I = csr_from_dense(np.eye(64))
    r = oracle.spgemm_f64(I, B)
    assert np.array_equal(r.row_ptr, B.row_ptr) and np.array_equal(r.col, B.col)
    assert np.array_equal(r.val, B.val.astype(np.float64))
    perm = rng.permutation(64)
    Pm = np.zeros((64, 64))
    Pm[np.arange(64), perm] = 1.0
    r = oracle.spgemm_f64(csr_from_dense(Pm), B)
    Db = B.to_scipy().toarray()
    V = np.zeros((64, 40))
    rows = np.repeat(np.arange(64), np.diff(r.row_ptr))
    V[rows, r.col] = r.val
    assert np.array_equal(V, Db[perm].astype(np.float64))


def test_oracle_diagonal_scaling_closed_form():
    rng = np.random.default_rng(8)
    B = rand_csr(rng, 50, 50, 0.1)
    d = rng.uniform(0.5, 2.0, 50).astype(np.float32)
    r = oracle.spgemm_f64(csr_from_dense(np.diag(d)), B)
    rows = np.repeat(np.arange(50), np.diff(B.row_ptr))
    assert np.array_equal(r.col, B.col)
    assert np.allclose(r.val, d[rows].astype(np.float64) * B.val.astype(np.float64), rtol=1e-15)


def test_oracle_keeps_cancellation_zero():
    # [1, 1] * [x, -x]^T = 0: the entry stays in the structure
    A = csr_from_coo(1, 2, [0, 0], [0, 1], [1.0, 1.0])
    B = csr_from_coo(2, 1, [0, 1], [0, 0], [0.75, -0.75])
    r = oracle.spgemm_f64(A, B)
    assert r.row_ptr.tolist() == [0, 1] and r.col.tolist() == [0] and r.val.tolist() == [0.0]
    assert r.absum.tolist() == [1.5]
    r32 = oracle.spgemm_f32(A, B)
    assert r32.col.tolist() == [0] and r32.val.tolist() == [0.0]


def test_oracle_empty_rows_and_rectangular():
    A = csr_from_coo(4, 3, [0, 2, 2], [1, 0, 2], [2.0, 3.0, 4.0])
    B = csr_from_coo(3, 5, [0, 1, 1], [4, 0, 3], [1.0, 5.0, 6.0])
    r = oracle.spgemm_f64(A, B)
    assert r.shape == (4, 5)
    # row 0: 2*B[1,:] = {0: 10, 3: 12}; row 2: 3*B[0,:] + 4*B[2,:] (empty) = {4: 3}
    assert r.row_ptr.tolist() == [0, 2, 2, 3, 3]
    assert r.col.tolist() == [0, 3, 4] and r.val.tolist() == [10.0, 12.0, 3.0]
    E = csr_from_coo(3, 5, [], [], [])
    r = oracle.spgemm_f64(A, E)
    assert r.row_ptr.tolist() == [0, 0, 0, 0, 0]


def test_oracle_f32_order_matches_python_loop():
    """FP32 accumulation in A-row order, then B-row order (one sequential
    any-hit per product), checked bit-for-bit against a plain loop."""
    rng = np.random.default_rng(9)
    A = rand_csr(rng, 30, 40, 0.3)
    B = rand_csr(rng, 40, 35, 0.3)
    r = oracle.spgemm_f32(A, B)
    for i in range(30):
        acc = {}
        for e in range(A.row_ptr[i], A.row_ptr[i + 1]):
            k = A.col[e]
            for q in range(B.row_ptr[k], B.row_ptr[k + 1]):
                j = int(B.col[q])
                acc[j] = np.float32(acc.get(j, np.float32(0)) + np.float32(A.val[e] * B.val[q]))
        cols = sorted(acc)
        seg = slice(r.row_ptr[i], r.row_ptr[i + 1])
        assert r.col[seg].tolist() == cols
        assert np.array_equal(r.val[seg], np.array([acc[j] for j in cols], np.float32))


def test_oracle_stream_and_rows_agree():
    rng = np.random.default_rng(10)
    A = rand_csr(rng, 300, 200, 0.05)
    B = rand_csr(rng, 200, 250, 0.05)
    r32 = oracle.spgemm_f32(A, B)
    nnz, chk = oracle.stream_f32(A, B, 0, 300)
    j = r32.col.astype(np.float64) + 1
    assert nnz == int(r32.row_ptr[-1])
    assert chk == pytest.approx(float((j * r32.val.astype(np.float64)).sum()), rel=1e-12, abs=1e-9)
    nnz2, _ = oracle.stream_f32(A, B, 100, 200)
    assert nnz2 == int(r32.row_ptr[200] - r32.row_ptr[100])
    r64 = oracle.spgemm_f64(A, B)
    sel = [5, 17, 299, 0, 17]
    rs = oracle.rows_f64(A, B, sel)
    for s, i in enumerate(sel):
        a, b = rs.row_ptr[s], rs.row_ptr[s + 1]
        c, d = r64.row_ptr[i], r64.row_ptr[i + 1]
        assert np.array_equal(rs.col[a:b], r64.col[c:d])
        assert np.array_equal(rs.val[a:b], r64.val[c:d])


def test_oracle_detects_planted_mistakes():
    """A wrong product (transposed operand / dropped last product per row) is
    caught by the brute-force comparison used above."""
    rng = np.random.default_rng(11)
    A = rand_csr(rng, 40, 40, 0.15)
    B = rand_csr(rng, 40, 40, 0.15)
    D, _ = dense_product(A, B)
    wrong = oracle.spgemm_f64(B, A)  # operands swapped
    V = np.zeros((40, 40))
    rows = np.repeat(np.arange(40), np.diff(wrong.row_ptr))
    V[rows, wrong.col] = wrong.val
    assert not np.allclose(V, D, rtol=1e-12, atol=1e-13)
    # drop the last entry of every A row: the dense check must fail too
    keep = np.ones(A.nnz, bool)
    ends = A.row_ptr[1:][np.diff(A.row_ptr) > 0] - 1
    keep[ends] = False
    rows_a = np.repeat(np.arange(40), np.diff(A.row_ptr))
    A2 = csr_from_coo(40, 40, rows_a[keep], A.col[keep], A.val[keep])
    r2 = oracle.spgemm_f64(A2, B)
    V2 = np.zeros((40, 40))
    rows2 = np.repeat(np.arange(40), np.diff(r2.row_ptr))
    V2[rows2, r2.col] = r2.val
    assert not np.allclose(V2, D, rtol=1e-12, atol=1e-13)
