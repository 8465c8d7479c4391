"""GPU parity of the HIP SpGEMM paths against the CPU oracle and the golden
fixtures.  Structure must match bit-exactly; values must be within
1e-5 * sum|a*b| of the FP64 oracle, entry by entry (REL_TOL in helpers)."""
import numpy as np
import pytest

import oracle
import rtspmspm_amd as rt
from tests.helpers import (FIXTURES, check_against_golden, check_parity, csr_from_coo, golden,
                           load_fixture)

pytestmark = pytest.mark.gpu


def run(A, B, **kw):
    C, info = rt.spmspm(A, B, **kw)
    return C, info


# ---- 4,096 x 4,096, 8 uniform columns per row, seed 0 ----------
@pytest.fixture(scope="module")
def cfg1():
    A = rt.gen_uniform_k(4096, 8, 0)
    return A, oracle.spgemm_f64(A, A)


def test_cfg1_plain(cfg1):
    A, ref = cfg1
    C, info = run(A, A, mode="plain")
    assert info["products"] == 4096 * 8 * 8
    err = check_parity(C, ref)
    print(f"cfg1 plain: nnz {C.nnz}, max err/sum|ab| {err:.2e}")


@pytest.mark.parametrize("out", ["dense", "slot"])
def test_cfg1_ray(cfg1, out):
    A, ref = cfg1
    C, _ = run(A, A, mode="ray", ray_output=out)
    err = check_parity(C, ref)
    print(f"cfg1 ray/{out}: max err/sum|ab| {err:.2e}")


@pytest.mark.parametrize("abl", ["no_atomic", "traverse", "no_hit", "empty"])
def test_ray_ablations_run(cfg1, abl):
    A, _ = cfg1
    C, info = run(A, A, mode="ray", ablation=abl, ray_output="dense")
    assert C.shape == (4096, 4096) and info["ms"] > 0


# ---- the nine shipped fixtures, A*A ---------------------------------------
@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_plain(name):
    A = load_fixture(name)
    C, _ = run(A, A, mode="plain")
    check_against_golden(C, golden()[name])
    check_parity(C, oracle.spgemm_f64(A, A))
    if golden()[name]["all_ones"] and golden()[name]["max_row_nnz"] < 2**24:
        assert np.array_equal(C.val, oracle.spgemm_f32(A, A).val)  # integer sums: exact


@pytest.mark.parametrize("name", ["wiki-Vote", "mario002", "web-Google", "ca-CondMat"])
def test_fixture_ray(name):
    A = load_fixture(name)
    C, _ = run(A, A, mode="ray")
    check_against_golden(C, golden()[name])
    check_parity(C, oracle.spgemm_f64(A, A))


# ---- structure edge cases ---------------------------------------------------
def test_cancellation_zero_is_kept():
    A = csr_from_coo(1, 2, [0, 0], [0, 1], [1.0, 1.0])
    B = csr_from_coo(2, 1, [0, 1], [0, 0], [0.75, -0.75])
    for mode in ("plain", "ray"):
        C, _ = run(A, B, mode=mode)
        assert C.row_ptr.tolist() == [0, 1] and C.col.tolist() == [0] and C.val.tolist() == [0.0]


def test_empty_and_rectangular():
    A = csr_from_coo(4, 3, [0, 2, 2], [1, 0, 2], [2.0, 3.0, 4.0])
    B = csr_from_coo(3, 5, [0, 1, 1], [4, 0, 3], [1.0, 5.0, 6.0])
    for mode in ("plain", "ray"):
        C, _ = run(A, B, mode=mode)
        check_parity(C, oracle.spgemm_f64(A, B))
    E = csr_from_coo(3, 5, [], [], [])
    C, _ = run(A, E)
    assert C.nnz == 0 and C.row_ptr.tolist() == [0] * 5
    Z = csr_from_coo(0, 3, [], [], [])
    C, _ = run(Z, csr_from_coo(3, 2, [0], [1], [1.0]))
    assert C.shape == (0, 2) and C.nnz == 0
    one = csr_from_coo(1, 1, [0], [0], [3.0])
    C, _ = run(one, one)
    assert C.val.tolist() == [9.0]


def test_shape_mismatch_is_an_error():
    A = csr_from_coo(2, 3, [0], [0], [1.0])
    with pytest.raises(rt.RtError) as e:
        run(A, A)
    assert e.value.code == rt.RT_E_SHAPE


# ---- every bin, including the LDS-bitmap and dense-tile kernels --------------
@pytest.mark.parametrize("n,k,seed", [(3000, 40, 1), (20000, 100, 2), (60000, 160, 3)])
def test_long_rows_uniform(n, k, seed):
    """k^2 products per row: 1,600 (hash bins) up to 25,600 (bitmap + tiles)."""
    A = rt.gen_uniform_k(n, k, seed)
    C, info = run(A, A)
    assert info["products"] == n * k * k
    check_parity(C, oracle.spgemm_f64(A, A))


def test_long_rows_dense_and_compressed_steps():
    """Rows of ~39K entries over 40K columns: the two full 16384-column tiles
    hold more than 7168 entries each (dense steps), the last one is a
    compressed step."""
    A = rt.gen_fixed_normal(200, 40000, 400, 41)
    B = rt.gen_uniform_k(40000, 400, 42)
    C, _ = run(A, B)
    assert np.diff(C.row_ptr).min() > 32768  # tile 0 of every row is dense
    check_parity(C, oracle.spgemm_f64(A, B))


def test_long_rows_sparse_over_many_tiles():
    """~5K entries spread over 62 tiles: compressed steps cut by the LDS
    budget (accumulators, staged bitmap words and group table) before the
    64-tile span, each staged in two rounds of 32 tiles."""
    A = rt.gen_fixed_normal(100, 5000, 250, 43)
    B = rt.gen_fixed_normal(5000, 1_000_000, 20, 44)
    C, info = run(A, B)
    assert info["products"] == 100 * 250 * 20
    check_parity(C, oracle.spgemm_f64(A, B))


def test_mixed_row_lengths_rmat():
    A = rt.gen_rmat(14, 16, seed=5)
    C, _ = run(A, A)
    err = check_parity(C, oracle.spgemm_f64(A, A))
    print(f"rmat14: nnz {C.nnz}, max err/sum|ab| {err:.2e}")


def test_wide_output_multiple_bitmap_windows():
    """B with 1.5M columns: the symbolic bitmap takes three 2^19-column windows."""
    A = rt.gen_fixed_normal(64, 3000, 300, 6)
    B = rt.gen_fixed_normal(3000, 1_500_000, 40, 7)
    C, _ = run(A, B)
    check_parity(C, oracle.spgemm_f64(A, B))


def test_long_rows_window_span_limit():
    """~4.8K entries over 92 tiles: a compressed step ends at its 64-tile span
    before its entry or word bounds."""
    A = rt.gen_fixed_normal(64, 3000, 120, 45)
    B = rt.gen_fixed_normal(3000, 1_500_000, 40, 46)
    C, _ = run(A, B)
    assert np.diff(C.row_ptr).min() > 4096  # every row takes the long-row kernels
    check_parity(C, oracle.spgemm_f64(A, B))


def _mixed_run_operands(seed):
    """B rows of 1, 15-17, 63-65, 127-129, 500 and 2000 entries over 300,000
    columns; A rows of 3, 8, 30 or 200 entries picking B rows at random."""
    rng = np.random.default_rng(seed)
    ncols, nb, na = 300_000, 600, 400
    lens = [1, 15, 16, 17, 63, 64, 65, 127, 128, 129, 500, 2000]
    br, bc = [], []
    for k in range(nb):
        cols = np.sort(rng.choice(ncols, lens[k % len(lens)], replace=False))
        br.append(np.full(cols.size, k))
        bc.append(cols)
    br, bc = np.concatenate(br), np.concatenate(bc)
    B = csr_from_coo(nb, ncols, br, bc, rng.uniform(-1, 1, br.size))
    ar, ac = [], []
    for i in range(na):
        ks = rng.choice(nb, [3, 8, 30, 200][i % 4], replace=False)
        ar.append(np.full(ks.size, i))
        ac.append(ks)
    ar, ac = np.concatenate(ar), np.concatenate(ac)
    A = csr_from_coo(na, nb, ar, ac, rng.uniform(-1, 1, ar.size))
    return A, B


@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_chunk_walk_block_and_mark_modes(mode):
    """Runs that fill 64-product blocks exactly, nearly, or leave them mostly
    idle: chunks fall on both sides of the blocks-or-owner-marks choice, in
    the hash kernels (short A rows) and in the long-row kernels (A rows of
    200 entries, ~50K products over 300,000 columns)."""
    A, B = _mixed_run_operands(11)
    C, info = run(A, B, mode=mode)
    nnz = np.diff(C.row_ptr)
    assert nnz.max() > 4096 and nnz.min() < 256  # long rows and hash rows both present
    check_parity(C, oracle.spgemm_f64(A, B))


@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_long_row_order_with_hub_rows(mode):
    """B of 4.8M entries (38 MB, more than the L2s hold) with ten hub rows of
    5,000 entries; 1,200 long rows, a third of them gathering hubs: the
    long rows are reordered by the hubs they gather before both long-row
    kernels run."""
    rng = np.random.default_rng(21)
    nb, ncols = 120_000, 1_000_000
    lens = np.full(nb, 40)
    lens[:10] = 5000
    br = np.repeat(np.arange(nb), lens)
    # Short rows: strictly increasing columns from gaps in [1, ncols/40).
    gaps = rng.integers(1, ncols // 40, size=(nb - 10, 40))
    bc = np.concatenate([rng.choice(ncols, n, replace=False) for n in lens[:10]] +
                        [np.cumsum(gaps, axis=1).ravel()])
    B = csr_from_coo(nb, ncols, br, bc, rng.uniform(-1, 1, br.size))
    assert B.nnz * 8 > 32 << 20
    na = 1200
    ar, ac = [], []
    for i in range(na):
        ks = rng.choice(np.arange(10, nb), 110, replace=False)
        if i % 3 == 0:
            ks[:3] = rng.choice(10, 3, replace=False)
        ar.append(np.full(ks.size, i))
        ac.append(ks)
    ar, ac = np.concatenate(ar), np.concatenate(ac)
    A = csr_from_coo(na, nb, ar, ac, rng.uniform(-1, 1, ar.size))
    C, _ = run(A, B, mode=mode)
    assert np.diff(C.row_ptr).min() > 4096  # every row is a long row
    check_parity(C, oracle.spgemm_f64(A, B))


def _span_operands(seed):
    """B rows of N entries, N from 5 to 4,000 (rows of every hash bin), whose
    columns lie in a span of about 2N columns, 64N, 256N or the whole 2^22:
    the hash kernels order such rows with one bitmap window, with several,
    or with the sort.  Families of four B rows share a span; A rows take one
    of them or two (overlapping columns: merged products)."""
    rng = np.random.default_rng(seed)
    ncols = 1 << 22
    br, bc, ar, ac = [], [], [], []
    nb = na = 0
    for n in [5, 12, 40, 100, 200, 600, 1000, 2000, 4000]:
        for span in [2 * n + 8, 64 * n, 256 * n, ncols]:
            span = min(span, ncols)
            base = int(rng.integers(0, ncols - span + 1))
            for k in range(4):
                cols = base + np.sort(rng.choice(span, n, replace=False))
                br.append(np.full(n, nb + k))
                bc.append(cols)
            for pick in ([0], [1], [0, 1], [2, 3]):
                ar.append(np.full(len(pick), na))
                ac.append(nb + np.array(pick))
                na += 1
            nb += 4
    br, bc = np.concatenate(br), np.concatenate(bc)
    B = csr_from_coo(nb, ncols, br, bc, rng.uniform(-1, 1, br.size))
    ar, ac = np.concatenate(ar), np.concatenate(ac)
    A = csr_from_coo(na, nb, ar, ac, rng.uniform(-1, 1, ar.size))
    return A, B


@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_hash_rows_over_narrow_windowed_and_wide_spans(mode):
    A, B = _span_operands(31)
    C, _ = run(A, B, mode=mode)
    check_parity(C, oracle.spgemm_f64(A, B))
    Cg, _ = run(A, B, mode="plain", graph=True)
    assert np.array_equal(Cg.row_ptr, C.row_ptr) and np.array_equal(Cg.col, C.col)


def _radix_operands(seed, ncols):
    """B: 2,000 rows of 32 entries at uniform columns over ncols; A: 600 rows
    of 10-60 entries into B, so C's rows hold about 320-1,920 entries spread
    over the whole column range."""
    rng = np.random.default_rng(seed)
    nb, per = 2000, 32
    br = np.repeat(np.arange(nb), per)
    bc = np.concatenate([rng.choice(ncols, per, replace=False) for _ in range(nb)])
    B = csr_from_coo(nb, ncols, br, bc, rng.uniform(-1, 1, br.size))
    lens = rng.integers(10, 61, 600)
    ar = np.repeat(np.arange(lens.size), lens)
    ac = np.concatenate([rng.choice(nb, n, replace=False) for n in lens])
    A = csr_from_coo(lens.size, nb, ar, ac, rng.uniform(-1, 1, ar.size))
    return A, B


@pytest.mark.parametrize("ncols", [1 << 22, (1 << 24) - 3])
@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_hash_rows_radix_sorted_over_wide_spans(ncols, mode):
    """Rows of the 2048- and 8192-slot hash bins whose columns span 2^22-2^24
    columns, far more than 32 x their tables: they are ordered by the radix
    sort with four 7-bit digit passes (R-MAT's spans of 2^20 take three)."""
    A, B = _radix_operands(81, ncols)
    ref = oracle.spgemm_f64(A, B)
    nnz = np.diff(ref.row_ptr)
    assert ((nnz > 256) & (nnz <= 1024)).sum() > 100 and ((nnz > 1024) & (nnz <= 2048)).sum() > 100
    C, _ = run(A, B, mode=mode)
    check_parity(C, ref)


@pytest.mark.parametrize("ncols", [16384, 65536, 65537, 262144])
@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_bitmap_rows_at_column_limits(ncols, mode):
    """Symbolic hash rows count their columns in a bitmap of the table's
    32 x T bits when every B column fits it (65,536 columns for the 2048-slot
    bin, 262,144 for the 8192-slot one), and exact numeric rows of the
    2048-slot bin rank their columns through such a bitmap and sum in rank
    order; otherwise both hash.  B's row 0, gathered by every A row, holds
    columns 0 and ncols - 1, so every row reaches both ends of the bitmap (or
    just misses the limit, at 65,537)."""
    rng = np.random.default_rng(ncols)
    nb, per = 2000, 32
    br = np.repeat(np.arange(nb), per)
    cols = [np.concatenate([[0, ncols - 1], rng.choice(np.arange(1, ncols - 1), per - 2,
                                                       replace=False)])]
    cols += [rng.choice(ncols, per, replace=False) for _ in range(nb - 1)]
    B = csr_from_coo(nb, ncols, br, np.concatenate(cols), rng.uniform(-1, 1, br.size))
    lens = rng.integers(10, 61, 600)
    ar = np.repeat(np.arange(lens.size), lens)
    ac = np.concatenate([np.concatenate([[0], rng.choice(np.arange(1, nb), n - 1, replace=False)])
                         for n in lens])
    A = csr_from_coo(lens.size, nb, ar, ac, rng.uniform(-1, 1, ar.size))
    ref = oracle.spgemm_f64(A, B)
    nnz = np.diff(ref.row_ptr)
    assert ((nnz > 256) & (nnz <= 1024)).sum() > 100 and ((nnz > 1024) & (nnz <= 2048)).sum() > 100
    assert (ref.col[ref.row_ptr[:-1]] == 0).all() and (ref.col[ref.row_ptr[1:] - 1] == ncols - 1).all()
    C, _ = run(A, B, mode=mode)
    check_parity(C, ref)


def test_structure_is_deterministic():
    A = rt.gen_rmat(13, 16, seed=8)
    C1, _ = run(A, A)
    C2, _ = run(A, A)
    assert np.array_equal(C1.row_ptr, C2.row_ptr) and np.array_equal(C1.col, C2.col)


def test_device_resident_operands():
    A = rt.gen_uniform_k(5000, 12, 9)
    dA = rt.to_device(A)
    dC, info = rt.spmspm(dA, dA)
    assert isinstance(dC, rt.DeviceCSR)
    check_parity(dC.to_host(), oracle.spgemm_f64(A, A))


@pytest.mark.parametrize("make", [lambda: rt.gen_rmat(13, 16, seed=51),
                                  lambda: rt.gen_uniform_k(20000, 100, 52)])
def test_graph_captured_stages_equal_direct(make):
    """Plain method with both stages captured in hipGraphs (persistent row
    kernels, device-side binning): identical structure, parity values."""
    A = make()
    Cg, _ = run(A, A, mode="plain", graph=True)
    Cd, _ = run(A, A, mode="plain", graph=False)
    assert np.array_equal(Cg.row_ptr, Cd.row_ptr) and np.array_equal(Cg.col, Cd.col)
    check_parity(Cg, oracle.spgemm_f64(A, A))


def test_batched_graph_and_direct_agree():
    ar, br, bc, A, B = rt.gen_batch_poisson(32, 64, 5000, 4.0, 32.0, 7)
    dA, dB = rt.to_device(A), rt.to_device(B)
    Cg = rt.spmspm_batched(ar, br, bc, dA, dB, graph=True)[0].to_host()
    Cd = rt.spmspm_batched(ar, br, bc, dA, dB, graph=False)[0].to_host()
    assert np.array_equal(Cg.row_ptr, Cd.row_ptr) and np.array_equal(Cg.col, Cd.col)


def test_batched_matches_each_pair():
    ar, br, bc, A, B = rt.gen_batch_poisson(24, 16, 3000, 4.0, 32.0, 4)
    dA, dB = rt.to_device(A), rt.to_device(B)
    dC, info = rt.spmspm_batched(ar, br, bc, dA, dB)
    C = dC.to_host()
    a0 = b0 = 0
    for p in range(len(ar)):
        Ap = rt.CSR((int(ar[p]), int(br[p])), A.row_ptr[a0:a0 + ar[p] + 1] - A.row_ptr[a0],
                    A.col[A.row_ptr[a0]:A.row_ptr[a0 + ar[p]]], A.val[A.row_ptr[a0]:A.row_ptr[a0 + ar[p]]])
        Bp = rt.CSR((int(br[p]), int(bc[p])), B.row_ptr[b0:b0 + br[p] + 1] - B.row_ptr[b0],
                    B.col[B.row_ptr[b0]:B.row_ptr[b0 + br[p]]], B.val[B.row_ptr[b0]:B.row_ptr[b0 + br[p]]])
        s, e = C.row_ptr[a0], C.row_ptr[a0 + ar[p]]
        Cp = rt.CSR((int(ar[p]), int(bc[p])), C.row_ptr[a0:a0 + ar[p] + 1] - s, C.col[s:e], C.val[s:e])
        check_parity(Cp, oracle.spgemm_f64(Ap, Bp))
        a0 += int(ar[p])
        b0 += int(br[p])
