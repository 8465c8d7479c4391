"""CPU-side tests: the C ABI library loads and exports what the header
declares, generators, Matrix Market I/O, argument errors, the CLI."""
import os
import re
import subprocess

import numpy as np
import pytest

import rtspmspm_amd as rt

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "rt_spmspm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import ctypes
    lib = ctypes.CDLL(rt.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(rt.EXPORTED)


def test_gen_uniform_k_is_exact_and_deterministic():
    A = rt.gen_uniform_k(4096, 8, 0)
    assert A.shape == (4096, 4096) and A.nnz == 4096 * 8
    assert np.all(np.diff(A.row_ptr) == 8)
    for i in range(0, 4096, 97):
        seg = A.col[A.row_ptr[i]:A.row_ptr[i + 1]]
        assert np.all(np.diff(seg) > 0) and seg.min() >= 0 and seg.max() < 4096
    assert np.all(np.abs(A.val) < 1.0)
    B = rt.gen_uniform_k(4096, 8, 0)
    assert np.array_equal(A.col, B.col) and np.array_equal(A.val, B.val)
    C = rt.gen_uniform_k(4096, 8, 1)
    assert not np.array_equal(A.col, C.col)


def test_gen_rmat_properties():
    A = rt.gen_rmat(12, 16, seed=1)
    assert A.shape == (4096, 4096)
    assert 0.8 * 4096 * 16 < A.nnz <= 4096 * 16  # duplicates merged
    for i in range(4096):
        seg = A.col[A.row_ptr[i]:A.row_ptr[i + 1]]
        assert np.all(np.diff(seg) > 0)
    deg = np.diff(A.row_ptr)
    assert deg[:64].mean() > 4 * deg.mean()  # a=0.57 concentrates mass in low rows


def test_gen_stencil_and_aggregation():
    n = 8
    A = rt.gen_stencil27(n, 2)
    assert A.shape == (n ** 3, n ** 3) and A.nnz == (3 * n - 2) ** 3
    rows = np.repeat(np.arange(n ** 3), np.diff(A.row_ptr))
    diag = A.val[A.col == rows]
    assert np.all(diag == 26.0)
    off = A.val[A.col != rows]
    assert np.all(off >= -1.01) and np.all(off <= -0.99)
    P, R = rt.gen_aggregation(n)
    assert P.shape == (n ** 3, (n // 2) ** 3) and P.nnz == n ** 3
    assert R.shape == ((n // 2) ** 3, n ** 3) and R.nnz == n ** 3
    assert np.array_equal(R.to_scipy().toarray(), P.to_scipy().toarray().T)


def test_gen_lognormal_poisson_batch():
    A = rt.gen_lognormal_rows(5000, 20000, 2.5, 1.0, 1, 2000, 3, 0)
    L = np.diff(A.row_ptr)
    assert L.min() >= 1 and L.max() <= 2000 and 8 < np.median(L) < 16  # exp(2.5) ~ 12.2
    B = rt.gen_fixed_normal(20000, 5000, 4, 3, 1)
    assert np.all(np.diff(B.row_ptr) == 4)
    Pm = rt.gen_poisson(3000, 10.0, 4, 0)
    assert abs(np.diff(Pm.row_ptr).mean() - 10.0) < 0.5
    ar, br, bc, SA, SB = rt.gen_batch_poisson(16, 256, 65536, 4.0, 32.0, 4)
    assert np.all(ar == br) and np.all(br == bc)
    assert np.all(ar >= 256) and np.all(ar <= 65536)
    assert SA.shape[0] == ar.sum() and SB.shape[0] == br.sum()


def test_mtx_roundtrip_and_reader(tmp_path):
    A = rt.gen_uniform_k(300, 5, 3)
    p = str(tmp_path / "a.mtx")
    rt.mtx_write(p, A)
    B = rt.mtx_read(p)
    assert np.array_equal(A.row_ptr, B.row_ptr) and np.array_equal(A.col, B.col)
    assert np.array_equal(A.val, B.val)  # 9 significant digits round-trip fp32
    # column-sorted, duplicate-bearing, pattern and symmetric inputs
    q = tmp_path / "b.mtx"
    q.write_text("%%MatrixMarket matrix coordinate real general\n% c\n3 3 4\n"
                 "2 1 1.5\n1 2 2\n1 2 3\n3 3 -1\n")
    M = rt.mtx_read(str(q))
    assert M.row_ptr.tolist() == [0, 1, 2, 3]
    assert M.col.tolist() == [1, 0, 2] and M.val.tolist() == [5.0, 1.5, -1.0]
    q.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2\n2 1\n3 3\n")
    M = rt.mtx_read(str(q))
    assert M.col.tolist() == [1, 0, 2] and M.val.tolist() == [1.0, 1.0, 1.0]
    with pytest.raises(rt.RtError) as e:
        rt.mtx_read(str(tmp_path / "missing.mtx"))
    assert e.value.code == rt.RT_E_IO


def test_mtx_ref_compat_drops_zeros(tmp_path):
    C = rt.CSR((2, 3), np.array([0, 2, 3]), np.array([0, 2, 1], np.int32),
               np.array([0.0, 1.0 / 3.0, 2.5], np.float32))
    p = tmp_path / "c.mtx"
    rt.mtx_write(str(p), C, ref_compat=True)
    lines = p.read_text().splitlines()
    assert lines[1] == "2 3 2"
    assert lines[2:] == ["1 3 0.333333", "2 2 2.5"]
    rt.mtx_write(str(p), C)
    assert p.read_text().splitlines()[1] == "2 3 3"


@pytest.mark.parametrize("name", ["wiki-Vote", "mario002"])
def test_mtx_reader_on_shipped_files_matches_fixture(name):
    path = f"/root/reference/optixSpMSpM/src/data/{name}/{name}_small.mtx"
    if not os.path.exists(path):
        pytest.skip("reference data not mounted")
    from tests.helpers import load_fixture
    A = rt.mtx_read(path)
    F = load_fixture(name)
    assert A.shape == F.shape
    assert np.array_equal(A.row_ptr, F.row_ptr) and np.array_equal(A.col, F.col)
    assert np.array_equal(A.val, F.val)


def test_argument_errors_without_gpu():
    A = rt.CSR((2, 3), np.array([0, 1, 1]), np.array([0], np.int32), np.array([1.0], np.float32))
    with pytest.raises(rt.RtError) as e:
        rt.spmspm(A, A)
    assert e.value.code == rt.RT_E_SHAPE
    with pytest.raises(rt.RtError) as e:
        rt.gen_uniform_k(10, 11, 0)
    assert e.value.code == rt.RT_E_INVALID


def test_cli_usage_exits_1():
    exe = os.path.join(REPO, "rtspmspm_amd", "optixSpMSpM")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 1 and "--mat1" in r.stderr
    r = subprocess.run([exe, "--bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "Unknown option" in r.stderr


def test_mtx_ref_header_reproduces_printresult_overwrite(tmp_path):
    # printResult (optixSpMSpM.cpp:452-494) writes the size line as
    # "rows cols " and, after the entries, rewrites the header in place with
    # the size of a map it never filled: "rows cols 0".  That line is one byte
    # longer, so it eats the first byte of the first entry line.
    head = "%%MatrixMarket matrix coordinate real general\n"
    C = rt.CSR((2, 3), np.array([0, 2, 3]), np.array([0, 2, 1], np.int32),
               np.array([0.0, 1.0 / 3.0, 2.5], np.float32))
    p = tmp_path / "c.mtx"
    rt.mtx_write(str(p), C, ref_compat=2)
    first = head + "2 3 \n" + "1 3 0.333333\n" + "2 2 2.5\n"
    over = head + "2 3 0\n"
    want = over + first[len(over):]
    assert p.read_bytes().decode() == want
    assert want == head + "2 3 0\n 3 0.333333\n2 2 2.5\n"
    # no entries: the rewrite extends the file by its extra byte
    E = rt.CSR((4, 5), np.zeros(5, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32))
    rt.mtx_write(str(p), E, ref_compat=2)
    assert p.read_bytes().decode() == head + "4 5 0\n"
    with pytest.raises(rt.RtError) as e:
        rt.mtx_write(str(p), C, ref_compat=3)
    assert e.value.code == rt.RT_E_INVALID


def test_tune_cache_lines_are_validated(tmp_path):
    good = "v2 r20c20a4b4SP 16 64 256 1024 2048 16 64 256 1024 2048 0 0 1 0 0 1 31"
    good4 = "v2 r20c20a4b4SL 16 64 256 1024 2048 16 64 256 1024 2048 0 0 0 0 0 2 31"
    bad = [
        "v2 k 16 64 256 1024 4096 16 64 256 1024 4096 0 0 0 0 0 2 31",     # 4096 > 4096/2
        "v2 k 16 64 256 1024 2048 16 64 256 1024 2048 0 0 0 0 2 0 31",     # shape 2 in bin 3
        "r20c20a4b4SP 16 64 256 1024 2048 16 64 256 1024 2048 0",        # old format, no version
        "v1 r20c20a4b4SP 16 64 256 1024 2048 16 64 256 1024 2048 0 0 0 0 0 0 31",
        "v2 k 17 64 256 1024 2048 17 64 256 1024 2048 0 0 0 0 0 0 31",     # 17 > 32/2
        "v2 k 16 64 256 1024 8192 16 64 256 1024 8192 0 0 0 0 0 0 31",     # 8192 > 8192/2
        "v2 k 16 64 64 1024 2048 16 64 64 1024 2048 0 0 0 0 0 0 31",       # not increasing
        "v2 k 16 64 256 1024 4096 16 64 256 1024 2048 0 0 0 0 0 0 31",     # sym[4] > num[4]
        "v2 k 0 64 256 1024 2048 0 64 256 1024 2048 0 0 0 0 0 0 31",       # not positive
        "v2 k 16 64 256 1024 2048 16 64 256 1024 2048 2 0 0 0 0 0 31",     # learned_heavy 2
        "v2 k 16 64 256 1024 2048 16 64 256 1024 2048 0 0 3 0 0 0 31",     # shape 3
        "v2 k 16 64 256 1024 2048 16 64 256 1024 2048 0 0 0 0 0 0 40",     # bins mask
        "v2 k 16 64 256 1024 2048 16 64 256 1024 2048 0 0 0 0 0 0",        # field missing
        "v2 k 16 64 256 1024 2048 16 64 256 1024 2048 0 0 0 0 0 0 31 7",   # extra field
        "v2 k sixteen 64 256 1024 2048 16 64 256 1024 2048 0 0 0 0 0 0 31",
    ]
    p = tmp_path / "tune.txt"
    p.write_text("# comment\n\n" + "\n".join([good, good4] + bad) + "\n")
    assert rt.tune_cache_check(str(p)) == (2, len(bad))
    assert rt.tune_cache_check(str(tmp_path / "absent.txt")) == (0, 0)


@pytest.mark.parametrize("case", ["col_range", "unsorted", "rowptr_start", "rowptr_order"])
def test_row_features_rejects_malformed_operands_on_host(case):
    # The checks run before any upload or launch, so no GPU is needed.
    rp = np.array([0, 2, 3], np.int64)
    col = np.array([0, 2, 1], np.int32)
    if case == "col_range":
        col = np.array([0, 3, 1], np.int32)
    elif case == "unsorted":
        col = np.array([2, 0, 1], np.int32)
    elif case == "rowptr_start":
        rp = np.array([1, 2, 3], np.int64)
    else:
        rp = np.array([0, 3, 2], np.int64)
    bad = rt.CSR((2, 3), rp, col, np.ones(3, np.float32))
    good = rt.CSR((3, 3), np.array([0, 1, 2, 3]), np.array([0, 1, 2], np.int32),
                  np.ones(3, np.float32))
    with pytest.raises(rt.RtError) as e:
        rt.row_features(bad, good)
    assert e.value.code == rt.RT_E_INVALID


def test_plan_create_rejects_null_row_pointer():
    import ctypes
    lib = ctypes.CDLL(rt.LIB_PATH)
    A = rt.rt_csr(4, 4, 0, None, None, None, rt.MEM_DEVICE, 0)
    plan = ctypes.c_void_p()
    C = rt.rt_csr()
    rc = lib.rt_plan_create(ctypes.byref(A), ctypes.byref(A), ctypes.byref(plan),
                            ctypes.byref(C), None)
    assert rc == rt.RT_E_INVALID and not plan.value
    assert b"row_ptr is NULL" in rt._L.rt_last_error()


@pytest.mark.parametrize("target,exe", [("rtspmspm_amd/csrc", "asan_host"),
                                        ("oracle", "oracle_asan")])
def test_host_code_under_address_sanitizer(tmp_path, target, exe):
    """The CPU-side code (generators, Matrix Market I/O, argument checks,
    tuning-cache parser; and the oracle) built with AddressSanitizer and
    UndefinedBehaviorSanitizer and driven through its cases."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, target), "asan"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([os.path.join(REPO, "_build", "asan", exe), str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "checks: ok" in r.stdout
