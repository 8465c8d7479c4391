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
