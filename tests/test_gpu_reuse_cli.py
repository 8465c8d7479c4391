"""GPU tests: numeric-only re-run on a fixed structure (rt_plan_*) and the
optixSpMSpM command-line drop-in."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import rtspmspm_amd as rt
from tests.helpers import check_against_golden, check_parity, golden, load_fixture

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "rtspmspm_amd", "optixSpMSpM")


def with_values(M, vals):
    return rt.CSR(M.shape, M.row_ptr, M.col, np.asarray(vals, np.float32))


@pytest.mark.parametrize("make", [lambda: rt.gen_rmat(13, 16, seed=31),
                                  lambda: rt.gen_uniform_k(30000, 90, 32)])
def test_plan_reruns_numeric_with_new_values(make):
    A = make()
    dA = rt.to_device(A)
    plan = rt.Plan(dA, dA)
    check_parity(plan.C.to_host(), oracle.spgemm_f64(A, A))
    rng = np.random.default_rng(5)
    for _ in range(2):
        A2 = with_values(A, rng.uniform(-2, 2, A.nnz))
        B2 = with_values(A, rng.uniform(-2, 2, A.nnz))
        d2, e2 = rt.to_device(A2), rt.to_device(B2)
        ms = plan.execute(d2, e2)
        assert ms > 0
        check_parity(plan.C.to_host(), oracle.spgemm_f64(A2, B2))
    plan.close()


def test_plan_rejects_foreign_result():
    A = rt.gen_uniform_k(1000, 4, 33)
    dA = rt.to_device(A)
    plan = rt.Plan(dA, dA)
    other, _ = rt.spmspm(dA, dA)
    with pytest.raises(rt.RtError):
        rt._check(rt._L.rt_plan_execute(plan._p, dA._s.val, dA._s.val,
                                        rt.ctypes.byref(other._s), None))
    plan.close()


@pytest.mark.parametrize("mode,banner", [("ray", "ATOMIC"), ("plain", "PLAIN")])
def test_cli_matches_reference_contract(tmp_path, mode, banner):
    name = "wiki-Vote"
    A = load_fixture(name)
    a = tmp_path / "a.mtx"
    rt.mtx_write(str(a), A)
    out = tmp_path / "c.mtx"
    log = tmp_path / "log.txt"
    r = subprocess.run([CLI, "-m1", str(a), "-m2", str(a), "-o", str(out), "-l", str(log),
                        "--mode", mode], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith(banner) and r.stdout.rstrip().endswith("success!")
    ms = float(log.read_text())
    assert ms > 0 and "\n" not in log.read_text()
    C = rt.mtx_read(str(out))
    check_against_golden(C, golden()[name])


def test_cli_ref_compat_and_errors(tmp_path):
    A = load_fixture("mario002")
    a = tmp_path / "a.mtx"
    rt.mtx_write(str(a), A)
    out = tmp_path / "c.mtx"
    r = subprocess.run([CLI, "-m1", str(a), "-m2", str(a), "-o", str(out), "--ref-compat"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    C = rt.mtx_read(str(out))
    ref = oracle.spgemm_f64(A, A)
    assert C.nnz == int(np.count_nonzero(ref.val.astype(np.float32)))  # exact zeros dropped
    r = subprocess.run([CLI, "-m1", str(tmp_path / "missing.mtx"), "-m2", str(a)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Caught exception" in r.stderr
    rect = tmp_path / "r.mtx"
    rt.mtx_write(str(rect), rt.gen_fixed_normal(10, 20, 3, 1))
    r = subprocess.run([CLI, "-m1", str(rect), "-m2", str(rect)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "shape mismatch" in r.stderr


def test_autotune_records_and_is_used(tmp_path, monkeypatch):
    """One class each for single and batched calls: the report names the
    chosen table sizes, thresholds and lanes per row, no candidate's result
    differed from the first's, and the cached choice runs at parity."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("RTSPMSPM_TUNE_CACHE", str(cache))
    A = rt.gen_rmat(13, 16, seed=61)
    dA = rt.to_device(A)
    ref = oracle.spgemm_f64(A, A)
    keys = []
    for mode in ("plain", "learned"):
        report = rt.autotune(dA, dA, mode=mode, reps=1)
        print(report)
        key = report.split()[1]
        assert report.startswith("class ") and key.endswith("S" + mode[0].upper())
        assert "bins (slots/max nnz/lanes): 32/" in report and "rejected" not in report
        keys.append(key)
        C, _ = rt.spmspm(A, A, mode=mode, use_tuned=True)
        check_parity(C, ref)
    ar, br, bc, SA, SB = rt.gen_batch_poisson(24, 64, 4000, 4.0, 32.0, 62)
    dSA, dSB = rt.to_device(SA), rt.to_device(SB)
    report = rt.autotune_batched(ar, br, bc, dSA, dSB, mode="learned", reps=1)
    print(report)
    key = report.split()[1]
    assert key.endswith("BL") and "rejected" not in report
    keys.append(key)
    Ct = rt.spmspm_batched(ar, br, bc, dSA, dSB, mode="learned", use_tuned=True)[0].to_host()
    Cp = rt.spmspm_batched(ar, br, bc, dSA, dSB, mode="plain")[0].to_host()
    assert np.array_equal(Ct.row_ptr, Cp.row_ptr) and np.array_equal(Ct.col, Cp.col)
    lines = [x for x in cache.read_text().split("\n") if x.strip()]
    assert sorted(x.split()[1] for x in lines) == sorted(keys)
    assert all(x.startswith("v2 ") for x in lines)
    assert rt.tune_cache_check(str(cache)) == (3, 0)


def test_bin4_half_size_table_at_parity(tmp_path, monkeypatch):
    """Bin 4's 4096-slot table (shape 2, 256 lanes per row, long-row cut
    2048), forced through a tuning-cache line for both methods: the hash rows
    of 1024-2048 entries (symbolic and numeric) and the long rows beside
    them match the oracle."""
    A = rt.gen_rmat(13, 16, seed=63)
    ref = oracle.spgemm_f64(A, A)
    nnz = np.diff(ref.row_ptr)
    assert ((nnz > 1024) & (nnz <= 2048)).sum() > 1000 and (nnz > 2048).sum() > 1000
    dA = rt.to_device(A)
    monkeypatch.setenv("RTSPMSPM_TUNE_CACHE", str(tmp_path / "probe.txt"))
    key = rt.autotune(dA, dA, mode="plain", reps=1).split()[1]
    forced = tmp_path / "forced.txt"
    forced.write_text("".join(
        f"v2 {key[:-1]}{m} 16 64 256 1024 2048 16 64 256 1024 2048 0 0 0 0 0 2 31\n"
        for m in "PL"))
    assert rt.tune_cache_check(str(forced)) == (2, 0)
    monkeypatch.setenv("RTSPMSPM_TUNE_CACHE", str(forced))
    for mode in ("plain", "learned"):
        C, _ = rt.spmspm(A, A, mode=mode, use_tuned=True)
        check_parity(C, ref)


def test_ae_driver_writes_harness_tables(tmp_path):
    """The experiment driver: per-run latencies, mean/median table, speedup
    table and JSON summary, with the ray ablation breakdown."""
    import csv
    import json
    import shutil
    data = tmp_path / "npz"
    data.mkdir()
    shutil.copy(os.path.join(REPO, "tests", "golden", "wiki-Vote.npz"), data)
    out = tmp_path / "out"
    r = subprocess.run(["python", "-m", "rtspmspm_amd.ae_bench", "--npz", str(data), "--runs", "3",
                        "--methods", "ray,plain", "--ablations", "--out", str(out)],
                       cwd=REPO, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = list(csv.DictReader(open(out / "all_result.csv")))
    assert len(rows) == 6 and {x["Software"] for x in rows} == {"ray", "plain"}
    assert all(float(x["Runtime(ms)"]) > 0 for x in rows)
    avg = list(csv.DictReader(open(out / "avg_result.csv")))
    assert len(avg) == 2
    fin = list(csv.DictReader(open(out / "final_result.csv")))
    assert fin[0]["DataSet"] == "wiki-Vote" and float(fin[0]["plain_speedup_vs_plain"]) == 1.0
    s = json.load(open(out / "summary.json"))["wiki-Vote"]
    assert set(s["ablation_ms"]) == {"full", "no_atomic", "traverse", "no_hit", "empty"}
