"""Host-side pieces of bench.py (no GPU): the autotuner report parser and the
single-GPU guard."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_parse_thresholds_takes_the_numeric_cut():
    line = ("class r20c20a4b4SP best 114.5 ms of 10 candidates; ms per candidate: default:120.4, "
            "cut2048:114.7; bins (slots/max nnz/lanes): 32/16/8 128/64/16 512/256/64 2048/512/128 "
            "8192/1024/512; thresholds sym 16 64 256 512 1024 num 16 64 256 512 1024 "
            "learned_heavy 0 shapes 0 0 0 0 0 learned_bins 31")
    assert bench.parse_thresholds(line) == [16, 64, 256, 512, 1024]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
