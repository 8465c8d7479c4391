"""Experiment driver with the reference harness's semantics.

    python -m rtspmspm_amd.ae_bench [--data DIR] [--npz DIR] [--synthetic] [--runs 20]
                                    [--methods ray,plain,learned] [--ablations]
                                    [--out results/]

The reference's scripts/AE_test.py runs each program 20 times per dataset
(:53-57), records one latency per run from the program's log (:192-231),
averages them (average_runtime_to_csv, :59-70) and reports the speedup of
the ray-traced method over the cuSPARSE baseline (compute_speedup,
:124-142), writing all_result.csv, avg_result.csv and final_result.csv.
Here the "programs" are the library's methods on device-resident operands:
  ray      the faithful one-work-item-per-nonzero launch (the reference's method)
  plain    the exact two-phase method (the comparator, in the cuSPARSE role)
  learned  learned row sizing
and the speedup column is plain_ms / method_ms.  Datasets are the
*_small.mtx files under --data (<DATA>/<name>/<name>_small.mtx, the
reference's layout, AE_test.py:157-182), CSR .npz files under --npz,
and/or the seeded synthetic
workloads (--synthetic), each multiplied as A*A like the harness.
--ablations also times the ray method's ablation modes (the reference's
ARCHSUP/NOMEM/NOINT/NOTHING builds, optixSpMSpM.cu:106-231) for a
per-dataset cost breakdown.  Per-run latencies are the library's hipEvent
times from inputs resident to C complete; the median is reported beside
the mean.
"""
import argparse
import csv
import glob
import json
import os
import statistics
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import rtspmspm_amd as rt  # noqa: E402
from rtspmspm_amd import workloads as W  # noqa: E402

ABLATIONS = ["full", "no_atomic", "traverse", "no_hit", "empty"]


def load_npz(path):
    """A CSR stored as arrays shape/row_ptr/col/val (no pickles)."""
    import numpy as np
    z = np.load(path, allow_pickle=False)
    return rt.CSR(tuple(int(x) for x in z["shape"]), z["row_ptr"], z["col"], z["val"])


def datasets(args):
    out = []
    if args.data:
        for path in sorted(glob.glob(os.path.join(args.data, "*", "*_small.mtx"))):
            out.append((os.path.basename(os.path.dirname(path)), lambda p=path: rt.mtx_read(p)))
    if args.npz:
        for path in sorted(glob.glob(os.path.join(args.npz, "*.npz"))):
            out.append((os.path.splitext(os.path.basename(path))[0], lambda p=path: load_npz(p)))
    if args.synthetic:
        out += [("uniform4k", W.uniform4k), ("rmat14", lambda: W.rmat(14, seed=1)),
                ("rmat16", lambda: W.rmat(16, seed=1)),
                ("stencil32", lambda: W.galerkin(32)[0])]
    return out


def time_method(dA, method, runs, ablation="full"):
    ms = []
    for r in range(runs + 1):
        C, info = rt.spmspm(dA, dA, mode=method, ablation=ablation)
        C.free()
        if r:  # first run is a warm-up
            ms.append(info["ms"])
    return ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", default=None, help="directory of <name>/<name>_small.mtx files")
    ap.add_argument("--npz", default=None,
                    help="directory of CSR .npz files (e.g. tests/golden: the nine SuiteSparse "
                         "matrices of the reference's evaluation)")
    ap.add_argument("--synthetic", action="store_true", help="include the seeded workloads")
    ap.add_argument("--runs", type=int, default=20)
    ap.add_argument("--methods", default="ray,plain,learned")
    ap.add_argument("--ablations", action="store_true")
    ap.add_argument("--out", default="results")
    args = ap.parse_args()
    if not args.data and not args.npz and not args.synthetic:
        args.synthetic = True
    methods = args.methods.split(",")
    os.makedirs(args.out, exist_ok=True)
    rows, summary = [], {}
    for name, load in datasets(args):
        A = load()
        if A.shape[0] != A.shape[1]:
            print(f"{name}: not square, skipped (A*A only, as AE_test.py:216-217)", flush=True)
            continue
        dA = rt.to_device(A)
        res = {}
        for m in methods:
            ms = time_method(dA, m, args.runs)
            rows += [(m, name, "A*A", t) for t in ms]
            res[m] = {"mean_ms": statistics.fmean(ms), "median_ms": statistics.median(ms)}
        if args.ablations:
            res["ablation_ms"] = {a: statistics.median(time_method(dA, "ray", args.runs, a))
                                  for a in ABLATIONS}
        if "plain" in res:
            for m in methods:
                res[m]["speedup_vs_plain"] = res["plain"]["mean_ms"] / res[m]["mean_ms"]
        summary[name] = {"n": A.shape[0], "nnz": A.nnz, **res}
        print(name, json.dumps(res), flush=True)
    with open(os.path.join(args.out, "all_result.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Software", "DataSet", "Scenario", "Runtime(ms)"])
        w.writerows(rows)
    with open(os.path.join(args.out, "avg_result.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Software", "DataSet", "Scenario", "MeanRuntime(ms)", "MedianRuntime(ms)"])
        for name, res in summary.items():
            for m in methods:
                w.writerow([m, name, "A*A", res[m]["mean_ms"], res[m]["median_ms"]])
    with open(os.path.join(args.out, "final_result.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["DataSet"] + [f"{m}_speedup_vs_plain" for m in methods])
        for name, res in summary.items():
            w.writerow([name] + [res[m].get("speedup_vs_plain", "") for m in methods])
    with open(os.path.join(args.out, "summary.json"), "w") as f:
        json.dump(summary, f, indent=1)


if __name__ == "__main__":
    main()
