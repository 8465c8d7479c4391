"""Benchmark: SpGEMM GFLOP/s and HBM-roofline fraction on one MI355X.

Workload (one "step"): C = A*A with A the R-MAT matrix of scale 20, edge
factor 16, quadrant probabilities .57/.19/.19/.05, duplicates merged,
self-loops kept, FP32 values U(-1,1), seed 1 -- inputs resident in HBM when
the timed region starts, C complete in HBM (CSR) when it ends.

Work unit: intermediate products P = sum over A(i,k) of nnz(B(k,:));
flops = 2P; value = 2P / t in GFLOP/s.  Algorithmic bytes of the whole
product (A read once, C written once, CSR with int64 row pointers and
4+4-byte entries): 8(m+1) + 8 nnz(A) + 8(m+1) + 8 nnz(C).

Before timing, the library's autotuner picks the row-bin thresholds for this
matrix class for each method timed (untimed, cached in a temporary tuning
file); both methods then run with their tuned settings.

Prints ONE JSON line on stdout (human notes go to stderr).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpGEMM GFLOP/s & % HBM roofline: learned+batched vs plain two-phase, 1xMI355X"
HBM_PEAK_GBS = 8000.0
TH = [16, 64, 256, 1024, 4096]  # default row-bin thresholds of the library


def parse_thresholds(report):
    """The numeric-pass thresholds from an autotuner report line."""
    words = report.split("thresholds sym", 1)[1].split()
    return [int(x) for x in words[6:11]]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def row_products(A):
    lens = np.diff(A.row_ptr)
    return np.add.reduceat(np.append(lens[A.col], 0), A.row_ptr[:-1]) * (np.diff(A.row_ptr) > 0)


def kernel_bytes(name, A, P, cnnz, ncols, TH=TH):
    """Algorithmic bytes of one launch of a named kernel: row pointers (16 B
    per row), A entries it must read (col 4 B, +val 4 B for numeric), C
    entries it writes (8 B) or row counts (8 B).  TH: the numeric thresholds
    the product ran with (the symbolic ones are equal in every tuning
    candidate)."""
    anz = np.diff(A.row_ptr)
    U = np.minimum(P, ncols)

    def binmask(key, b):
        lo = 0 if b == 0 else TH[b - 1]
        hi = TH[b] if b < 5 else np.iinfo(np.int64).max
        return (key > lo) & (key <= hi)

    sizes = {"t32": 0, "t128": 1, "t512": 2, "t2048": 3, "t8192": 4}
    if name.startswith("num_hash_") or name.startswith("learn_hash_"):
        # learned tables are binned by predicted capacity (~ nnz); nnz bins here
        m = binmask(cnnz, sizes[name.split("_")[-1]])
        return float(16 * m.sum() + 8 * anz[m].sum() + 8 * cnnz[m].sum())
    if name == "num_heavy":
        m = cnnz > TH[4]
        return float(16 * m.sum() + 8 * anz[m].sum() + 8 * cnnz[m].sum())
    if name.startswith("sym_hash_"):
        m = binmask(U, sizes[name.split("_")[-1]])
        return float(24 * m.sum() + 4 * anz[m].sum())
    if name == "sym_bitmap":
        m = U > TH[4]
        return float(24 * m.sum() + 4 * anz[m].sum())
    return None


def cpu_baseline(A, P, threads, target_s=15.0):
    """FP32 Gustavson (reference semantics) of a strided row sample of the
    same product, on `threads` host cores; returns GFLOP/s and the sample."""
    import oracle
    m = A.shape[0]

    def sample(stride):
        idx = np.arange(0, m, stride)
        lens = np.diff(A.row_ptr)[idx]
        rp = np.zeros(idx.size + 1, np.int64)
        np.cumsum(lens, out=rp[1:])
        starts = A.row_ptr[idx]
        flat = np.repeat(starts - rp[:-1], lens) + np.arange(rp[-1])
        from rtspmspm_amd import CSR
        return CSR((idx.size, A.shape[1]), rp, A.col[flat], A.val[flat]), float(P[idx].sum())

    stride = 4096
    while True:
        S, ps = sample(stride)
        t0 = time.perf_counter()
        oracle.stream_f32(S, A, 0, S.shape[0], threads)
        t = time.perf_counter() - t0
        if t > target_s / 8 or stride == 1:
            break
        stride = max(1, int(stride * max(t, 1e-3) / target_s * 2) if t > 0.05 else stride // 16)
        stride = max(1, stride)
    if t < target_s / 2 and stride > 1:
        stride = max(1, int(stride * t / target_s))
        S, ps = sample(stride)
        t0 = time.perf_counter()
        oracle.stream_f32(S, A, 0, S.shape[0], threads)
        t = time.perf_counter() - t0
    return {"value": 2.0 * ps / t / 1e9, "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": (f"all {S.shape[0]} rows of A (" if stride == 1 else
                       f"1 row in {stride} of A ({S.shape[0]} rows, ") + f"{ps:.3e} products), "
                      f"FP32 Gustavson with sorted CSR rows, computed and discarded, {t:.1f} s"}


def timed_calls(fn, steps, warmup):
    """Run fn() warmup + steps times; returns ms per timed call (wall) and the
    last result.  fn returns (result, free_callable)."""
    import rtspmspm_amd as rt
    for _ in range(warmup):
        r, free = fn()
        free()
    rt.synchronize()
    t0 = time.perf_counter()
    last = None
    for s in range(steps):
        r, free = fn()
        if s + 1 < steps:
            free()
        else:
            last = (r, free)
    rt.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3, last


def run_workload(args):
    """The workloads other than the headline one: one JSON line with plain and
    learned (and ray for uniform4k, sequential-vs-batched for batch2048)."""
    import rtspmspm_amd as rt
    from rtspmspm_amd import workloads as W
    out = {"workload": args.workload, "steps": args.steps, "warmup": args.warmup}
    t0 = time.perf_counter()

    def one(dA, dB, mode):
        def f():
            C, info = rt.spmspm(dA, dB, mode=mode)
            return (C, info), C.free
        return f

    if args.workload in ("uniform4k", "rect"):
        if args.workload == "uniform4k":
            A = W.uniform4k()
            B = A
            desc = "A*A, A 4096 x 4096, 8 uniform columns per row, seed 0"
        else:
            A, B = W.rect()
            desc = "A 500000 x 2000000 lognormal rows * B 2000000 x 500000 4 per row, seed 3"
        P = int(W.row_products(A, B).sum())
        dA, dB = rt.to_device(A), rt.to_device(B)
        log(f"{args.workload}: generated in {time.perf_counter() - t0:.1f} s, products {P}")
        modes = ["plain", "learned"] + (["ray"] if args.workload == "uniform4k" else [])
        for mode in modes:
            ms, ((C, info), free) = timed_calls(one(dA, dB, mode), args.steps, args.warmup)
            nnz = C.nnz
            free()
            out[mode] = {"ms": round(ms, 4), "gflops": round(2 * P / ms / 1e6, 3),
                         "overflow_rows": info["overflow_rows"],
                         "predicted_rows": info["predicted_rows"]}
            log(f"  {mode}: {ms:.3f} ms, {2 * P / ms / 1e6:.2f} GFLOP/s, nnz(C) {nnz}")
        bytes_alg = W.csr_bytes(A.shape[0], A.nnz) + (0 if B is A else W.csr_bytes(B.shape[0], B.nnz)) \
            + W.csr_bytes(A.shape[0], nnz)
        out.update(config=desc, products=P, nnz_C=nnz, bytes_alg=bytes_alg)
    elif args.workload == "galerkin":
        A, Pm, R = W.galerkin(args.grid)
        dA, dP, dR = rt.to_device(A), rt.to_device(Pm), rt.to_device(R)
        P1 = int(W.row_products(A, Pm).sum())
        log(f"galerkin {args.grid}^3: generated in {time.perf_counter() - t0:.1f} s")
        for mode in ("plain", "learned"):
            def chain():
                AP, i1 = rt.spmspm(dA, dP, mode=mode)
                C, i2 = rt.spmspm(dR, AP, mode=mode)
                nnz_ap = AP.nnz
                AP.free()
                return (C, nnz_ap, i1, i2), C.free
            ms, ((C, nnz_ap, i1, i2), free) = timed_calls(chain, args.steps, args.warmup)
            nnz_c = C.nnz
            free()
            P2 = i2["products"]
            out[mode] = {"ms": round(ms, 4), "gflops": round(2 * (P1 + P2) / ms / 1e6, 3),
                         "overflow_rows": i1["overflow_rows"] + i2["overflow_rows"]}
            log(f"  {mode}: {ms:.3f} ms, {2 * (P1 + P2) / ms / 1e6:.2f} GFLOP/s, "
                f"nnz(AP) {nnz_ap}, nnz(C) {nnz_c}")
        bytes_alg = (W.csr_bytes(A.shape[0], A.nnz) + W.csr_bytes(Pm.shape[0], Pm.nnz)
                     + 2 * W.csr_bytes(A.shape[0], nnz_ap) + W.csr_bytes(R.shape[0], R.nnz)
                     + W.csr_bytes(R.shape[0], nnz_c))
        out.update(config=f"R*(A*P), 27-point stencil on {args.grid}^3, 2x2x2 aggregation",
                   products=P1 + P2, nnz_AP=nnz_ap, nnz_C=nnz_c, bytes_alg=bytes_alg)
    elif args.workload == "batch2048":
        ar, br, bc, SA, SB = W.batch(args.count)
        dSA, dSB = rt.to_device(SA), rt.to_device(SB)
        pairs = W.split_pairs(ar, br, bc, SA, SB)
        dpairs = [(rt.to_device(a), rt.to_device(b)) for a, b in pairs]
        log(f"batch {args.count}: generated in {time.perf_counter() - t0:.1f} s, "
            f"rows {int(ar.sum())}, nnz(A) {SA.nnz}")
        for mode in ("plain", "learned"):
            def batched():
                C, info = rt.spmspm_batched(ar, br, bc, dSA, dSB, mode=mode)
                return (C, info), C.free
            ms, ((C, info), free) = timed_calls(batched, args.steps, args.warmup)
            nnz_c, P = C.nnz, info["products"]
            free()
            out["batched_" + mode] = {"ms": round(ms, 3), "gflops": round(2 * P / ms / 1e6, 3),
                                      "overflow_rows": info["overflow_rows"]}
            log(f"  batched {mode}: {ms:.2f} ms, {2 * P / ms / 1e6:.1f} GFLOP/s, nnz(C) {nnz_c}")

        def sequential():
            res = [rt.spmspm(a, b, mode="plain")[0] for a, b in dpairs]
            return res, (lambda: [c.free() for c in res])
        ms_seq, (res, free) = timed_calls(sequential, max(1, args.steps // 2), 1)
        seq_nnz = sum(c.nnz for c in res)
        free()
        assert seq_nnz == nnz_c, "batched and sequential results differ in size"
        out["sequential_plain"] = {"ms": round(ms_seq, 3), "gflops": round(2 * P / ms_seq / 1e6, 3)}
        out["batched_speedup_vs_sequential"] = round(ms_seq / out["batched_learned"]["ms"], 3)
        log(f"  sequential plain ({args.count} launches): {ms_seq:.2f} ms -> batched learned "
            f"{out['batched_speedup_vs_sequential']:.2f}x")
        bytes_alg = (W.csr_bytes(SA.shape[0], SA.nnz) + W.csr_bytes(SB.shape[0], SB.nnz)
                     + W.csr_bytes(SA.shape[0], nnz_c))
        out.update(config=f"{args.count} pairs, n log-uniform [256, 65536], Poisson(U(4,32)), seed 4",
                   products=P, nnz_C=nnz_c, bytes_alg=bytes_alg)
    for k in ("plain", "learned", "ray", "batched_plain", "batched_learned"):
        if k in out:
            out[k]["roofline_frac"] = round(out["bytes_alg"] / out[k]["ms"] / 1e6 / HBM_PEAK_GBS, 6)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="rmat20",
                    choices=["rmat20", "uniform4k", "galerkin", "rect", "batch2048"],
                    help="rmat20 (default) prints the benchmark's contract line; the others "
                         "print one JSON line of per-method timings")
    ap.add_argument("--grid", type=int, default=128, help="galerkin grid size")
    ap.add_argument("--count", type=int, default=2048, help="batch2048 pair count")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--method", default="learned", choices=["plain", "learned", "ray"])
    ap.add_argument("--no-compare", dest="compare", action="store_false",
                    help="skip timing the plain two-phase comparator")
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel event pass")
    ap.add_argument("--no-tune", action="store_true",
                    help="run with the library's default row-bin thresholds")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the library's device_copy kernel over BYTES (counter "
                         "calibration for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import rtspmspm_amd as rt
        for _ in range(3):
            ms = rt.device_copy_probe(args.copy_calibration)
        log(f"device_copy {args.copy_calibration} B: {ms:.3f} ms, "
            f"{2 * args.copy_calibration / ms / 1e6:.1f} GB/s (read + write)")
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU row "
              "partition is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if "RANK" in os.environ and int(os.environ.get("WORLD_SIZE", "1")) > 1:
        sys.exit("bench.py is single-GPU")
    if args.workload != "rmat20":
        run_workload(args)
        return

    if not args.no_tune:
        import tempfile
        os.environ.setdefault("RTSPMSPM_TUNE_CACHE",
                              os.path.join(tempfile.gettempdir(), f"rtspmspm_bench_{os.getpid()}.txt"))
    import rtspmspm_amd as rt
    dev = rt.device_info()
    log(f"device {dev['name']}, {dev['cus']} CUs, {dev['total_mem'] / 2**30:.0f} GiB")
    t0 = time.perf_counter()
    A = rt.gen_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=1)
    P = row_products(A)
    ptot = int(P.sum())
    log(f"R-MAT s{args.scale}: n={A.shape[0]} nnz={A.nnz} products={ptot} "
        f"(gen {time.perf_counter() - t0:.1f} s)")
    dA = rt.to_device(A)
    tuned = not args.no_tune
    tuning = {}
    th_used = TH
    if tuned:
        for method in dict.fromkeys((["plain"] if args.compare and args.method != "plain" else [])
                                    + [args.method]):
            if method == "ray":
                continue
            t1 = time.perf_counter()
            tuning[method] = rt.autotune(dA, dA, mode=method, reps=2)
            log(f"autotune {method} ({time.perf_counter() - t1:.1f} s): {tuning[method]}")
        if args.method in tuning:
            th_used = parse_thresholds(tuning[args.method])

    def timed(method):
        """Warm-up, then K timed steps; returns (wall ms/step, device ms list,
        last C's per-row nnz, last info)."""
        for _ in range(args.warmup):
            C, info = rt.spmspm(dA, dA, mode=method, use_tuned=tuned)
            C.free()
        rt.synchronize()
        ms_dev = []
        t0 = time.perf_counter()
        for s in range(args.steps):
            C, info = rt.spmspm(dA, dA, mode=method, use_tuned=tuned)
            ms_dev.append(info["ms"])
            if s + 1 < args.steps:
                C.free()
        rt.synchronize()
        wall = time.perf_counter() - t0
        rows = np.diff(C.row_ptr_host())
        C.free()
        return wall / args.steps * 1e3, ms_dev, rows, info

    compare = {}
    if args.compare and args.method != "plain":
        p_ms, p_dev, p_rows, _ = timed("plain")
        compare = {"plain_ms_per_step": round(p_ms, 3),
                   "plain_gflops": round(2.0 * ptot / (p_ms / 1e3) / 1e9, 3)}
        log(f"plain: {p_ms:.2f} ms/step, {compare['plain_gflops']:.1f} GFLOP/s")
    ms_per_step, ms_dev, cnnz_row, info = timed(args.method)
    if "plain_ms_per_step" in compare:
        assert np.array_equal(p_rows, cnnz_row), "learned and plain row structure differ"
        compare["speedup_vs_plain"] = round(compare["plain_ms_per_step"] / ms_per_step, 4)
        compare["overflow_rows"] = info["overflow_rows"]
        compare["predicted_rows"] = info["predicted_rows"]
    nnz_c = int(cnnz_row.sum())
    m = A.shape[0]
    bytes_alg = 8 * (m + 1) + 8 * A.nnz + 8 * (m + 1) + 8 * nnz_c
    gflops = 2.0 * ptot / (ms_per_step / 1e3) / 1e9
    log(f"{args.method}: {ms_per_step:.2f} ms/step (device {np.median(ms_dev):.2f} ms median), "
        f"nnz(C)={nnz_c}, {gflops:.1f} GFLOP/s, pipeline {bytes_alg / ms_per_step / 1e6:.1f} GB/s"
        + (f", {compare['speedup_vs_plain']:.3f}x plain, overflow rows "
           f"{info['overflow_rows']}/{info['predicted_rows']}" if compare else ""))

    roofline = None
    kernels = {}
    if not args.no_profile:
        rt.prof_reset()
        C, _ = rt.spmspm(dA, dA, mode=args.method, profile=True, use_tuned=tuned)
        C.free()
        stats = rt.prof_stats()
        kernels = {k: {"launches": v[0], "ms": round(v[1], 4)} for k, v in stats.items()}
        dom = max(stats, key=lambda k: stats[k][1])
        launches, tot_ms = stats[dom]
        avg_s = tot_ms / launches / 1e3
        kb = kernel_bytes(dom, A, P, cnnz_row, A.shape[1], th_used)
        achieved = kb / avg_s / 1e9 if kb else None
        traffic = None
        pmc = os.path.join(ROOT, "profiles", f"{dom}_rmat{args.scale}_traffic.json")
        if os.path.exists(pmc):
            with open(pmc) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                    "traffic": traffic, "alg_bytes_per_launch": kb,
                    "avg_launch_ms": avg_s * 1e3}
        log("kernels: " + ", ".join(f"{k} {v['ms']:.2f} ms/{v['launches']}" for k, v in
                                    sorted(kernels.items(), key=lambda kv: -kv[1]["ms"])))

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(A, P, threads)
        log(f"cpu baseline: {cpu['value']:.2f} GFLOP/s on {threads} threads ({cpu['sample']})")

    out = {
        "metric": METRIC, "value": round(gflops, 3), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"R-MAT scale {args.scale}, edge factor 16, a/b/c/d .57/.19/.19/.05, "
                               "seed 1, C = A*A", "method": args.method, "rows": m,
                   "nnz_A": int(A.nnz), "products": ptot, "nnz_C": nnz_c,
                   "bytes_alg": bytes_alg,
                   "pipeline_gbs": round(bytes_alg / ms_per_step / 1e6, 2),
                   "pipeline_roofline_frac": round(bytes_alg / ms_per_step / 1e6 / HBM_PEAK_GBS, 5),
                   "device_ms_median": round(float(np.median(ms_dev)), 3),
                   "tuned": tuned, "num_thresholds": th_used,
                   "tuning": {k: v.split("; thresholds ")[-1] for k, v in tuning.items()},
                   **compare},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
