"""Benchmark: SpGEMM GFLOP/s and HBM-roofline fraction on one MI355X.

Workload (one "step"): C = A*A with A the R-MAT matrix of scale 20, edge
factor 16, quadrant probabilities .57/.19/.19/.05, duplicates merged,
self-loops kept, FP32 values U(-1,1), seed 1 -- inputs resident in HBM when
the timed region starts, C complete in HBM (CSR) when it ends.

Work unit: intermediate products P = sum over A(i,k) of nnz(B(k,:));
flops = 2P; value = 2P / t in GFLOP/s.  Algorithmic bytes of the whole
product (A read once, C written once, CSR with int64 row pointers and
4+4-byte entries): 8(m+1) + 8 nnz(A) + 8(m+1) + 8 nnz(C).

Before timing, the library's autotuner picks the row-bin settings for this
matrix class for each method timed (untimed, cached in a temporary tuning
file); both methods then run with their tuned settings.  The operands are
checked once (an untimed call); the timed calls pass them as trusted.

The same line carries, under config.workloads, the other full-size
workloads, each tuned and timed the same way: C = A*A of a 4,096 x 4,096
matrix with 8 uniform columns per row (plain, learned, ray); the Galerkin
triple product R*(A*P) on a 128^3 grid; the rectangular 500,000 x 2,000,000
times 2,000,000 x 500,000 product; and the batch of 2,048 independent
products (batched plain and learned against 2,048 sequential plain calls).

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

    sizes = {"t32": 0, "t128": 1, "t512": 2, "t2048": 3, "t4096": 4, "t8192": 4}
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


def host_threads():
    """Host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (on a shared GPU box it is the CPU share
    of one GPU)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


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


def _timed_product(call, steps, warmup):
    """Warm-up, then `steps` timed calls of call() -> (C, info, free).
    Returns (wall ms per call, median device ms, last info, last C's nnz)."""
    import rtspmspm_amd as rt
    for _ in range(warmup):
        C, info, free = call()
        free()
    rt.synchronize()
    dev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        C, info, free = call()
        dev.append(info["ms"])
        nnz = C.nnz
        free()
    rt.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3, float(np.median(dev)), info, nnz


def _timed_methods(calls, steps, warmup):
    """Times several methods of one product fairly: warm-ups, then `steps`
    rounds that call each method once in turn (so drift in clocks or
    temperature falls on all alike).  calls: name -> call() -> (C, info,
    free).  Returns name -> (median wall ms per call, median device ms,
    last info, last C's nnz)."""
    import rtspmspm_amd as rt
    for call in calls.values():
        for _ in range(warmup):
            C, info, free = call()
            free()
    rt.synchronize()
    wall = {k: [] for k in calls}
    dev = {k: [] for k in calls}
    last = {}
    for _ in range(steps):
        for k, call in calls.items():
            t0 = time.perf_counter()
            C, info, free = call()
            nnz = C.nnz
            free()
            rt.synchronize()
            wall[k].append((time.perf_counter() - t0) * 1e3)
            dev[k].append(info["ms"])
            last[k] = (info, nnz)
    return {k: (float(np.median(wall[k])), float(np.median(dev[k])), last[k][0], last[k][1])
            for k in calls}


def _entry(ms, dev_ms, flops, bytes_alg, info=None):
    e = {"ms": round(ms, 4), "device_ms": round(dev_ms, 4), "gflops": round(flops / ms / 1e6, 3),
         "roofline_frac": round(bytes_alg / ms / 1e6 / HBM_PEAK_GBS, 6)}
    if info is not None and info.get("predicted_rows"):
        e["predicted_rows"] = info["predicted_rows"]
        e["overflow_rows"] = info["overflow_rows"]
    return e


def _speedups(out, base="plain"):
    for k, v in out.items():
        if isinstance(v, dict) and "ms" in v and k != base and base in out:
            v["speedup_vs_" + base] = round(out[base]["ms"] / v["ms"], 4)


def wl_single(name, A, B, desc, steps, warmup, tune, modes=("plain", "learned")):
    """A single product C = A*B: each mode tuned (ray: untuned) and timed."""
    import rtspmspm_amd as rt
    from rtspmspm_amd import workloads as W
    P = int(W.row_products(A, B).sum())
    dA = rt.to_device(A)
    dB = dA if B is A else rt.to_device(B)
    C0, _ = rt.spmspm(dA, dB)  # operands checked once; the timed calls trust them
    nnz = C0.nnz
    C0.free()
    bytes_alg = (W.csr_bytes(A.shape[0], A.nnz) + (0 if B is A else W.csr_bytes(B.shape[0], B.nnz))
                 + W.csr_bytes(A.shape[0], nnz))
    out = {"config": desc, "products": P, "nnz_C": nnz, "bytes_alg": bytes_alg, "tuning": {}}
    calls = {}
    for mode in modes:
        if tune and mode != "ray":
            out["tuning"][mode] = rt.autotune(dA, dB, mode=mode, reps=1).split("; ", 1)[-1]

        def call(mode=mode):
            C, info = rt.spmspm(dA, dB, mode=mode, use_tuned=tune, trusted=True)
            return C, info, C.free
        calls[mode] = call
    for mode, (ms, dev_ms, info, got) in _timed_methods(calls, steps, warmup).items():
        assert got == nnz, f"{name} {mode}: nnz(C) {got} != {nnz}"
        out[mode] = _entry(ms, dev_ms, 2 * P, bytes_alg, info)
        log(f"  {name} {mode}: {ms:.3f} ms ({dev_ms:.3f} device), {2 * P / ms / 1e6:.1f} GFLOP/s")
    _speedups(out)
    rt.pool_trim()
    return out


def wl_galerkin(grid, steps, warmup, tune):
    """R*(A*P) on a grid^3 27-point stencil with 2x2x2 aggregation."""
    import rtspmspm_amd as rt
    from rtspmspm_amd import workloads as W
    A, Pm, R = W.galerkin(grid)
    dA, dP, dR = rt.to_device(A), rt.to_device(Pm), rt.to_device(R)
    AP0, _ = rt.spmspm(dA, dP)
    C0, i2 = rt.spmspm(dR, AP0)
    P1, P2 = int(W.row_products(A, Pm).sum()), i2["products"]
    nnz_ap, nnz_c = AP0.nnz, C0.nnz
    out = {"config": f"R*(A*P), 27-point stencil on {grid}^3, 2x2x2 aggregation",
           "products": P1 + P2, "nnz_AP": nnz_ap, "nnz_C": nnz_c, "tuning": {}}
    out["bytes_alg"] = (W.csr_bytes(A.shape[0], A.nnz) + W.csr_bytes(Pm.shape[0], Pm.nnz)
                        + 2 * W.csr_bytes(A.shape[0], nnz_ap) + W.csr_bytes(R.shape[0], R.nnz)
                        + W.csr_bytes(R.shape[0], nnz_c))
    calls = {}
    for mode in ("plain", "learned"):
        if tune:
            out["tuning"][mode] = [rt.autotune(dA, dP, mode=mode, reps=1).split("; ", 1)[-1],
                                   rt.autotune(dR, AP0, mode=mode, reps=1).split("; ", 1)[-1]]

        def call(mode=mode):
            AP, i1 = rt.spmspm(dA, dP, mode=mode, use_tuned=tune, trusted=True)
            C, i2 = rt.spmspm(dR, AP, mode=mode, use_tuned=tune, trusted=True)
            AP.free()
            info = {"ms": i1["ms"] + i2["ms"],
                    "predicted_rows": i1["predicted_rows"] + i2["predicted_rows"],
                    "overflow_rows": i1["overflow_rows"] + i2["overflow_rows"]}
            return C, info, C.free
        calls[mode] = call
    for mode, (ms, dev_ms, info, got) in _timed_methods(calls, steps, warmup).items():
        assert got == nnz_c
        out[mode] = _entry(ms, dev_ms, 2 * (P1 + P2), out["bytes_alg"], info)
        log(f"  galerkin {mode}: {ms:.3f} ms ({dev_ms:.3f} device)")
    AP0.free()
    C0.free()
    _speedups(out)
    rt.pool_trim()
    return out


def wl_batch(count, steps, warmup, tune):
    """`count` independent products in one batched call (plain and learned)
    against `count` sequential plain calls."""
    import rtspmspm_amd as rt
    from rtspmspm_amd import workloads as W
    t0 = time.perf_counter()
    ar, br, bc, SA, SB = W.batch(count)
    dSA, dSB = rt.to_device(SA), rt.to_device(SB)
    dpairs = [(rt.to_device(a), rt.to_device(b)) for a, b in W.split_pairs(ar, br, bc, SA, SB)]
    log(f"  batch {count}: operands ready in {time.perf_counter() - t0:.1f} s, "
        f"rows {int(ar.sum())}, nnz(A) {SA.nnz}")
    C0, i0 = rt.spmspm_batched(ar, br, bc, dSA, dSB)
    P, nnz_c = i0["products"], C0.nnz
    C0.free()
    bytes_alg = (W.csr_bytes(SA.shape[0], SA.nnz) + W.csr_bytes(SB.shape[0], SB.nnz)
                 + W.csr_bytes(SA.shape[0], nnz_c))
    out = {"config": f"{count} pairs, n log-uniform [256, 65536], Poisson(U(4,32)) per row, "
                     "seed 4", "products": P, "nnz_C": nnz_c, "bytes_alg": bytes_alg, "tuning": {}}
    calls = {}
    for mode in ("plain", "learned"):
        if tune:
            out["tuning"]["batched_" + mode] = rt.autotune_batched(
                ar, br, bc, dSA, dSB, mode=mode, reps=1).split("; ", 1)[-1]

        def call(mode=mode):
            C, info = rt.spmspm_batched(ar, br, bc, dSA, dSB, mode=mode, use_tuned=tune,
                                        trusted=True)
            return C, info, C.free
        calls[mode] = call
    for mode, (ms, dev_ms, info, got) in _timed_methods(calls, steps, warmup).items():
        assert got == nnz_c
        out["batched_" + mode] = _entry(ms, dev_ms, 2 * P, bytes_alg, info)
        log(f"  batched {mode}: {ms:.2f} ms ({dev_ms:.2f} device), {2 * P / ms / 1e6:.1f} GFLOP/s")

    def seq():
        res = [rt.spmspm(a, b, mode="plain", trusted=True)[0] for a, b in dpairs]
        n = sum(c.nnz for c in res)

        class _All:
            nnz = n
        return _All, {"ms": 0.0}, (lambda: [c.free() for c in res])
    ms, _, _, got = _timed_product(seq, max(1, steps // 2), 1)
    assert got == nnz_c, "batched and sequential results differ in size"
    out["sequential_plain"] = _entry(ms, 0.0, 2 * P, bytes_alg)
    del out["sequential_plain"]["device_ms"]
    for k in ("batched_plain", "batched_learned"):
        out[k]["speedup_vs_sequential"] = round(ms / out[k]["ms"], 4)
    out["batched_learned"]["speedup_vs_batched_plain"] = round(
        out["batched_plain"]["ms"] / out["batched_learned"]["ms"], 4)
    log(f"  sequential plain ({count} calls): {ms:.2f} ms -> batched learned "
        f"{out['batched_learned']['speedup_vs_sequential']:.2f}x")
    rt.pool_trim()
    return out


def run_workloads(names, args, tune):
    from rtspmspm_amd import workloads as W
    res = {}
    # Products of a few ms or less take more timed calls (a few microseconds
    # of host noise would otherwise decide their method ratios).
    small = max(args.steps, 20)
    for name in names:
        t0 = time.perf_counter()
        if name == "uniform4k":
            A = W.uniform4k()
            res[name] = wl_single(name, A, A, "A*A, A 4096 x 4096, 8 uniform columns per row, "
                                  "U(-1,1), seed 0", max(args.steps, 50), args.warmup, tune,
                                  ("plain", "learned", "ray"))
        elif name == "galerkin":
            res[name] = wl_galerkin(args.grid, small, args.warmup, tune)
        elif name == "rect":
            A, B = W.rect()
            res[name] = wl_single(name, A, B, "A 500000 x 2000000 lognormal(2.5,1) rows * "
                                  "B 2000000 x 500000 4 per row, N(0,1), seed 3",
                                  small, args.warmup, tune)
        elif name == "batch2048":
            res[name] = wl_batch(args.count, args.steps, args.warmup, tune)
        res[name]["seconds"] = round(time.perf_counter() - t0, 1)
    return res


WORKLOADS = ("uniform4k", "galerkin", "rect", "batch2048")


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
    ap.add_argument("--no-workloads", action="store_true",
                    help="leave the other workloads out of the headline line")
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
    if not args.no_tune:
        import tempfile
        os.environ.setdefault("RTSPMSPM_TUNE_CACHE",
                              os.path.join(tempfile.gettempdir(), f"rtspmspm_bench_{os.getpid()}.txt"))
    if args.workload != "rmat20":
        out = run_workloads([args.workload], args, not args.no_tune)[args.workload]
        print(json.dumps({"workload": args.workload, "steps": args.steps,
                          "warmup": args.warmup, **out}), flush=True)
        return
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
    C0, _ = rt.spmspm(dA, dA)  # operands checked once; the timed calls trust them
    C0.free()
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
            C, info = rt.spmspm(dA, dA, mode=method, use_tuned=tuned, trusted=True)
            C.free()
        rt.synchronize()
        ms_dev = []
        t0 = time.perf_counter()
        for s in range(args.steps):
            C, info = rt.spmspm(dA, dA, mode=method, use_tuned=tuned, trusted=True)
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
        C, _ = rt.spmspm(dA, dA, mode=args.method, profile=True, use_tuned=tuned, trusted=True)
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
        threads = host_threads()
        # 1 thread first: the oracle sets OpenMP's thread count for the
        # process, and the many-thread run restores it.
        one = cpu_baseline(A, P, 1, target_s=6.0)
        cpu = cpu_baseline(A, P, threads, target_s=12.0)
        cpu["nproc"] = os.cpu_count()
        cpu["value_1thread"] = one["value"]
        cpu["sample_1thread"] = one["sample"]
        log(f"cpu baseline: {cpu['value']:.2f} GFLOP/s on {threads} threads of {os.cpu_count()} "
            f"({cpu['sample']}); 1 thread {one['value']:.3f} GFLOP/s ({one['sample']})")

    workloads = None
    if not args.no_workloads:
        del dA
        rt.pool_trim()
        workloads = run_workloads(WORKLOADS, args, tuned)

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
                   "tuning": {k: v.split("; ", 1)[-1] for k, v in tuning.items()},
                   "timed_calls": "operands checked once before timing, then passed as trusted",
                   **compare, "workloads": workloads},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
