"""Train the row-nnz predictor used by the learned method and store it.

    python -m rtspmspm_amd.train_predictor [--out rtspmspm_amd/weights]

Training pairs (features, log2(1 + exact row nnz)) come from the exact
symbolic pass on the seeded synthetic matrices of the first two workloads:
  * A*A, A 4,096 x 4,096 with 8 uniform columns per row, seed 0;
  * A*A, A R-MAT scale 20, edge factor 16, .57/.19/.19/.05, seed 1
    (plus R-MAT scales 16 and 18 with seeds 11 and 12 of the same family).
Every row with at least one product is kept: short rows are sized for the
LDS-hash tables, long rows for the dense-tile kernel's output buffer.
Training runs on the GPU with the library's own forward/backward/Adam
kernels (rt_mlp_train).  Writes weights/rownnz_mlp.bin (loaded by the
library on first use of the learned method) and weights/rownnz_mlp.json.
"""
import argparse
import json
import os
import struct
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import rtspmspm_amd as rt  # noqa: E402

TH = 4096


def row_products(A, B):
    blen = np.diff(B.row_ptr)
    out = np.zeros(A.shape[0], np.int64)
    nz = np.diff(A.row_ptr) > 0
    sums = np.add.reduceat(np.append(blen[A.col], 0), A.row_ptr[:-1])
    out[nz] = sums[nz]
    return out


def collect(A, B):
    feat, nnz = rt.row_features(A, B)
    P = row_products(A, B)
    U = np.minimum(P, B.shape[1])
    keep = P > 0
    return feat[keep], np.log2(1.0 + nnz[keep]).astype(np.float32)


def training_sets():
    return [("uniform 4096 x 8, seed 0", lambda: rt.gen_uniform_k(4096, 8, 0)),
            ("R-MAT s20 ef16 seed 1", lambda: rt.gen_rmat(20, 16, seed=1)),
            ("R-MAT s18 ef16 seed 12", lambda: rt.gen_rmat(18, 16, seed=12)),
            ("R-MAT s16 ef16 seed 11", lambda: rt.gen_rmat(16, 16, seed=11))]


def save(out_dir, params, margin, meta):
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "rownnz_mlp.bin"), "wb") as f:
        f.write(b"RTMLP001")
        f.write(struct.pack("<i", params.size))
        f.write(struct.pack("<f", margin))
        f.write(np.ascontiguousarray(params, np.float32).tobytes())
    with open(os.path.join(out_dir, "rownnz_mlp.json"), "w") as f:
        json.dump(meta, f, indent=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "weights"))
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    Xs, ys, names = [], [], []
    for name, make in training_sets():
        t0 = time.time()
        A = make()
        X, y = collect(A, A)
        Xs.append(X)
        ys.append(y)
        names.append({"set": name, "rows": int(y.size)})
        print(f"{name}: {y.size} rows ({time.time() - t0:.1f} s)", flush=True)
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    t0 = time.time()
    loss = rt.mlp_train(X, y, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    params, margin = rt.mlp_get_weights()
    pred = rt.mlp_predict(params, X)
    over = float(np.mean(y > pred + margin))
    cap = np.minimum(np.ceil(np.exp2(pred + margin) - 1), np.exp2(X[:, 6] * 32) - 1)
    ratio = float(np.mean(np.maximum(cap, 1) / np.maximum(np.exp2(y) - 1, 1)))
    print(f"trained {args.epochs} epochs on {y.size} rows in {time.time() - t0:.1f} s: "
          f"mse {loss:.4f}, margin {margin:.3f}, overflow on training rows {over:.4f}, "
          f"mean capacity/nnz {ratio:.2f}", flush=True)
    save(args.out, params, margin,
         {"train_sets": names, "rows": int(y.size), "epochs": args.epochs, "batch": args.batch,
          "lr": args.lr, "seed": args.seed, "mse": loss, "margin": margin,
          "train_overflow_rate": over, "mean_capacity_over_nnz": ratio,
          "features": ["log2(1+nnzA)/16", "log2(1+P)/32", "log2(1+maxlenB)/24",
                       "log2(1+P/nnzA)/24", "occupied/64 of 64-bucket sampled column hash",
                       "log2(1+samples)/8", "log2(1+min(P,cols))/32", "log2(1+cols)/32"],
          "target": "log2(1 + row nnz)"})


if __name__ == "__main__":
    main()
