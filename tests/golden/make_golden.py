"""Writes the golden fixtures for the 9 matrices the reference ships.

Run once in a container that has the reference mounted:
    python tests/golden/make_golden.py /root/reference/optixSpMSpM/src/data

The reference has no tests and no expected outputs; its only data are the
nine ``*_small.mtx`` inputs (top-left quarters of SuiteSparse/SNAP matrices,
scripts/matrixSampling.py:26-41) that its harness multiplies as A*A
(scripts/AE_test.py:204-215).  For each one this script stores

  <name>.npz    the input as CSR arrays (int64 row_ptr, int32 col, float32 val)
  golden.json   properties of C = A*A computed with scipy.sparse:
                  structure  -- from the pattern product |A|*|A| (all stored
                                entries, explicit zeros included, set to 1),
                                so no cancellation can remove an entry:
                                nnz, sha256 of the per-row nnz (int64) and of
                                the column indices (int32, sorted per row);
                  values     -- from the FP64 product (inputs cast to float64):
                                sum, sum |.|, checksum sum (j+1)*v, and the
                                accuracy scale sum_e sum|a*b| (= the sum of the
                                entries of |A|*|A| with |values|) and its
                                (j+1)-weighted form.
The scipy product drops exact zeros from its result, so it pins values
only; the structure comes from the pattern product.
"""
import hashlib
import json
import os
import sys

import numpy as np
import scipy.io
import scipy.sparse as sp

NAMES = ["amazon0312", "ca-CondMat", "email-Enron", "mario002", "p2p-Gnutella31",
         "roadNet-CA", "web-Google", "webbase-1M", "wiki-Vote"]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(data_dir: str) -> None:
    here = os.path.dirname(os.path.abspath(__file__))
    out = {}
    for name in NAMES:
        path = os.path.join(data_dir, name, f"{name}_small.mtx")
        M = scipy.io.mmread(path).tocsr()
        M.sum_duplicates()
        M.sort_indices()
        rp = M.indptr.astype(np.int64)
        col = M.indices.astype(np.int32)
        val = M.data.astype(np.float32)
        np.savez_compressed(os.path.join(here, f"{name}.npz"), row_ptr=rp, col=col, val=val,
                            shape=np.array(M.shape, np.int64))
        A64 = sp.csr_matrix((val.astype(np.float64), col, rp), shape=M.shape)
        pat = sp.csr_matrix((np.ones_like(val, dtype=np.float64), col, rp), shape=M.shape)
        absA = sp.csr_matrix((np.abs(val).astype(np.float64), col, rp), shape=M.shape)
        P = (pat @ pat).tocsr()
        P.sort_indices()
        C = (A64 @ A64).tocsr()
        C.sort_indices()
        Ab = (absA @ absA).tocsr()
        Ab.sort_indices()
        row_nnz = np.diff(P.indptr).astype(np.int64)
        rows_of = np.repeat(np.arange(C.shape[0]), np.diff(C.indptr))
        jc = C.indices.astype(np.float64) + 1.0
        jb = Ab.indices.astype(np.float64) + 1.0
        products = int(np.diff(rp)[col].sum()) if M.shape[0] == M.shape[1] else None
        out[name] = {
            "shape": list(M.shape),
            "nnz_A": int(M.nnz),
            "products": products,
            "nnz_C": int(P.nnz),
            "row_nnz_sha256": sha(row_nnz),
            "col_sha256": sha(P.indices.astype(np.int32)),
            "val_sum": float(C.data.sum()),
            "val_abs_sum": float(np.abs(C.data).sum()),
            "checksum": float((jc * C.data).sum()),
            "absprod_sum": float(Ab.data.sum()),
            "absprod_checksum": float((jb * Ab.data).sum()),
            "all_ones": bool(np.all(val == 1.0)),
            "max_row_nnz": int(row_nnz.max()) if row_nnz.size else 0,
        }
        del rows_of
        print(name, out[name]["nnz_C"], out[name]["products"])
    with open(os.path.join(here, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/optixSpMSpM/src/data")
