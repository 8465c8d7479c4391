"""Shared test helpers: fixtures, reference products and parity checks."""
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_DIR = os.path.join(HERE, "golden")
FIXTURES = ["amazon0312", "ca-CondMat", "email-Enron", "mario002", "p2p-Gnutella31",
            "roadNet-CA", "web-Google", "webbase-1M", "wiki-Vote"]
# Relative accuracy bar per entry: |c - c64| <= REL_TOL * sum |a*b|.
REL_TOL = 1e-5


def golden():
    with open(os.path.join(GOLDEN_DIR, "golden.json")) as f:
        return json.load(f)


def load_fixture(name):
    from rtspmspm_amd import CSR
    z = np.load(os.path.join(GOLDEN_DIR, f"{name}.npz"), allow_pickle=False)
    return CSR(tuple(int(x) for x in z["shape"]), z["row_ptr"], z["col"], z["val"])


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def csr_from_dense(D):
    from rtspmspm_amd import CSR
    D = np.asarray(D, dtype=np.float32)
    rows, cols = D.shape
    mask = D != 0
    rp = np.zeros(rows + 1, np.int64)
    np.cumsum(mask.sum(1), out=rp[1:])
    r, c = np.nonzero(mask)
    return CSR((rows, cols), rp, c.astype(np.int32), D[r, c].astype(np.float32))


def csr_from_coo(rows, cols, r, c, v):
    """CSR keeping explicit zeros (pattern given by (r, c))."""
    from rtspmspm_amd import CSR
    r = np.asarray(r, np.int64)
    c = np.asarray(c, np.int64)
    v = np.asarray(v, np.float32)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    rp = np.zeros(rows + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    np.cumsum(rp, out=rp)
    return CSR((rows, cols), rp, c.astype(np.int32), v)


def check_parity(C, ref, rel_tol=REL_TOL):
    """Structure bit-exact, values within rel_tol * sum|a*b| of the FP64 oracle.
    Returns the largest |err| / sum|a*b| seen."""
    assert tuple(C.shape) == tuple(ref.shape), (C.shape, ref.shape)
    assert np.array_equal(np.asarray(C.row_ptr, np.int64), ref.row_ptr), "row_ptr differs"
    assert np.array_equal(np.asarray(C.col, np.int32), ref.col), "column indices differ"
    err = np.abs(C.val.astype(np.float64) - ref.val)
    bound = rel_tol * ref.absum
    bad = err > bound
    assert not bad.any(), (
        f"{int(bad.sum())} entries beyond tolerance; worst err {err[bad].max():.3e} "
        f"vs bound {bound[bad][np.argmax(err[bad])]:.3e}")
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = np.where(ref.absum > 0, err / ref.absum, 0.0)
    return float(ratio.max()) if ratio.size else 0.0


def check_against_golden(C, g):
    """Compare a result with the scipy-derived golden summary of a fixture."""
    assert list(C.shape) == g["shape"]
    assert C.nnz == g["nnz_C"]
    row_nnz = np.diff(np.asarray(C.row_ptr, np.int64))
    assert sha(row_nnz) == g["row_nnz_sha256"], "per-row nnz differs from the pattern product"
    assert sha(np.asarray(C.col, np.int32)) == g["col_sha256"], "columns differ"
    v = C.val.astype(np.float64)
    j = np.asarray(C.col, np.float64) + 1.0
    assert abs(v.sum() - g["val_sum"]) <= REL_TOL * g["absprod_sum"] + 1e-12
    assert abs((j * v).sum() - g["checksum"]) <= REL_TOL * g["absprod_checksum"] + 1e-12
