"""The five seeded SpGEMM workloads (host CSR operands) and their sizes.

They stand in for the reference's SuiteSparse datasets, which the harness
downloaded and multiplied as A*A (scripts/download_dataset.sh,
scripts/AE_test.py:25-42,204-215):

  uniform4k   A*A, A 4,096 x 4,096, exactly 8 uniform columns per row, U(-1,1), seed 0
  rmat20      A*A, A R-MAT scale 20, edge factor 16, a/b/c/d .57/.19/.19/.05, seed 1
  galerkin    R*(A*P): A 27-point stencil on a 128^3 grid, P 2x2x2 aggregation, R = P^T
  rect        A 500,000 x 2,000,000 (lognormal(2.5, 1) row lengths clipped to [1, 2000]),
              B 2,000,000 x 500,000 (4 per row), N(0,1), seed 3
  batch2048   2,048 independent pairs, n log-uniform in [256, 65,536],
              Poisson(lambda ~ U(4, 32)) per row, U(-1,1), seed 4

`scale` shrinks a workload for tests (grid size, rows, pair count).
"""
from __future__ import annotations

import numpy as np

from . import (gen_aggregation, gen_batch_poisson, gen_fixed_normal, gen_lognormal_rows, gen_rmat,
               gen_stencil27, gen_uniform_k)


def row_products(A, B) -> np.ndarray:
    """P_i = sum over A(i,k) of nnz(B(k,:)) (int64)."""
    blen = np.diff(B.row_ptr)
    out = np.zeros(A.shape[0], np.int64)
    nz = np.diff(A.row_ptr) > 0
    if A.nnz:
        sums = np.add.reduceat(blen[A.col].astype(np.int64), A.row_ptr[:-1].clip(max=A.nnz - 1))
        out[nz] = sums[nz]
    return out


def csr_bytes(rows: int, nnz: int) -> int:
    """CSR footprint with int64 row pointers and 4+4-byte entries."""
    return 8 * (rows + 1) + 8 * nnz


def uniform4k(n: int = 4096, k: int = 8, seed: int = 0):
    return gen_uniform_k(n, k, seed)


def rmat(scale: int = 20, seed: int = 1):
    return gen_rmat(scale, 16, 0.57, 0.19, 0.19, seed=seed)


def galerkin(grid: int = 128, seed: int = 2):
    """(A, P, R) of the 27-point stencil and its 2x2x2 aggregation."""
    A = gen_stencil27(grid, seed)
    P, R = gen_aggregation(grid)
    return A, P, R


def rect(rows: int = 500_000, inner: int = 2_000_000, seed: int = 3):
    A = gen_lognormal_rows(rows, inner, 2.5, 1.0, 1, min(2000, inner), seed, 0)
    B = gen_fixed_normal(inner, rows, 4, seed, 1)
    return A, B


def batch(count: int = 2048, nmin: int = 256, nmax: int = 65536, seed: int = 4):
    return gen_batch_poisson(count, nmin, nmax, 4.0, 32.0, seed)


def split_pairs(a_rows, b_rows, b_cols, A, B):
    """The stacked batch operands as a list of (A_b, B_b) host CSR pairs."""
    from . import CSR
    pairs = []
    a0 = b0 = 0
    for p in range(len(a_rows)):
        ar, br = int(a_rows[p]), int(b_rows[p])
        sa, ea = A.row_ptr[a0], A.row_ptr[a0 + ar]
        sb, eb = B.row_ptr[b0], B.row_ptr[b0 + br]
        Ap = CSR((ar, br), A.row_ptr[a0:a0 + ar + 1] - sa, A.col[sa:ea], A.val[sa:ea])
        Bp = CSR((br, int(b_cols[p])), B.row_ptr[b0:b0 + br + 1] - sb, B.col[sb:eb], B.val[sb:eb])
        pairs.append((Ap, Bp))
        a0 += ar
        b0 += br
    return pairs
