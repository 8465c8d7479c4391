"""CPU pins of the FP64 predictor restatement (oracle/mlp.py)."""
import numpy as np
import pytest

from oracle import mlp
from tests.helpers import csr_from_coo


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mlp.NPARAM)
    X = rng.uniform(0, 1, (37, mlp.NF))
    y = rng.uniform(0, 20, 37)
    loss, g = mlp.loss_grad(p, X, y)
    h = 1e-6
    idx = np.concatenate([rng.choice(mlp.NPARAM, 60, replace=False), [mlp.OB3, mlp.OW1, mlp.OW2]])
    for q in idx:
        e = np.zeros_like(p)
        e[q] = h
        num = (mlp.loss_grad(p + e, X, y)[0] - mlp.loss_grad(p - e, X, y)[0]) / (2 * h)
        assert abs(num - g[q]) <= 1e-6 * max(1.0, abs(num)), (q, num, g[q])
    assert loss > 0


def test_mlp_forward_residual_on_upper_bound():
    p = np.zeros(mlp.NPARAM)
    X = np.zeros((3, mlp.NF))
    X[:, 6] = [0.0, 0.25, 0.5]
    out, _ = mlp.forward(p, X)
    assert np.allclose(out, [0.0, 8.0, 16.0])


def test_features_closed_form():
    # A row 0 -> B rows 0 (3 entries) and 1 (1 entry); A row 1 empty.
    A = csr_from_coo(2, 2, [0, 0], [0, 1], [1.0, 2.0])
    B = csr_from_coo(2, 10, [0, 0, 0, 1], [1, 4, 7, 4], [1.0, 1.0, 1.0, 1.0])
    F = mlp.features(A, B)
    assert np.isclose(F[0, 0], np.log2(3) / 16)
    assert np.isclose(F[0, 1], np.log2(5) / 32)          # P = 4
    assert np.isclose(F[0, 2], np.log2(4) / 24)          # longest B row 3
    assert np.isclose(F[0, 3], np.log2(3) / 24)          # mean 2
    assert np.isclose(F[0, 5], np.log2(5) / 8)           # 4 sampled columns
    assert F[0, 4] in (2 / 64, 3 / 64)                   # 3 distinct columns, may collide
    assert np.isclose(F[0, 6], np.log2(5) / 32)          # U = min(4, 10)
    assert np.isclose(F[0, 7], np.log2(11) / 32)
    assert np.allclose(F[1, [0, 1, 2, 3, 4, 5, 6]], 0.0)


def _bucket(c):
    """The 64-bucket column hash the sampled-occupancy feature is defined on."""
    return ((c * 0x9E3779B1) & 0xFFFFFFFF) >> 26


def test_features_sampling_of_long_rows():
    # An A row of 32 entries samples entries 0, 2, ..., 30 (16 of them, i*nA//16);
    # a B row of 20 entries samples positions t*20//8 = 0, 2, 5, 7, 10, 12, 15, 17.
    # Columns are chosen so that every sampled column hashes to bucket 5 and
    # every unsampled one elsewhere: a wrong entry or position stride shows up
    # as extra bits in the occupancy feature, and the sample count pins the
    # per-row cap of 8.
    sampled_pos = {0, 2, 5, 7, 10, 12, 15, 17}

    def increasing(n, want5, start):
        out, c = [], start
        for t in range(n):
            while (_bucket(c) == 5) != want5(t):
                c += 1
            out.append(c)
            c += 1
        return out

    rows, cols = [], []
    for k in range(32):                           # B row k, A entry k
        if k % 2 == 0:                            # sampled entry: a B row of 20
            cs = increasing(20, lambda t: t in sampled_pos, 1000 * k)
        else:                                     # unsampled entry: one column
            cs = increasing(1, lambda t: False, 1000 * k)
        rows += [k] * len(cs)
        cols += cs
    B = csr_from_coo(32, 1 << 16, rows, cols, np.ones(len(cols)))
    A = csr_from_coo(1, 32, [0] * 32, list(range(32)), np.ones(32))
    F = mlp.features(A, B)
    assert np.isclose(F[0, 0], np.log2(33) / 16)
    assert np.isclose(F[0, 5], np.log2(1 + 16 * 8) / 8)   # 16 B rows, 8 samples each
    assert F[0, 4] == 1 / 64                               # only bucket 5 seen
    # The same row with its odd entries sampled instead (a planted stride
    # error, restated here) would see the unsampled columns' buckets.
    odd = {_bucket(c) for r, c in zip(rows, cols) if r % 2 == 1}
    assert 5 not in odd and len(odd) > 1
