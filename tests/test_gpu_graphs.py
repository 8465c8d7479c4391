"""GPU tests of the graph-captured paths: the learned method's three
captured stages against its direct launches, and replays of cached
executable graphs -- the same stage run again on other operands of the same
shape (the cached graph is updated with the new addresses), with the
program's own eager work between replays, every replay checked."""
import numpy as np
import pytest

import oracle
import rtspmspm_amd as rt
from rtspmspm_amd import workloads as W
from tests.helpers import check_parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make", [lambda: (rt.gen_rmat(13, 16, seed=111),) * 2,
                                  lambda: (rt.gen_uniform_k(20000, 100, 112),) * 2,
                                  lambda: (rt.gen_fixed_normal(100, 5000, 250, 43),
                                           rt.gen_fixed_normal(5000, 1_000_000, 20, 44))])
@pytest.mark.parametrize("long", [None, "learned"])
def test_learned_graph_equals_direct(make, long):
    A, B = make()
    Cg, ig = rt.spmspm(A, B, mode="learned", graph=True, learned_long=long)
    Cd, idd = rt.spmspm(A, B, mode="learned", graph=False, learned_long=long)
    assert np.array_equal(Cg.row_ptr, Cd.row_ptr) and np.array_equal(Cg.col, Cd.col)
    assert ig["products"] == idd["products"] and ig["predicted_rows"] == idd["predicted_rows"]
    check_parity(Cg, oracle.spgemm_f64(A, B))


def _pair_parity(C, ar, br, bc, SA, SB):
    a0 = 0
    for (Ap, Bp), n in zip(W.split_pairs(ar, br, bc, SA, SB), ar):
        s, e = C.row_ptr[a0], C.row_ptr[a0 + n]
        Cq = rt.CSR((int(n), Bp.shape[1]), C.row_ptr[a0:a0 + n + 1] - s, C.col[s:e], C.val[s:e])
        check_parity(Cq, oracle.spgemm_f64(Ap, Bp))
        a0 += int(n)


@pytest.mark.parametrize("mode", ["plain", "learned"])
def test_cached_graph_replays_on_new_operands(mode):
    """Three batches of one shape (same pair sizes and structure, other
    values) run in turn through the batched graph path, twice over, with
    eager direct products in between; every result matches the oracle."""
    sets = []
    for seed in (121, 122, 123):
        ar, br, bc, SA, SB = rt.gen_batch_poisson(24, 64, 3000, 4.0, 32.0, 120)
        rng = np.random.default_rng(seed)
        SA.val = rng.uniform(-1, 1, SA.nnz).astype(np.float32)
        SB.val = rng.uniform(-1, 1, SB.nnz).astype(np.float32)
        sets.append((ar, br, bc, SA, SB, rt.to_device(SA), rt.to_device(SB)))
    E = rt.gen_rmat(11, 16, seed=124)
    for rep in range(2):
        for ar, br, bc, SA, SB, dA, dB in sets:
            C, info = rt.spmspm_batched(ar, br, bc, dA, dB, mode=mode)
            _pair_parity(C.to_host(), ar, br, bc, SA, SB)
            Ce, _ = rt.spmspm(E, E, mode="plain", graph=False)  # eager work between replays
            assert Ce.nnz > 0


def test_single_product_graph_replays_on_new_operands():
    """The plain graph path on two different matrices with the same row
    count, alternately: the second run of each stage updates the cached
    executable graph with the other operands' addresses."""
    mats = [rt.gen_rmat(13, 16, seed=s) for s in (131, 132)]
    refs = [oracle.spgemm_f64(M, M) for M in mats]
    for rep in range(2):
        for M, ref in zip(mats, refs):
            C, _ = rt.spmspm(M, M, mode="plain", graph=True)
            check_parity(C, ref)
            C, _ = rt.spmspm(M, M, mode="learned", graph=True)
            check_parity(C, ref)
