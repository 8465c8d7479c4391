"""GPU tests of the C boundary beyond the product itself: the learned
long-row kernel forced on, the out-of-memory paths (every entry point
returns RT_E_NOMEM and releases what it took), concurrent calls from two
host threads on two streams, trusted operands, plan argument checks and
the ray ablations' output."""
import threading

import numpy as np
import pytest

import oracle
import rtspmspm_amd as rt
from tests.helpers import check_parity, csr_from_coo

pytestmark = pytest.mark.gpu


def _hub_operands():
    """1,200 long rows over a B of 4.8M entries with ten 5,000-entry hub rows."""
    rng = np.random.default_rng(21)
    nb, ncols = 120_000, 1_000_000
    lens = np.full(nb, 40)
    lens[:10] = 5000
    br = np.repeat(np.arange(nb), lens)
    gaps = rng.integers(1, ncols // 40, size=(nb - 10, 40))
    bc = np.concatenate([rng.choice(ncols, n, replace=False) for n in lens[:10]] +
                        [np.cumsum(gaps, axis=1).ravel()])
    B = csr_from_coo(nb, ncols, br, bc, rng.uniform(-1, 1, br.size))
    ar, ac = [], []
    for i in range(1200):
        ks = rng.choice(np.arange(10, nb), 110, replace=False)
        if i % 3 == 0:
            ks[:3] = rng.choice(10, 3, replace=False)
        ar.append(np.full(ks.size, i))
        ac.append(ks)
    ar, ac = np.concatenate(ar), np.concatenate(ac)
    A = csr_from_coo(1200, nb, ar, ac, rng.uniform(-1, 1, ar.size))
    return A, B


LONG_CASES = {
    "dense_and_compressed": lambda: (rt.gen_fixed_normal(200, 40000, 400, 41),
                                     rt.gen_uniform_k(40000, 400, 42)),
    "sparse_many_tiles": lambda: (rt.gen_fixed_normal(100, 5000, 250, 43),
                                  rt.gen_fixed_normal(5000, 1_000_000, 20, 44)),
    "hub_rows": _hub_operands,
    "rmat13": lambda: (rt.gen_rmat(13, 16, seed=25),) * 2,
}


@pytest.mark.parametrize("case", sorted(LONG_CASES))
@pytest.mark.parametrize("long", ["learned", "exact"])
def test_learned_long_rows_forced(case, long):
    """The learned method with its long rows sized by the predictor and run
    by the learned dense-tile kernel (k_num_heavy<*, true>), or sent to the
    exact bitmap path: C equals the oracle either way."""
    A, B = LONG_CASES[case]()
    C, info = rt.spmspm(A, B, mode="learned", learned_long=long)
    assert np.diff(C.row_ptr).max() > 4096  # long rows are present
    err = check_parity(C, oracle.spgemm_f64(A, B))
    print(f"{case}/{long}: predicted {info['predicted_rows']} overflow {info['overflow_rows']} "
          f"max err/sum|ab| {err:.2e}")


def _live():
    s = rt.pool_stats()
    return s["live_bytes"], s["live_blocks"]


def _each_failure(call, limit=400):
    """Arms the n-th allocation to fail for n = 1, 2, ... until a call runs
    through without reaching it; every failing call must raise RT_E_NOMEM and
    leave the pool's live allocations as they were.  Returns the count of
    allocations the call makes."""
    base = _live()
    for n in range(1, limit):
        rt.debug_fail_alloc(n)
        try:
            out = call()
        except rt.RtError as e:
            assert e.code == rt.RT_E_NOMEM, (n, str(e))
            rt.synchronize()
            assert _live() == base, f"allocation {n}: live pool blocks leaked"
            continue
        finally:
            rt.debug_fail_alloc(0)
        for o in out:
            o.free()
        rt.synchronize()
        assert _live() == base
        return n - 1
    raise AssertionError("call kept allocating")


def test_out_of_memory_paths_release_everything():
    A = rt.gen_rmat(12, 16, seed=71)  # hash bins and long rows
    dA = rt.to_device(A)
    ar, br, bc, SA, SB = rt.gen_batch_poisson(12, 64, 3000, 4.0, 32.0, 72)
    dSA, dSB = rt.to_device(SA), rt.to_device(SB)
    calls = {
        "plain": lambda: [rt.spmspm(dA, dA, mode="plain")[0]],
        "plain_graph": lambda: [rt.spmspm(dA, dA, mode="plain", graph=True)[0]],
        "learned": lambda: [rt.spmspm(dA, dA, mode="learned")[0]],
        "learned_long": lambda: [rt.spmspm(dA, dA, mode="learned", learned_long="learned")[0]],
        "ray": lambda: [rt.spmspm(dA, dA, mode="ray", ray_output="slot")[0]],
        "batched_plain": lambda: [rt.spmspm_batched(ar, br, bc, dSA, dSB, mode="plain")[0]],
        "batched_learned": lambda: [rt.spmspm_batched(ar, br, bc, dSA, dSB, mode="learned")[0]],
        "host_operands": lambda: [],
        "plan": lambda: [rt.Plan(dA, dA)],
    }
    counts = {}
    for name, call in calls.items():
        if name == "host_operands":
            def call():
                rt.spmspm(A, A)  # host result: nothing stays in the pool
                return []
        counts[name] = _each_failure(call)
    print("allocations per call:", counts)
    assert all(v > 3 for v in counts.values()), counts
    C, _ = rt.spmspm(dA, dA)  # the library still works afterwards
    check_parity(C.to_host(), oracle.spgemm_f64(A, A))


def test_two_host_threads_on_two_streams():
    """Different products run concurrently from two host threads, each on
    its own stream (one of them profiling); each result equals the oracle."""
    ops = [(rt.gen_rmat(12, 16, seed=81),) * 2,
           (rt.gen_uniform_k(20000, 60, 82), rt.gen_uniform_k(20000, 60, 83))]
    refs = [oracle.spgemm_f64(a, b) for a, b in ops]
    results, errors = {}, []

    def work(t):
        try:
            a, b = ops[t]
            da, db = rt.to_device(a), rt.to_device(b)
            st = rt.Stream()
            out = []
            for it in range(4):
                mode = ("plain", "learned")[it % 2]
                C, _ = rt.spmspm(da, db, mode=mode, stream=st, profile=(t == 0))
                out.append(C.to_host())
                C.free()
            st.close()
            results[t] = out
        except Exception as e:  # reported by the main thread
            errors.append(e)

    th = [threading.Thread(target=work, args=(t,)) for t in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=300)
    assert not errors, errors
    for t in range(2):
        for C in results[t]:
            check_parity(C, refs[t])
    rt.prof_reset()


def test_trusted_operands_give_the_same_result():
    A = rt.gen_rmat(12, 16, seed=91)
    dA = rt.to_device(A)
    ref = oracle.spgemm_f64(A, A)
    for mode in ("plain", "learned"):
        C, _ = rt.spmspm(dA, dA, mode=mode, trusted=True)
        check_parity(C.to_host(), ref)
    ar, br, bc, SA, SB = rt.gen_batch_poisson(8, 64, 2000, 4.0, 32.0, 92)
    dSA, dSB = rt.to_device(SA), rt.to_device(SB)
    Ct = rt.spmspm_batched(ar, br, bc, dSA, dSB, trusted=True)[0].to_host()
    Cv = rt.spmspm_batched(ar, br, bc, dSA, dSB)[0].to_host()
    assert np.array_equal(Ct.row_ptr, Cv.row_ptr) and np.array_equal(Ct.col, Cv.col)


def test_plan_execute_checks_operand_sizes():
    A = rt.gen_uniform_k(2000, 6, 93)
    dA = rt.to_device(A)
    plan = rt.Plan(dA, dA)
    small = rt.to_device(rt.gen_uniform_k(2000, 5, 94))
    with pytest.raises(ValueError):
        plan.execute(small, dA)
    with pytest.raises(ValueError):
        plan.execute(dA, small)
    assert plan.execute(dA, dA) > 0
    plan.close()


@pytest.mark.parametrize("abl", ["no_atomic", "traverse", "no_hit", "empty"])
def test_ray_ablation_outputs_as_documented(abl):
    """Dense output: the ablations touch no entry, so C has no entries (the
    reference's ablation builds leave their dense result zero).  Slot output:
    C's exact structure."""
    A = rt.gen_uniform_k(4096, 8, 0)
    ref = oracle.spgemm_f64(A, A)
    C, _ = rt.spmspm(A, A, mode="ray", ablation=abl, ray_output="dense")
    assert C.shape == (4096, 4096) and C.nnz == 0
    C, _ = rt.spmspm(A, A, mode="ray", ablation=abl, ray_output="slot")
    assert np.array_equal(C.row_ptr, ref.row_ptr) and np.array_equal(C.col, ref.col)
