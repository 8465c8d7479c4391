"""GPU tests of the learned row-sizing path: feature and network kernels
against the FP64 CPU restatement, training, and C identical to the exact
method whatever the predictor says (overflow rows rerun exactly)."""
import numpy as np
import pytest

import oracle
import rtspmspm_amd as rt
from oracle import mlp
from tests.helpers import check_parity

pytestmark = pytest.mark.gpu


def test_features_match_cpu_restatement():
    A = rt.gen_rmat(11, 16, seed=3)
    F, nnz = rt.row_features(A, A)
    ref = mlp.features(A, A)
    assert np.allclose(F, ref, rtol=0, atol=2e-7)
    assert np.array_equal(nnz, oracle.count(A, A))


def test_network_forward_and_gradient_match_fp64():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, mlp.NPARAM).astype(np.float32)
    X = rng.uniform(0, 1, (1000, mlp.NF)).astype(np.float32)
    y = rng.uniform(0, 20, 1000).astype(np.float32)
    loss, g = rt.mlp_loss_grad(p, X, y)
    loss64, g64 = mlp.loss_grad(p.astype(np.float64), X.astype(np.float64), y.astype(np.float64))
    assert abs(loss - loss64) <= 1e-4 * loss64
    scale = np.abs(g64).max()
    assert np.allclose(g, g64, rtol=1e-3, atol=1e-5 * scale), np.abs(g - g64).max()
    pred = rt.mlp_predict(p, X)
    out64, _ = mlp.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.allclose(pred, out64, rtol=1e-5, atol=1e-4)


@pytest.fixture(scope="module")
def trained():
    """Weights trained in-test on R-MAT scale 14 (features from the exact pass)."""
    A = rt.gen_rmat(14, 16, seed=21)
    F, nnz = rt.row_features(A, A)
    y = np.log2(1.0 + nnz).astype(np.float32)
    keep = nnz > 0
    loss0, _ = rt.mlp_loss_grad(np.zeros(rt.mlp_param_count(), np.float32), F[keep], y[keep])
    loss = rt.mlp_train(F[keep], y[keep], epochs=30, batch=2048, lr=3e-3, seed=0)
    params, margin = rt.mlp_get_weights()
    return params, margin, loss0, loss


def test_training_reduces_loss(trained):
    params, margin, loss0, loss = trained
    print(f"mse untrained (bound only) {loss0:.4f} -> trained {loss:.4f}, margin {margin:.3f}")
    assert loss < 0.5 * loss0
    assert 0.0 <= margin < 4.0


@pytest.mark.parametrize("case", ["cfg1", "rmat13", "rect"])
def test_learned_equals_exact(trained, case):
    params, margin, _, _ = trained
    rt.mlp_set_weights(params, margin)
    if case == "cfg1":
        A = rt.gen_uniform_k(4096, 8, 0)
        B = A
    elif case == "rmat13":
        A = rt.gen_rmat(13, 16, seed=22)
        B = A
    else:
        A = rt.gen_lognormal_rows(20000, 80000, 2.5, 1.0, 1, 2000, 3, 0)
        B = rt.gen_fixed_normal(80000, 20000, 4, 3, 1)
    Cl, il = rt.spmspm(A, B, mode="learned")
    Cp, _ = rt.spmspm(A, B, mode="plain")
    assert np.array_equal(Cl.row_ptr, Cp.row_ptr) and np.array_equal(Cl.col, Cp.col)
    check_parity(Cl, oracle.spgemm_f64(A, B))
    print(f"{case}: predicted rows {il['predicted_rows']}, overflow {il['overflow_rows']}")
    assert il["predicted_rows"] > 0


def test_forced_overflow_and_bad_weights_still_exact(trained):
    params, margin, _, _ = trained
    A = rt.gen_rmat(13, 16, seed=23)
    ref = oracle.spgemm_f64(A, A)
    try:
        rt.mlp_set_weights(params, -30.0)  # capacities of 1: nearly every row overflows
        C, info = rt.spmspm(A, A, mode="learned")
        check_parity(C, ref)
        assert info["overflow_rows"] > 0.5 * info["predicted_rows"]
        rt.mlp_set_weights(np.zeros_like(params), 0.0)  # predictor = upper bound
        C, info = rt.spmspm(A, A, mode="learned")
        check_parity(C, ref)
        assert info["overflow_rows"] == 0
    finally:
        rt.mlp_set_weights(params, margin)


def test_rejects_invalid_weights():
    with pytest.raises(rt.RtError):
        rt.mlp_set_weights(np.full(rt.mlp_param_count(), np.nan, np.float32), 0.0)
    with pytest.raises(rt.RtError):
        rt.mlp_set_weights(np.zeros(5, np.float32), 0.0)


def test_trained_network_tightens_tables(trained):
    """What the network buys over sizing by the upper bound U = min(P, cols):
    total table capacity, and the overflow rate it pays for it."""
    params, margin, _, _ = trained
    A = rt.gen_rmat(14, 16, seed=24)
    F, nnz = rt.row_features(A, A)
    U = np.round(np.exp2(F[:, 6].astype(np.float64) * 32) - 1)
    sel = (nnz > 0) & (U <= 4096)
    pred = rt.mlp_predict(params, F[sel])
    cap = np.clip(np.ceil(np.exp2(pred + margin) - 1), 1, U[sel])
    over = float(np.mean(nnz[sel] > cap))
    print(f"capacity: bound {U[sel].sum():.0f} slots, network {cap.sum():.0f} "
          f"({cap.sum() / U[sel].sum():.2f}x), exact {nnz[sel].sum()}, overflow {over:.3f}")
    assert cap.sum() < U[sel].sum()
    assert over < 0.05


def test_trained_network_gains_time_where_rows_have_many_products():
    """What the trained network gains, in time.  C = A*A for a 27-point
    stencil: 729 products and 125 entries per interior row.  A network
    trained in-test on the 24^3 stencil square predicts the rows' sizes, so
    the learned method sizes them into 512-slot tables, where the upper bound
    min(P, cols) = 729 needs 2048-slot ones; it then also beats the exact
    method, whose symbolic pass over the 729 products it replaces with
    features, a prediction and a copy of C.  C equals the exact method's."""
    shipped = rt.mlp_get_weights()
    try:
        T = rt.gen_stencil27(24, 5)
        F, nnz = rt.row_features(T, T)
        keep = nnz > 0
        rt.mlp_train(F[keep], np.log2(1.0 + nnz[keep]).astype(np.float32), epochs=60,
                     batch=1024, lr=3e-3, seed=0)
        trained = rt.mlp_get_weights()
        S = rt.gen_stencil27(48, 6)
        Cl, info = rt.spmspm(S, S, mode="learned")
        assert info["predicted_rows"] == S.shape[0] and info["overflow_rows"] == 0
        check_parity(Cl, oracle.spgemm_f64(S, S))

        A = rt.gen_stencil27(96, 7)
        dA = rt.to_device(A)

        def median_ms(mode):
            for _ in range(2):
                C, _ = rt.spmspm(dA, dA, mode=mode, trusted=True)
                C.free()
            ms = []
            for _ in range(7):
                C, i = rt.spmspm(dA, dA, mode=mode, trusted=True)
                ms.append(i["ms"])
                C.free()
            return float(np.median(ms))
        plain = median_ms("plain")
        learned = median_ms("learned")
        rt.mlp_set_weights(np.zeros_like(trained[0]), 0.0)  # upper-bound sizing
        bound = median_ms("learned")
        rt.mlp_set_weights(*trained)
        Cp = rt.spmspm(dA, dA, mode="plain")[0].row_ptr_host()
        Cq = rt.spmspm(dA, dA, mode="learned")[0].row_ptr_host()
        assert np.array_equal(Cp, Cq)
        print(f"stencil27 96^3 A*A: exact {plain:.3f} ms, learned {learned:.3f} ms "
              f"({plain / learned:.2f}x), learned with upper-bound sizing {bound:.3f} ms "
              f"(network gain {bound / learned:.2f}x)")
        assert learned < 0.8 * bound
        assert learned < 1.1 * plain
    finally:
        rt.mlp_set_weights(*shipped)
