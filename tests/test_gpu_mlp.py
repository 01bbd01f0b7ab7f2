"""The MLP's MFMA kernels against the numpy oracle (oracle/mlp_oracle.py)."""
import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
@pytest.mark.parametrize("shape", [(16, 16, 32), (70, 50, 45), (129, 65, 200), (1, 300, 7),
                                   (100, 32, 2048), (300, 20, 700), (64, 96, 1536), (257, 33, 1000)])
def test_gemm_matches_bf16_oracle(gpu, ta, tb, shape):
    m, n, k = shape
    print(f"gemm m={m} n={n} k={k} ta={ta} tb={tb}", flush=True)
    rng = np.random.default_rng(m * 7 + n + k)
    a = rng.standard_normal((k, m) if ta else (m, k)).astype(np.float32)
    b = rng.standard_normal((n, k) if tb else (k, n)).astype(np.float32)
    got = gpu.gemm(a, b, ta, tb)
    want = mo.matmul_bf16(a.T if ta else a, b.T if tb else b)
    # identical bf16 operands on both sides: only the FP32 summation order differs
    scale = mo.matmul_bf16(np.abs(a.T if ta else a), np.abs(b.T if tb else b)) + 1e-30
    assert np.all(np.abs(got - want) <= 1e-5 * scale), np.max(np.abs(got - want) / scale)


def test_gemm_lane_layout_identity(gpu):
    # A = identity picks rows of B exactly (bf16-representable B): catches any
    # transposed or shifted operand/result placement
    rng = np.random.default_rng(1)
    b = mo.bf16_round(rng.standard_normal((32, 48)).astype(np.float32))
    got = gpu.gemm(np.eye(32, dtype=np.float32), b)
    assert np.array_equal(got, b)
    perm = rng.permutation(32)
    got = gpu.gemm(np.eye(32, dtype=np.float32)[perm], b)
    assert np.array_equal(got, b[perm])


def _samples(rng, T, C, n):
    s = np.zeros(n, mo_sample_dtype())
    s["cur"] = rng.integers(0, T, n)
    s["octant"] = rng.integers(0, 8, n)
    s["depth"] = rng.uniform(0, 1, n)
    s["hist"] = rng.integers(0, T, (n, 4))
    s["hist"][rng.uniform(size=(n, 4)) < 0.2] = mo.NONE
    s["target"] = rng.integers(0, C, n)
    return s


def mo_sample_dtype():
    import rtx

    return rtx.MLP_SAMPLE_DTYPE


@pytest.mark.parametrize("T,C,n", [(40, 96, 300), (60, 2500, 700), (20, 17, 33)])
def test_loss_and_gradients_match_oracle(gpu, T, C, n):
    """Loss and every gradient block against the oracle.  The output layer is
    fused (logits only in registers): 2,500 classes span three row-statistics
    chunks and sixteen class splits of the dQ pass; 17 classes and 33 rows
    leave most of every tile empty."""
    rng = np.random.default_rng(5 + C)
    mlp = gpu.mlp(T, C, seed=9)
    try:
        vec = mlp.params()
        P = mo.unpack(vec.copy(), T, C)
        P["o"][:] = rng.standard_normal(C).astype(np.float32) * 0.1
        for k in ("b1", "b2", "b3"):
            P[k][:] = rng.standard_normal(P[k].shape).astype(np.float32) * 0.1
        vec = mo.pack(P, T, C)
        mlp.set_params(vec)
        s = _samples(rng, T, C, n)
        loss = mlp.loss_grad(s)
        g = mo.unpack(mlp.grads(), T, C)
        lo, G = mo.loss_and_grad(mo.unpack(vec, T, C), s, bf16=True)
        assert abs(loss - lo) <= 1e-4 * max(1.0, abs(lo)), (loss, lo)
        for k in G:
            d = np.linalg.norm(g[k] - G[k]) / (np.linalg.norm(G[k]) + 1e-12)
            assert d < 1e-2, (k, d)
        # the padded rows of W1 stay untouched by the gradient
        assert np.all(g["W1"][mo.IN:] == 0)
    finally:
        mlp.close()


def test_predict_matches_oracle_top4(gpu):
    T, C, n = 30, 200, 500
    rng = np.random.default_rng(6)
    mlp = gpu.mlp(T, C, seed=3)
    try:
        s = _samples(rng, T, C, n)
        top, prob = mlp.predict(s)
        P = mo.unpack(mlp.params(), T, C)
        otop, oprob = mo.top4(P, s, bf16=True)
        *_, L = mo.forward(P, s, bf16=True)
        Ls = -np.sort(-L, axis=1)
        clear = np.all(np.abs(np.diff(Ls[:, :5], axis=1)) > 1e-3, axis=1)  # no near-ties in the top 5
        assert clear.mean() > 0.5
        assert np.array_equal(top[clear], otop[clear])
        assert np.allclose(prob[clear], oprob[clear], rtol=1e-3, atol=1e-5)
        assert np.all(np.diff(prob, axis=1) <= 1e-7)
    finally:
        mlp.close()


def test_training_learns_a_treelet_rule(gpu):
    # next class = a fixed function of (cur, octant): the network must fit it
    T, C, n = 64, 128, 8192
    rng = np.random.default_rng(7)
    rule = rng.integers(0, C, (T, 8))
    s = _samples(rng, T, C, n)
    s["target"] = rule[s["cur"], s["octant"]]
    mlp = gpu.mlp(T, C, seed=1)
    try:
        buf = mlp.upload(s)
        l0 = mlp.loss_grad(buf, n)
        last = mlp.train(buf, n, epochs=30, batch=512, lr=3e-3)
        top, _ = mlp.predict(buf, n)
        acc = np.mean(top[:, 0] == s["target"])
        print(f"loss {l0:.3f} -> {last:.3f}, top-1 accuracy {acc:.3f}")
        assert last < 0.25 * l0 and acc > 0.9
    finally:
        mlp.close()
