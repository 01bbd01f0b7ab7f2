"""CPU checks of the MLP oracle: BF16 rounding and gradients by finite differences."""
import numpy as np
import pytest

import mlp_oracle as mo


def test_bf16_round_known_values():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # bf16 spacing at 1.0
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 0.0, -0.0, 3.0e38],
                 np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 * ulp, 1.0, -2.5, 0.0, -0.0, 2.9980e38], np.float32)
    got = mo.bf16_round(x)
    assert got[0] == one and got[1] == one  # ties to even: 1 + 2^-8 -> 1
    assert got[2] == np.float32(1.0 + 2 * ulp)  # 1 + 1.5 ulp -> the even neighbour 1 + 2 ulp
    assert got[3] == one and got[4] == np.float32(-2.5)
    assert np.signbit(got[6]) and got[5] == 0
    assert abs(float(got[7]) - float(want[7])) / 3e38 < 2 ** -7


def test_bf16_round_brute_force():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    got = mo.bf16_round(x)
    u = x.view(np.uint32)
    lo = (u & 0xFFFF0000).view(np.float32).astype(np.float64)
    hi = ((u & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dl, dh = np.abs(xd - lo), np.abs(hi - xd)
    lo_even = ((u >> 16) & 1) == 0
    want = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where(lo_even, lo, hi)))
    assert np.array_equal(got.astype(np.float64), want)


def _tiny(rng, T=5, C=7, n=6, dtype=np.float64):
    lay, total = mo.layout(T, C)
    vec = rng.standard_normal(total).astype(dtype) * 0.3
    P = mo.unpack(vec, T, C)
    P["W1"][mo.IN:] = 0
    s = np.zeros(n, dtype=[("cur", "<u4"), ("octant", "<u4"), ("depth", "<f4"), ("hist", "<u4", 4),
                           ("target", "<u4")])
    s["cur"] = rng.integers(0, T, n)
    s["octant"] = rng.integers(0, 8, n)
    s["depth"] = rng.uniform(0, 1, n)
    s["hist"] = rng.integers(0, T, (n, 4))
    s["hist"][0, 2:] = mo.NONE
    s["target"] = rng.integers(0, C, n)
    return vec, P, s


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_gradient_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    T, C = 5, 7
    vec, P, s = _tiny(rng, T, C)
    _, G = mo.loss_and_grad(P, s, bf16=False)
    g = mo.pack(G, T, C)
    lay, total = mo.layout(T, C)
    eps = 1e-6
    checked = 0
    for name, (off, shape) in lay.items():
        n = int(np.prod(shape))
        idx = rng.choice(n, size=min(n, 12), replace=False)
        for i in idx:
            if name == "W1" and (i // mo.HID) >= mo.IN:
                continue
            v = vec.copy()
            v[off + i] += eps
            lp = mo.loss(mo.unpack(v, T, C), s, bf16=False)
            v[off + i] -= 2 * eps
            lm = mo.loss(mo.unpack(v, T, C), s, bf16=False)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[off + i]) <= 1e-6 + 1e-5 * abs(fd), (name, i, fd, g[off + i])
            checked += 1
    assert checked > 60


def test_oracle_none_history_reads_last_row():
    rng = np.random.default_rng(3)
    T, C = 5, 7
    _, P, s = _tiny(rng, T, C)
    X = mo.inputs(s, P["E"])
    assert np.array_equal(X[0, 41 + 64:41 + 96], P["E"][T])
    assert np.array_equal(X[0, 0:32], P["E"][s["cur"][0]])
    assert X[0, 32 + s["octant"][0]] == 1 and X[0, 32:40].sum() == 1
    assert np.all(X[:, mo.IN:] == 0)


def test_bf16_forward_close_to_exact():
    rng = np.random.default_rng(4)
    T, C = 5, 7
    vec, P, s = _tiny(rng, T, C, n=32, dtype=np.float32)
    l32 = mo.loss(P, s, bf16=True)
    l64 = mo.loss(mo.unpack(vec.astype(np.float64), T, C), s, bf16=False)
    assert abs(l32 - l64) < 0.02 * max(1.0, abs(l64))
