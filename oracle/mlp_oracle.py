"""CPU restatement of the next-treelet MLP (forward, loss, backward) in numpy.

TEST INFRASTRUCTURE ONLY: the checker for the device MLP of
treelet-prefetching-for-rt_amd/csrc/mlp.hip; never used by the product path.

The network has no counterpart in the reference (its RT unit uses popularity
heuristics only, src/gpgpu-sim/shader.cc:3419-3559), so it is pinned by
mathematics rather than reference vectors:
  * with bf16=False and float64 parameters, `loss_and_grad` is checked against
    central finite differences of `loss` (tests/test_mlp_oracle.py);
  * `bf16_round` is checked against the bit patterns of round-to-nearest-even;
  * the device GEMM is checked against `matmul_bf16` on the GPU.

Layout of the parameter vector (FP32), as in rtx_mlp_get_params:
  E [(T+1) x 32] | W1 [192 x 128] | b1 [128] | W2 [128 x 128] | b2 [128] |
  W3 [128 x 32] | b3 [32] | O [C x 32] | o [C]
Input row x (169 used of 192): E[cur] | one-hot octant (8) | depth | E[h1..h4]
where a history entry of 0xFFFFFFFF reads row T ("none").
"""
import numpy as np

EMB, HID, IN, IN_PAD = 32, 128, 169, 192
NONE = 0xFFFFFFFF


def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def matmul_bf16(a, b, bf16=True):
    """op(a) @ op(b) with operands rounded to bf16 and a float64 accumulation, cast to float32."""
    if bf16:
        a, b = bf16_round(a), bf16_round(b)
    return (np.asarray(a, np.float64) @ np.asarray(b, np.float64)).astype(np.float32)


def layout(n_treelets, n_classes):
    sizes = [("E", (n_treelets + 1, EMB)), ("W1", (IN_PAD, HID)), ("b1", (HID,)), ("W2", (HID, HID)),
             ("b2", (HID,)), ("W3", (HID, EMB)), ("b3", (EMB,)), ("O", (n_classes, EMB)), ("o", (n_classes,))]
    out, off = {}, 0
    for name, shape in sizes:
        n = int(np.prod(shape))
        out[name] = (off, shape)
        off += n
    return out, off


def unpack(vec, n_treelets, n_classes):
    lay, total = layout(n_treelets, n_classes)
    assert len(vec) == total
    return {k: vec[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in lay.items()}


def pack(parts, n_treelets, n_classes):
    lay, total = layout(n_treelets, n_classes)
    vec = np.zeros(total, dtype=parts["E"].dtype)
    for k, (o, s) in lay.items():
        vec[o:o + int(np.prod(s))] = parts[k].reshape(-1)
    return vec


def inputs(samples, E):
    """The 192-wide input rows of build_input."""
    n, T = len(samples), E.shape[0] - 1
    X = np.zeros((n, IN_PAD), dtype=E.dtype)
    X[:, 0:32] = E[samples["cur"]]
    X[np.arange(n), 32 + samples["octant"].astype(np.int64)] = 1
    X[:, 40] = samples["depth"]
    for h in range(4):
        t = samples["hist"][:, h].astype(np.int64)
        t = np.where(t == NONE, T, t)
        X[:, 41 + 32 * h:73 + 32 * h] = E[t]
    return X


def _mm(a, b, bf16):
    if bf16:
        return matmul_bf16(a, b)
    return a @ b


def forward(P, samples, bf16=True):
    X = inputs(samples, P["E"])
    H1 = np.maximum(_mm(X, P["W1"], bf16) + P["b1"], 0)
    H2 = np.maximum(_mm(H1, P["W2"], bf16) + P["b2"], 0)
    Q = _mm(H2, P["W3"], bf16) + P["b3"]
    L = _mm(Q, P["O"].T, bf16) + P["o"]
    return X, H1, H2, Q, L


def loss(P, samples, bf16=True):
    *_, L = forward(P, samples, bf16)
    m = L.max(axis=1, keepdims=True)
    lse = np.log(np.exp(L - m).sum(axis=1)) + m[:, 0]
    return float(np.mean(lse - L[np.arange(len(L)), samples["target"].astype(np.int64)]))


def loss_and_grad(P, samples, bf16=True):
    """Mean softmax cross-entropy and the gradient of every parameter block,
    with the same operand rounding as the device backward pass."""
    n = len(samples)
    X, H1, H2, Q, L = forward(P, samples, bf16)
    tgt = samples["target"].astype(np.int64)
    m = L.max(axis=1, keepdims=True)
    ex = np.exp(L - m)
    se = ex.sum(axis=1, keepdims=True)
    lval = float(np.mean(np.log(se[:, 0]) + m[:, 0] - L[np.arange(n), tgt]))
    dL = ex / se
    dL[np.arange(n), tgt] -= 1
    dL /= n
    G = {}
    G["O"] = _mm(dL.T, Q, bf16)
    G["o"] = dL.sum(axis=0)
    dQ = _mm(dL, P["O"], bf16)
    G["W3"] = _mm(H2.T, dQ, bf16)
    G["b3"] = dQ.sum(axis=0)
    dH2 = _mm(dQ, P["W3"].T, bf16) * (H2 > 0)
    G["W2"] = _mm(H1.T, dH2, bf16)
    G["b2"] = dH2.sum(axis=0)
    dH1 = _mm(dH2, P["W2"].T, bf16) * (H1 > 0)
    G["W1"] = _mm(X.T, dH1, bf16)
    G["b1"] = dH1.sum(axis=0)
    dX = _mm(dH1, P["W1"].T, bf16)
    dE = np.zeros_like(P["E"])
    T = P["E"].shape[0] - 1
    np.add.at(dE, samples["cur"].astype(np.int64), dX[:, 0:32])
    for h in range(4):
        t = samples["hist"][:, h].astype(np.int64)
        t = np.where(t == NONE, T, t)
        np.add.at(dE, t, dX[:, 41 + 32 * h:73 + 32 * h])
    G["E"] = dE
    return lval, G


def top4(P, samples, bf16=True):
    """Top-4 classes (ties to the lower class) and softmax probabilities."""
    *_, L = forward(P, samples, bf16)
    order = np.argsort(-L, axis=1, kind="stable")[:, :4]
    m = L.max(axis=1, keepdims=True)
    p = np.exp(L - m) / np.exp(L - m).sum(axis=1, keepdims=True)
    return order.astype(np.uint32), np.take_along_axis(p, order, axis=1)
