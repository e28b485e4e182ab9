"""FP64 reference of the task ranker (24 -> 64 ReLU -> 1, BCE with logits).

TEST INFRASTRUCTURE ONLY: the checker for the GPU ranker kernels' forward
pass and gradients.  The flat parameter layout is the C ABI's
tfsm_ranker_weights: w1 [64][24] row-major, b1 [64], w2 [64], b2.
"""
import numpy as np

F, H = 24, 64
P = H * F + 2 * H + 1


def split(flat):
    flat = np.asarray(flat, dtype=np.float64)
    w1 = flat[:H * F].reshape(H, F)
    b1 = flat[H * F:H * F + H]
    w2 = flat[H * F + H:H * F + 2 * H]
    b2 = flat[H * F + 2 * H]
    return w1, b1, w2, b2


def bf16_round(flat):
    """Round FP32 values to BF16 (round to nearest even), returned as FP32."""
    u = np.ascontiguousarray(flat, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + (((u >> 16) & 1) + 0x7FFF)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def forward(flat, X):
    w1, b1, w2, b2 = split(flat)
    X = np.asarray(X, dtype=np.float64)
    h = np.maximum(X @ w1.T + b1, 0.0)
    return h @ w2 + b2


def loss_and_grad(flat, X, y):
    """Mean BCE-with-logits loss and its gradient (flat layout)."""
    w1, b1, w2, b2 = split(flat)
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    a = X @ w1.T + b1
    h = np.maximum(a, 0.0)
    z = h @ w2 + b2
    loss = np.mean(np.logaddexp(0.0, z) - y * z)
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / len(X)
    gw2 = h.T @ dz
    gb2 = dz.sum()
    dh = np.outer(dz, w2) * (a > 0)
    gw1 = dh.T @ X
    gb1 = dh.sum(axis=0)
    return loss, np.concatenate([gw1.ravel(), gb1, gw2, [gb2]])
