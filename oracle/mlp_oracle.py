"""FP64 restatement of the learned-frontier MLP (test infrastructure only).

6 features -> 64 hidden (ReLU) -> 1 logit, weighted binary cross-entropy
    loss = sum_i c_i BCE(sigmoid(z_i), y_i) / sum_i c_i,  c_i = pos_w if y_i else 1
Parameter vector layout (as the engine's): W1[16][64] (feature rows 6..15
zero), b1[64], w2[64], b2.  The GPU computes in BF16 on MFMA with FP32
accumulation; `bf16()` rounds like the device so tests can compare the two
at matched input precision.  Pinned by finite differences in tests/test_mlp.py.
"""
import numpy as np

NFEAT, NH = 6, 64
W1, B1, W2, B2 = 0, 16 * NH, 16 * NH + NH, 16 * NH + 2 * NH
NPARAM = B2 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return p[W1:W1 + 16 * NH].reshape(16, NH)[:NFEAT], p[B1:B1 + NH], p[W2:W2 + NH], p[B2]


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward(p, X):
    w1, b1, w2, b2 = unpack(p)
    X = np.asarray(X, np.float64)
    z1 = X @ w1 + b1
    a = np.maximum(z1, 0)
    z = a @ w2 + b2
    return z1, a, z


def loss_and_grad(p, X, y, pos_w=1.0):
    """Returns (loss, grad[NPARAM]) in FP64."""
    w1, b1, w2, b2 = unpack(p)
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    c = np.where(y > 0, pos_w, 1.0)
    z1, a, z = forward(p, X)
    wsum = c.sum()
    bce = np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))
    loss = (c * bce).sum() / wsum
    dz = c * (1.0 / (1.0 + np.exp(-z)) - y) / wsum
    g = np.zeros(NPARAM)
    g[B2] = dz.sum()
    g[W2:W2 + NH] = a.T @ dz
    dh = np.outer(dz, w2) * (z1 > 0)
    g[B1:B1 + NH] = dh.sum(0)
    gw1 = np.zeros((16, NH))
    gw1[:NFEAT] = X.T @ dh
    g[W1:W1 + 16 * NH] = gw1.ravel()
    return loss, g


def prob(p, X):
    return 1.0 / (1.0 + np.exp(-forward(p, X)[2]))
