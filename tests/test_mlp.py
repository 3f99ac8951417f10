"""Pins the FP64 MLP oracle (oracle/mlp_oracle.py) by central finite
differences, then (GPU) checks the engine's BF16/MFMA forward, gradients and
training against it."""
import numpy as np
import pytest

import mlp_oracle as M


def rand_params(rng):
    p = np.zeros(M.NPARAM)
    w1 = np.zeros((16, M.NH))
    w1[:M.NFEAT] = rng.normal(0, 0.5, (M.NFEAT, M.NH))
    p[M.W1:M.W1 + 16 * M.NH] = w1.ravel()
    p[M.B1:M.B1 + M.NH] = rng.normal(0, 0.1, M.NH)
    p[M.W2:M.W2 + M.NH] = rng.normal(0, 0.3, M.NH)
    p[M.B2] = 0.1
    return p


def data(rng, n):
    X = rng.uniform(-1, 1, (n, M.NFEAT))
    y = (rng.random(n) < 0.3).astype(np.uint8)
    return X, y


@pytest.mark.parametrize("pos_w", [1.0, 3.0])
def test_oracle_gradient_finite_differences(pos_w):
    rng = np.random.default_rng(1)
    p = rand_params(rng)
    X, y = data(rng, 50)
    _, g = M.loss_and_grad(p, X, y, pos_w)
    idx = list(range(0, 6 * M.NH, 37)) + list(range(M.B1, M.NPARAM, 11)) + [M.B2]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        num = (M.loss_and_grad(pp, X, y, pos_w)[0] - M.loss_and_grad(pm, X, y, pos_w)[0]) / (2 * eps)
        assert abs(num - g[i]) < 1e-7 + 1e-5 * abs(num), (i, num, g[i])
    # padded feature rows carry no gradient
    assert np.all(g[M.W1 + M.NFEAT * M.NH:M.B1] == 0)


def test_oracle_prob_matches_closed_form():
    # one hidden unit active: p = sigmoid(w2 * (x.w1 + b1) + b2)
    p = np.zeros(M.NPARAM)
    p[M.W1 + 0 * M.NH + 0] = 2.0
    p[M.W2 + 0] = 1.5
    p[M.B2] = -0.5
    X = np.array([[0.3, 0, 0, 0, 0, 0], [-0.3, 0, 0, 0, 0, 0]])
    pr = M.prob(p, X)
    assert pr[0] == pytest.approx(1 / (1 + np.exp(-(1.5 * 0.6 - 0.5))))
    assert pr[1] == pytest.approx(1 / (1 + np.exp(0.5)))


def bf16_params(p):
    q = np.asarray(p, np.float32).astype(np.float64).copy()
    q[M.W1:M.B1] = M.bf16(q[M.W1:M.B1])
    return q


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(engine):
    rng = np.random.default_rng(2)
    p = rand_params(rng).astype(np.float32)
    X, _ = data(rng, 1000)
    X = X.astype(np.float32)
    got = engine.mlp_forward(p, X)
    exp = M.prob(bf16_params(p), M.bf16(X))  # the device rounds W1 and X to bf16
    assert np.max(np.abs(got - exp)) < 2e-3


@pytest.mark.gpu
@pytest.mark.parametrize("pos_w", [1.0, 4.0])
def test_gpu_gradients_match_oracle(engine, pos_w):
    rng = np.random.default_rng(3)
    p = rand_params(rng).astype(np.float32)
    X, y = data(rng, 3000)
    X = X.astype(np.float32)
    g, loss = engine.mlp_grad(p, X, y, pos_w)
    l_ref, g_ref = M.loss_and_grad(bf16_params(p), M.bf16(X), y, pos_w)
    assert loss == pytest.approx(l_ref, rel=1e-3)
    # dW1 uses bf16 copies of X and of the hidden gradient: relative tolerance on the block norm
    for lo, hi in ((M.W1, M.W1 + M.NFEAT * M.NH), (M.B1, M.B1 + M.NH), (M.W2, M.W2 + M.NH), (M.B2, M.B2 + 1)):
        err = np.linalg.norm(g[lo:hi] - g_ref[lo:hi])
        assert err <= 1e-2 * np.linalg.norm(g_ref[lo:hi]) + 1e-6, (lo, err)
    assert np.all(g[M.W1 + M.NFEAT * M.NH:M.B1] == 0)


@pytest.mark.gpu
def test_gpu_training_learns_separable_rule(engine):
    # moves iff f2 (last gain) > 0.2: the trained net must recover it
    rng = np.random.default_rng(4)
    X, _ = data(rng, 20000)
    y = (X[:, 2] > 0.2).astype(np.uint8)
    p, loss = engine.mlp_train(X.astype(np.float32), y, epochs=30, batch=1024, lr=1e-2, seed=1)
    assert loss[-1] < 0.5 * loss[0]
    pr = engine.mlp_forward(p, X.astype(np.float32))
    assert ((pr > 0.5) == (y > 0)).mean() > 0.97
