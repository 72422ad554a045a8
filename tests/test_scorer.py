"""The learned-ordering pair scorer: the FP64 reference (oracle/mlp_ref.py)
pinned by finite differences, and the matrix-core kernels (forward,
gradient, training) checked against it on the GPU."""
import numpy as np
import pytest

from oracle import mlp_ref as M


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0.0, 1.0, (n, M.NF))
    w = rng.normal(0.0, 1.0, M.NF)
    y = np.tanh(X @ w * 0.3) + 0.1 * X[:, 0] ** 2
    return X, y


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM)
    p[M.P_W1:M.P_B1] = rng.uniform(-1, 1, M.NF * M.NH) * np.sqrt(6 / M.NF)
    p[M.P_B1:M.P_W2] = rng.uniform(-0.1, 0.1, M.NH)
    p[M.P_W2:M.P_B2] = rng.uniform(-1, 1, M.NH * M.NH) * np.sqrt(6 / M.NH)
    p[M.P_B2:M.P_W3] = rng.uniform(-0.1, 0.1, M.NH)
    p[M.P_W3:M.P_B3] = rng.uniform(-1, 1, M.NH) * 0.3
    p[M.P_B3] = 0.05
    return p


def test_reference_backward_matches_finite_differences():
    X, y = _data(48, 1)
    p = _params(2)
    _, g = M.loss_grad(p, X, y, bf=False)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(np.arange(M.P_W1, M.P_B1), 12, replace=False),
                          rng.choice(np.arange(M.P_W2, M.P_B2), 12, replace=False),
                          np.arange(M.P_B1, M.P_B1 + 4), np.arange(M.P_B2, M.P_B2 + 4),
                          np.arange(M.P_W3, M.P_W3 + 4), [M.P_B3]])
    h = 1e-6
    for i in idx:
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        fd = (M.loss_grad(a, X, y, bf=False)[0] - M.loss_grad(b, X, y, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_matches_hardware_convention():
    # BF16 keeps 8 significand bits: spacing 2^-7 on [1, 2)
    x = np.array([1.0, 1.0078125, 1.00390625, 1.01171875, 1.005859375, -3.14159], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0078125
    assert r[2] == 1.0                      # halfway, to even (down)
    assert r[3] == 1.015625                 # halfway, to even (up)
    assert r[4] == 1.0078125                # nearer the upper neighbour
    assert r[5] == -3.140625


@pytest.mark.gpu
def test_scorer_kernels_match_reference():
    import c2a_amd as G
    X, y = _data(700, 4)                    # not a multiple of 64: padded rows
    p = _params(5).astype(np.float32)
    s = G.Scorer(7)
    s.set_params(p)
    print("scorer: infer / grad over %d rows" % len(y), flush=True)
    got = s.infer(X.astype(np.float32))
    ref, _ = M.forward(p.astype(np.float64), X.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(got - ref)) <= 2e-4 * max(1.0, np.max(np.abs(ref)))
    loss, g = s.grad(X.astype(np.float32), y.astype(np.float32))
    rl, rg = M.loss_grad(p.astype(np.float64), X.astype(np.float32).astype(np.float64), y)
    assert abs(loss - rl) <= 1e-4 * max(1.0, rl)
    for lo, hi, name in [(M.P_W1, M.P_B1, "W1"), (M.P_B1, M.P_W2, "b1"), (M.P_W2, M.P_B2, "W2"),
                         (M.P_B2, M.P_W3, "b2"), (M.P_W3, M.P_B3, "w3"), (M.P_B3, M.NPARAM, "b3")]:
        err = np.max(np.abs(g[lo:hi] - rg[lo:hi]))
        scale = max(1e-3, np.max(np.abs(rg[lo:hi])))
        assert err <= 5e-3 * scale, (name, err, scale)


@pytest.mark.gpu
def test_scorer_training_fits():
    import c2a_amd as G
    X, y = _data(8192, 6)
    s = G.Scorer(11)
    p0 = s.params()
    l0 = M.loss_grad(p0.astype(np.float64), X.astype(np.float32).astype(np.float64), y)[0]
    losses = s.train(X.astype(np.float32), y.astype(np.float32), epochs=20, batch=512, lr=3e-3,
                     seed=1)
    print("scorer training: loss %.4f -> %.4f" % (l0, losses[-1]), flush=True)
    l1 = M.loss_grad(s.params().astype(np.float64), X.astype(np.float32).astype(np.float64), y)[0]
    assert losses[-1] < 0.2 * l0 and l1 < 0.2 * l0
    # reproducible
    s2 = G.Scorer(11)
    s2.train(X.astype(np.float32), y.astype(np.float32), epochs=20, batch=512, lr=3e-3, seed=1)
    assert np.array_equal(s.params(), s2.params())
