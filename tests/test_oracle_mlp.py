"""Pins the fp64 network restatement (oracle/mlp.py) that checks the device
MLP: central finite differences of the loss, a closed-form case, planted
mistakes."""
import numpy as np

from oracle import mlp as R


def _data(rng, rows=7):
    X = rng.uniform(0, 1, size=(rows, 5))
    Y = rng.normal(size=(rows, 8))
    M = (rng.uniform(size=(rows, 8)) < 0.7).astype(np.float64)
    return X, Y, M


def test_param_count_matches_header():
    assert R.n_params() == 9224


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    theta = rng.normal(scale=0.3, size=R.n_params())
    X, Y, M = _data(rng)
    L, g = R.loss_grad(theta, X, Y, M)
    assert abs(L - R.loss(theta, X, Y, M)) < 1e-12
    idx = rng.choice(theta.size, size=60, replace=False)
    # every layer's weights and biases are sampled
    idx = np.concatenate([idx, [0, 5 * 64, 5 * 64 + 64, R.n_params() - 1, R.n_params() - 9]])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (R.loss(tp, X, Y, M) - R.loss(tm, X, Y, M)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_identity_hidden_layers_give_affine_map():
    """Hidden layers = identity on nonnegative inputs, so the network is
    x -> x W1' W4 + b4 with W1' the first layer padded; closed form."""
    rng = np.random.default_rng(4)
    theta = np.zeros(R.n_params())
    layers = R.unpack(theta)
    W1 = np.abs(rng.normal(size=(5, 64)))
    layers[0][0][:] = W1
    layers[1][0][:] = np.eye(64)
    layers[2][0][:] = np.eye(64)
    W4 = rng.normal(size=(64, 8))
    layers[3][0][:] = W4
    layers[3][1][:] = rng.normal(size=8)
    X = rng.uniform(0, 1, size=(9, 5))
    P, _ = R.forward(theta, X)
    assert np.allclose(P, X @ W1 @ W4 + layers[3][1], rtol=1e-12, atol=1e-12)


def test_planted_mistakes_are_caught():
    """A gradient with the ReLU mask dropped, or the mean over rows instead
    of over valid entries, fails the finite-difference comparison."""
    rng = np.random.default_rng(5)
    theta = rng.normal(scale=0.5, size=R.n_params())
    X, Y, M = _data(rng)
    _, g = R.loss_grad(theta, X, Y, M)

    def fd(i, h=1e-6):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        return (R.loss(tp, X, Y, M) - R.loss(tm, X, Y, M)) / (2 * h)

    # wrong normaliser: rows instead of valid entries
    g_bad = g * M.sum() / M.shape[0]
    assert any(abs(fd(i) - g_bad[i]) > 1e-6 * max(1.0, abs(fd(i))) for i in range(0, 300, 7))
    # dropped ReLU mask: recompute backprop without it
    P, (acts, zs) = R.forward(theta, X)
    cnt = M.sum()
    dz = 2.0 * (P - Y) * M * M / cnt
    layers = R.unpack(theta)
    W4 = layers[3][0]
    dz3 = dz @ W4.T  # no mask
    gW3_bad = acts[2].T @ dz3
    off = R.n_params() - (64 * 8 + 8) - (64 * 64 + 64)
    gW3_true = g[off: off + 64 * 64].reshape(64, 64)
    assert not np.allclose(gW3_bad, gW3_true)
