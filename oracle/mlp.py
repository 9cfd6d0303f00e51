"""fp64 restatement of the dispatch cost model's network (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the device MLP's forward pass and
gradients (pmlg_mlp_forward / pmlg_mlp_loss_grad).  The network has no
counterpart in the reference (its dispatch is threshold-based,
flint-extras/src/nmod_poly_mat_extra/mul.c:52-96); this file follows the
parameter layout of include/pmlgpu.h: for each layer l, W_l (in x out,
row-major) then b_l, layers 5 -> 64 -> 64 -> 64 -> 8, ReLU on the hidden
layers, masked mean-squared-error loss.

Pinned in tests/test_oracle_mlp.py by central finite differences of the loss
(fp64), by a network whose hidden layers are the identity on nonnegative
inputs (a closed-form affine map), and by planted mistakes.
"""
import numpy as np

DIMS = (5, 64, 64, 64, 8)


def n_params(dims=DIMS):
    return sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))


def unpack(theta, dims=DIMS):
    """List of (W, b) views into the flat parameter vector."""
    theta = np.asarray(theta, dtype=np.float64)
    out, off = [], 0
    for i in range(len(dims) - 1):
        a, b = dims[i], dims[i + 1]
        W = theta[off: off + a * b].reshape(a, b)
        off += a * b
        bias = theta[off: off + b]
        off += b
        out.append((W, bias))
    assert off == theta.size
    return out


def forward(theta, X, dims=DIMS):
    """Output of the network and the cache of pre-activations."""
    X = np.asarray(X, dtype=np.float64)
    layers = unpack(theta, dims)
    acts, zs = [X], []
    h = X
    for li, (W, b) in enumerate(layers):
        z = h @ W + b
        zs.append(z)
        h = np.maximum(z, 0.0) if li + 1 < len(layers) else z
        acts.append(h)
    return h, (acts, zs)


def loss(theta, X, Y, M, dims=DIMS):
    P, _ = forward(theta, X, dims)
    M = np.asarray(M, dtype=np.float64)
    cnt = M.sum()
    d = (P - np.asarray(Y, dtype=np.float64)) * M
    return float((d * d).sum() / cnt) if cnt > 0 else 0.0


def loss_grad(theta, X, Y, M, dims=DIMS):
    """Loss and its gradient with respect to the flat parameters (backprop)."""
    P, (acts, zs) = forward(theta, X, dims)
    M = np.asarray(M, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    cnt = M.sum()
    inv = 1.0 / cnt if cnt > 0 else 0.0
    d = (P - Y) * M
    L = float((d * d).sum() * inv)
    dz = 2.0 * d * M * inv
    layers = unpack(theta, dims)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        grads[li] = (acts[li].T @ dz, dz.sum(axis=0))
        if li > 0:
            dz = (dz @ W.T) * (zs[li - 1] > 0)
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return L, flat
