"""Plain numpy restatement of the learned guesses' perceptron (tanh or SiLU
hidden layers, linear output, mean-squared-error loss), used by the tests to
check the HIP forward and backward passes.  Test infrastructure only."""
import numpy as np


def unflatten(flat, dims):
    layers, k = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = flat[k:k + a * b].reshape(a, b)
        k += a * b
        layers.append((W, flat[k:k + b]))
        k += b
    return layers


def _act(Z, activation):
    return np.tanh(Z) if activation == "tanh" else Z / (1.0 + np.exp(-Z))


def _dact(Z, activation):
    if activation == "tanh":
        return 1.0 - np.tanh(Z) ** 2
    s = 1.0 / (1.0 + np.exp(-Z))
    return s * (1.0 + Z * (1.0 - s))


def forward(flat, dims, X, activation="tanh", pre=None):
    H = [X]
    layers = unflatten(flat, dims)
    for i, (W, b) in enumerate(layers):
        Z = H[-1] @ W + b
        if pre is not None:
            pre.append(Z)
        H.append(_act(Z, activation) if i < len(layers) - 1 else Z)
    return H


def loss_grad(flat, dims, X, Y, activation="tanh"):
    """Loss = mean((f(X) - Y)^2) and its gradient (flat, same layout)."""
    Zs = []
    H = forward(flat, dims, X, activation, Zs)
    layers = unflatten(flat, dims)
    n = Y.size
    dZ = 2.0 * (H[-1] - Y) / n
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        grads.append((H[i].T @ dZ, dZ.sum(axis=0)))
        if i > 0:
            dZ = (dZ @ W.T) * _dact(Zs[i - 1], activation)
    grads.reverse()
    g = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return float(np.mean((H[-1] - Y) ** 2)), g
