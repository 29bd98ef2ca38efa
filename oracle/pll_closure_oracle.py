"""CPU restatement of the learned closure's perceptron (csrc/k_closure.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernels, never by the product path.

10 -> 64 -> 64 -> 1, tanh hidden layers, softplus output, loss = mean of
(softplus(o) - y)^2.  With bf16=True the inputs, weights, activations h1 and h2
and the back-propagated deltas are rounded to BF16 (nearest, ties to even) at
the points where the GPU rounds them before an MFMA product; sums are FP64.
With bf16=False everything is FP64: the form the finite-difference check pins.
"""
import numpy as np

N_IN, N_IN_PAD, N_H = 10, 16, 64
O_W1 = 0
O_B1 = O_W1 + N_IN_PAD * N_H
O_W2 = O_B1 + N_H
O_B2 = O_W2 + N_H * N_H
O_W3 = O_B2 + N_H
O_B3 = O_W3 + N_H
N_PARAM = O_B3 + 1


def bf16(a):
    """Round float32 values to BF16 (nearest, ties to even); returns float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) >> 16, u >> 16)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[O_W1:O_B1].reshape(N_IN_PAD, N_H).copy()
    W1[N_IN:] = 0.0
    return dict(W1=W1, b1=p[O_B1:O_W2], W2=p[O_W2:O_B2].reshape(N_H, N_H), b2=p[O_B2:O_W3],
                W3=p[O_W3:O_B3], b3=p[O_B3])


def _r(a, on):
    return bf16(a).astype(np.float64) if on else np.asarray(a, dtype=np.float64)


def softplus(o):
    return np.where(o > 20, o, np.log1p(np.exp(np.minimum(o, 20))))


def forward(p, x, bf16_round=True):
    w = unpack(p)
    xb = np.zeros((x.shape[0], N_IN_PAD))
    xb[:, :N_IN] = _r(x, bf16_round)
    W1, W2, W3 = (_r(w[k], bf16_round) for k in ("W1", "W2", "W3"))
    h1 = np.tanh(xb @ W1 + w["b1"])
    if bf16_round:
        h1 = h1.astype(np.float32).astype(np.float64)
    h1b = _r(h1, bf16_round)
    h2 = np.tanh(h1b @ W2 + w["b2"])
    if bf16_round:
        h2 = h2.astype(np.float32).astype(np.float64)
    h2b = _r(h2, bf16_round)
    o = h2b @ W3 + w["b3"]
    return dict(xb=xb, W1=W1, W2=W2, W3=W3, h1=h1, h1b=h1b, h2=h2, h2b=h2b, o=o)


def predict(p, x, bf16_round=True):
    return softplus(forward(p, x, bf16_round)["o"])


def loss(p, x, y, bf16_round=True):
    f = forward(p, x, bf16_round)
    return float(np.mean((softplus(f["o"]) - y) ** 2))


def gradient(p, x, y, bf16_round=True):
    """(gradient over the parameter vector, loss), the back propagation of
    csrc/k_closure.hip step by step."""
    f = forward(p, x, bf16_round)
    n = x.shape[0]
    o = f["o"]
    d = softplus(o) - y
    g = 2.0 * d / n / (1.0 + np.exp(-o))
    grad = np.zeros(N_PARAM)
    grad[O_W3:O_B3] = f["h2b"].T @ g
    grad[O_B3] = g.sum()
    d2 = g[:, None] * f["W3"][None, :] * (1.0 - f["h2"] ** 2)
    grad[O_B2:O_W3] = d2.sum(axis=0)
    d2b = _r(d2, bf16_round)
    grad[O_W2:O_B2] = (f["h1b"].T @ d2b).ravel()
    d1 = (d2b @ f["W2"].T) * (1.0 - f["h1"] ** 2)
    grad[O_B1:O_W2] = d1.sum(axis=0)
    d1b = _r(d1, bf16_round)
    gW1 = f["xb"].T @ d1b
    gW1[N_IN:] = 0.0
    grad[O_W1:O_B1] = gW1.ravel()
    return grad, float(np.mean(d ** 2))
