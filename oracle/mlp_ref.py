"""FP64 restatement of the learned hop proposal's network: 16 -> 64 -> 64 -> 1 with
ReLU, parameters packed as W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3, and
the binary cross-entropy (with logits) training loss.  No reference counterpart:
the reference's chargeHop proposes hops by a fixed rule (S/main26.c:1926-2390).

TEST INFRASTRUCTURE ONLY (the checker): tests/ and bench.py may use it to check the
HIP forward/backward; the product never imports it."""
import numpy as np

SIZES = [(64, 16), (64,), (64, 64), (64,), (64,), (1,)]
NPARAM = sum(int(np.prod(s)) for s in SIZES)


def unpack(w):
    out, o = [], 0
    for s in SIZES:
        k = int(np.prod(s))
        out.append(np.asarray(w[o:o + k], np.float64).reshape(s))
        o += k
    return out


def init(rng, scale=1.0):
    """He-style initial weights (float32)."""
    w = np.zeros(NPARAM, np.float32)
    W1 = rng.normal(0, scale * np.sqrt(2 / 16), (64, 16))
    W2 = rng.normal(0, scale * np.sqrt(2 / 64), (64, 64))
    W3 = rng.normal(0, scale * np.sqrt(1 / 64), 64)
    w[:] = np.concatenate([W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64), W3, [0.0]])
    return w


def forward(w, X):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    X = np.asarray(X, np.float64)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    return h2 @ W3 + b3[0], (X, a1, h1, a2, h2)


def loss_grad(w, X, y):
    """mean BCE-with-logits loss and its gradient w.r.t. the packed parameters."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    z, (X, a1, h1, a2, h2) = forward(w, X)
    y = np.asarray(y, np.float64)
    n = len(y)
    loss = np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    dz = (1 / (1 + np.exp(-z)) - y) / n
    gW3 = h2.T @ dz
    gb3 = np.array([dz.sum()])
    dh2 = np.outer(dz, W3) * (a2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3])
    return loss, g


def loss_only(w, X, y):
    z, _ = forward(w, X)
    y = np.asarray(y, np.float64)
    return np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
