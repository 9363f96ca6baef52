"""FP64 numpy restatement of the warm-start classifier (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, as the checker.  The classifier has no counterpart in the reference: it
only proposes the starting set of Subhalo_t::Unbind (src/subhalo_unbind.cpp:463),
which the exact iteration then verifies.

Network: x[8] -> ReLU(W1 x + b1)[64] -> ReLU(W2 h1 + b2)[64] -> z = w3.h2 + b3,
loss = mean binary cross-entropy with logits, softplus(z) - y z.
Flat parameter layout: W1[64][8] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3.
Pinned by central finite differences of the loss (tests/test_learned_cpu.py).
"""
import numpy as np

NIN, NH = 8, 64
LAYOUT = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)), ("b3", (1,))]
NPAR = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(flat):
    out, k = {}, 0
    for name, shape in LAYOUT:
        size = int(np.prod(shape))
        out[name] = np.asarray(flat[k:k + size], dtype=np.float64).reshape(shape)
        k += size
    return out


def bf16_round(w):
    """Round-to-nearest-even of float32 values to bfloat16, returned as float32."""
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def forward(flat, X):
    p = unpack(flat)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0.0)
    z = h2 @ p["w3"] + p["b3"][0]
    return z, (X, a1, h1, a2, h2)


def loss(flat, X, y):
    z, _ = forward(flat, X)
    y = np.asarray(y, dtype=np.float64)
    return float(np.mean(np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z))) - y * z))


def gradient(flat, X, y):
    """(gradient of the mean loss in the flat layout, mean loss)."""
    z, (X, a1, h1, a2, h2) = forward(flat, X)
    p = unpack(flat)
    y = np.asarray(y, dtype=np.float64)
    n = len(X)
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / n
    g = {}
    g["w3"] = h2.T @ dz
    g["b3"] = np.array([dz.sum()])
    dh2 = np.outer(dz, p["w3"]) * (a2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ p["W2"]) * (a1 > 0)
    g["W1"] = dh1.T @ X
    g["b1"] = dh1.sum(0)
    flatg = np.concatenate([g[name].ravel() for name, _ in LAYOUT])
    lval = float(np.mean(np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z))) - y * z))
    return flatg, lval
