"""CPU oracle of the learned-proposal network (test infrastructure only: nothing in the
shipped library imports it; only tests/ use it as the checker).

The network has no counterpart in the reference; it is the proposal model of the batched
engine: x[128] -> relu(x W1 + b1)[256] -> relu(h1 W2 + b2)[256] -> h2 W3 + b3 [64] ->
z = e / |e|, weights [in x out] row-major in one flat vector (W1, b1, W2, b2, W3, b3).
Training loss: InfoNCE over a batch of n pairs, logits L = z_q z_k^T / tau, positives on
the diagonal, loss = mean_i (logsumexp_j L_ij - L_ii); both towers share the weights.

Everything is float64 numpy.  The backward is derived by hand; tests/test_learn_oracle.py
pins it against central finite differences of the loss, so the GPU gradients are checked
against a gradient that is itself checked.
"""
import numpy as np

K_IN, K_H1, K_H2, K_EMB = 128, 256, 256, 64
SHAPES = [("W1", (K_IN, K_H1)), ("b1", (K_H1,)), ("W2", (K_H1, K_H2)), ("b2", (K_H2,)),
          ("W3", (K_H2, K_EMB)), ("b3", (K_EMB,))]
PARAM_COUNT = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16_round(a):
    """Rounds float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward(p, x):
    """Returns (z, cache) for inputs x [n x 128] (float64)."""
    w = unpack(p)
    a1 = x @ w["W1"] + w["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ w["W2"] + w["b2"]
    h2 = np.maximum(a2, 0.0)
    e = h2 @ w["W3"] + w["b3"]
    nrm = np.sqrt(np.maximum((e * e).sum(1, keepdims=True), 1e-12))
    z = e / nrm
    return z, (x, a1, h1, a2, h2, e, nrm, z)


def backward(p, cache, dz):
    """Parameter gradient (flat) and input gradient for one tower given dL/dz."""
    w = unpack(p)
    x, a1, h1, a2, h2, e, nrm, z = cache
    de = (dz - z * (z * dz).sum(1, keepdims=True)) / nrm
    g = {"W3": h2.T @ de, "b3": de.sum(0)}
    dh2 = de @ w["W3"].T
    da2 = dh2 * (a2 > 0)
    g["W2"] = h1.T @ da2
    g["b2"] = da2.sum(0)
    dh1 = da2 @ w["W2"].T
    da1 = dh1 * (a1 > 0)
    g["W1"] = x.T @ da1
    g["b1"] = da1.sum(0)
    return pack(g)


def infonce(p, xq, xk, tau):
    """(loss, flat gradient) of the InfoNCE objective over pairs (xq[i], xk[i])."""
    zq, cq = forward(p, xq)
    zk, ck = forward(p, xk)
    n = xq.shape[0]
    L = zq @ zk.T / tau
    m = L.max(1, keepdims=True)
    lse = m[:, 0] + np.log(np.exp(L - m).sum(1))
    loss = float(np.mean(lse - np.diag(L)))
    P = np.exp(L - lse[:, None])
    dL = (P - np.eye(n)) / n
    dzq = dL @ zk / tau
    dzk = dL.T @ zq / tau
    return loss, backward(p, cq, dzq) + backward(p, ck, dzk)


def he_uniform(seed):
    """A deterministic initialisation of the same shape family as the library's (not the
    same values: the library seeds its own generator)."""
    rng = np.random.RandomState(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            fan = 120 if name == "W1" else shp[0]
            a = np.sqrt(6.0 / fan)
            v = rng.uniform(-a, a, size=shp)
            if name == "W1":
                v[120:] = 0.0
            d[name] = v
        else:
            d[name] = np.zeros(shp)
    return pack(d)
