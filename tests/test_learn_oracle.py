"""The CPU oracle of the learned-proposal network, pinned by finite differences: its
hand-derived InfoNCE gradient must match central differences of its own loss."""
import numpy as np

import mlp_oracle as M


def _pairs(n, seed):
    rng = np.random.RandomState(seed)
    x = np.zeros((n, M.K_IN))
    x[:, :120] = rng.uniform(0, 1, size=(n, 120))
    x[:, 3:120:4] = rng.randint(0, 2, size=(n, 30))   # presence flags
    xk = x + rng.normal(0, 0.05, size=x.shape)
    xk[:, 120:] = 0
    return x, xk


def test_infonce_gradient_matches_finite_differences():
    p = M.he_uniform(3)
    xq, xk = _pairs(6, 4)
    tau = 0.2
    loss, g = M.infonce(p, xq, xk, tau)
    assert np.isfinite(loss) and g.shape == (M.PARAM_COUNT,)
    rng = np.random.RandomState(5)
    # every bias entry of the last layer plus random weights of each block
    sizes = [int(np.prod(s)) for _, s in M.SHAPES]
    starts = np.cumsum([0] + sizes)
    idx = list(range(starts[5], starts[6]))
    for b in range(6):
        idx += list(rng.randint(starts[b], starts[b + 1], size=12))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.infonce(pp, xq, xk, tau)[0] - M.infonce(pm, xq, xk, tau)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_forward_invariants():
    p = M.he_uniform(1)
    xq, xk = _pairs(5, 2)
    z, _ = M.forward(p, xq)
    assert np.allclose((z * z).sum(1), 1.0)
    # padded inputs (rows 120..127 of W1) carry no gradient when the pads are zero
    _, g = M.infonce(p, xq, xk, 0.1)
    w1 = g[:M.K_IN * M.K_H1].reshape(M.K_IN, M.K_H1)
    assert np.abs(w1[120:]).max() == 0.0


def test_bf16_rounding_helpers():
    v = np.array([1.0, -2.5, 3.14159, 1e-3, 65504.0], np.float32)
    b = M.bf16_bits(v)
    assert np.allclose(M.from_bits(b), M.bf16_round(v))
    assert M.from_bits(M.bf16_bits(np.float32(1.0 + 2 ** -9)))[()] == 1.0       # ties to even
    assert M.from_bits(M.bf16_bits(np.float32(1.0 + 3 * 2 ** -9)))[()] == 1.0 + 2 ** -7
