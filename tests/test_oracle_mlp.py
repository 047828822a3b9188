"""CPU: the oracle's restatement of the rank predictor network, pinned by
closed forms and by finite differences (no GPU)."""
import math

import numpy as np

from oracle import lib as orc

NP = orc.MLP_NPARAMS
W1, B1, W2, B2, W3, B3 = 0, 512, 576, 4672, 4736, 4800


def gelu(x):
    return 0.5 * x * (1.0 + math.erf(x / math.sqrt(2.0)))


def test_forward_closed_form():
    """Only the path x0 -> unit 0 -> unit 0 -> output is switched on:
    out = 2 gelu(gelu(3 x0 + 0.5) - 0.25) + 0.1."""
    p = np.zeros(NP)
    p[W1 + 0 * 8 + 0] = 3.0
    p[B1 + 0] = 0.5
    p[W2 + 0 * 64 + 0] = 1.0
    p[B2 + 0] = -0.25
    p[W3 + 0] = 2.0
    p[B3] = 0.1
    feat = np.array([[v] + [0.0] * 7 for v in (-2.0, -0.3, 0.0, 0.7, 1.9)], np.float32)
    out = orc.mlp_predict(p, np.zeros(8), np.ones(8), feat)
    ref = [2 * gelu(gelu(3 * float(f[0]) + 0.5) - 0.25) + 0.1 for f in feat]
    # units 1..63 see z = 0 and gelu(0) = 0, so they add nothing
    assert np.allclose(out, ref, rtol=1e-14, atol=1e-14)
    assert gelu(1.0) == 0.5 * (1 + math.erf(1 / math.sqrt(2)))  # GELU(x) = x Phi(x)


def test_normalisation_is_applied():
    rng = np.random.default_rng(0)
    p = rng.uniform(-0.3, 0.3, NP)
    feat = rng.normal(size=(7, 8)).astype(np.float32)
    mean = rng.normal(size=8)
    std = rng.uniform(0.5, 2.0, 8)
    a = orc.mlp_predict(p, mean, std, feat)
    b = orc.mlp_predict(p, np.zeros(8), np.ones(8), ((feat - mean) / std).astype(np.float32))
    assert np.allclose(a, b, rtol=1e-6, atol=1e-6)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.uniform(-0.4, 0.4, NP)
    feat = rng.normal(size=(13, 8)).astype(np.float32)
    tgt = rng.uniform(0, 4, 13).astype(np.float32)
    mean, std = np.zeros(8), np.ones(8)
    loss, g = orc.mlp_loss_grad(p, mean, std, feat, tgt)
    out = orc.mlp_predict(p, mean, std, feat)
    assert abs(loss - np.mean(0.5 * (out - tgt) ** 2)) < 1e-14
    idx = np.concatenate([rng.choice(np.arange(a, b), 8, replace=False) for a, b in
                          ((W1, B1), (B1, W2), (W2, B2), (B2, W3), (W3, B3))] + [[B3]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (orc.mlp_loss_grad(pp, mean, std, feat, tgt)[0] - orc.mlp_loss_grad(pm, mean, std, feat, tgt)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
