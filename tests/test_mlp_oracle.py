"""The MLP oracle (oracle/mlp.py) pinned by mathematics: its float64
backward against central finite differences of its loss, and its AUC
against a brute-force pair count."""
import numpy as np

from oracle import mlp


def _params(rng, nd=7):
    p = []
    for s, fan in zip(mlp.SIZES, [nd, nd, 256, 256, 256, 256, 256, 256]):
        p.append(rng.uniform(-1, 1, s) / np.sqrt(fan))
    p = np.concatenate(p)
    W1 = p[:mlp.H * mlp.IN_PAD].reshape(mlp.H, mlp.IN_PAD)
    W1[:, nd:] = 0
    return p


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    Q = rng.uniform(-np.pi, np.pi, size=(7, 12))
    X = mlp.inputs(Q).astype(np.float64)
    y = (rng.uniform(size=12) < 0.5).astype(np.float64)
    g, L = mlp.gradients(p, X, y, mode="f64")
    assert abs(L - mlp.loss(p, X, y, mode="f64")) < 1e-12
    # every bias, a sample of every weight tensor
    offs = np.cumsum([0] + mlp.SIZES)
    idx = []
    for t in range(8):
        lo, hi = offs[t], offs[t + 1]
        cand = np.arange(lo, hi)
        if t == 0:
            cand = cand[(cand - lo) % mlp.IN_PAD < 7]
        idx.extend(rng.choice(cand, size=min(20, len(cand)), replace=False))
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss(p + e, X, y, "f64") - mlp.loss(p - e, X, y, "f64")) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0], np.float32)
    b = mlp.bf16(x)
    assert b[0] == 1.0 and b[1] == 1.0  # halfway rounds to even
    assert b[2] == np.float32(1.0078125)
    assert abs(b[3] + 3.140625) < 1e-7 and b[4] == 0


def test_auc_against_pair_count():
    rng = np.random.default_rng(1)
    s = rng.integers(0, 20, size=300).astype(float)
    lab = rng.uniform(size=300) < 0.4
    pos, neg = s[lab], s[~lab]
    want = ((pos[:, None] > neg[None, :]).sum() + 0.5 * (pos[:, None] == neg[None, :]).sum()) / (len(pos) * len(neg))
    assert abs(mlp.auc(s, lab) - want) < 1e-12
