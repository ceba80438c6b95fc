"""CPU restatement of the collision-probability MLP (gpuprm_amd/csrc/
surrogate.hip): forward, binary cross-entropy and backward in numpy.

TEST INFRASTRUCTURE ONLY (the checker of the GPU's forward/backward kernels).

Parameters use the kernels' flat order: W1 [256 x 16] (input columns beyond
ndof zero), b1 [256], W2 [256 x 256], b2, W3, b3, w4 [256], b4 [1].

mode "bf16": values are rounded to bfloat16 where the kernels round them
(inputs, weight copies, hidden activations, the SiLU derivatives kept for
the backward pass, activation gradients), products accumulate in float32.  mode "f64": plain float64 mathematics, used to pin the
backward formulas by central finite differences.
"""
from __future__ import annotations

import numpy as np

H = 256
IN_PAD = 16
SIZES = [H * IN_PAD, H, H * H, H, H * H, H, H, 1]


def split(p):
    out, o = [], 0
    for s in SIZES:
        out.append(p[o:o + s])
        o += s
    W1, b1, W2, b2, W3, b3, w4, b4 = out
    return W1.reshape(H, IN_PAD), b1, W2.reshape(H, H), b2, W3.reshape(H, H), b3, w4, b4


def bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _silu(a):
    return a / (1 + np.exp(-a))


def _dsilu(a):
    s = 1 / (1 + np.exp(-a))
    return s * (1 + a * (1 - s))


def inputs(Q):
    """Q (ndof, n) raw joint angles -> (n, 16) network input."""
    nd, n = Q.shape
    X = np.zeros((n, IN_PAD), np.float32)
    X[:, :nd] = (Q.T.astype(np.float32) * np.float32(0.3183098861837907)).astype(np.float32)
    return X


def forward(p, X, mode="bf16"):
    W1, b1, W2, b2, W3, b3, w4, b4 = split(p)
    if mode == "bf16":
        r, dt = bf16, np.float32
    else:
        r, dt = (lambda v: np.asarray(v, np.float64)), np.float64
    X = r(X).astype(dt)
    cache = {"X": X}
    h = X
    for l, (W, b) in enumerate(((W1, b1), (W2, b2), (W3, b3)), 1):
        Wr = r(W).astype(dt)
        a = (h @ Wr.T + b.astype(dt)).astype(dt)
        cache["a%d" % l] = a
        cache["W%d" % l] = Wr
        h = r(_silu(a)).astype(dt)
        cache["h%d" % l] = h
    z = h @ w4.astype(dt) + b4.astype(dt)[0]
    cache["z"] = z
    return 1 / (1 + np.exp(-z)), cache


def loss(p, X, y, mode="bf16"):
    pr, c = forward(p, X, mode)
    z = c["z"]
    return np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))


def gradients(p, X, y, mode="bf16"):
    """(flat gradient in parameter order, loss)."""
    W1, b1, W2, b2, W3, b3, w4, b4 = split(p)
    r = bf16 if mode == "bf16" else (lambda v: np.asarray(v, np.float64))
    dt = np.float32 if mode == "bf16" else np.float64
    pr, c = forward(p, X, mode)
    n = X.shape[0]
    y = np.asarray(y, dt)
    dz = ((pr - y) / n).astype(dt)
    gw4 = c["h3"].T @ dz
    gb4 = np.array([dz.sum()], dt)
    s3, s2, s1 = (r(_dsilu(c["a%d" % l])).astype(dt) for l in (3, 2, 1))
    dA3 = r(dz[:, None] * w4.astype(dt)[None, :] * s3).astype(dt)
    gW3 = dA3.T @ c["h2"]
    gb3 = dA3.sum(0)
    dA2 = r((dA3 @ c["W3"]) * s2).astype(dt)
    gW2 = dA2.T @ c["h1"]
    gb2 = dA2.sum(0)
    dA1 = r((dA2 @ c["W2"]) * s1).astype(dt)
    gW1 = dA1.T @ c["X"]
    gb1 = dA1.sum(0)
    z = c["z"]
    L = np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3, gw4, gb4]).astype(dt)
    return g, L


def auc(score, label):
    """Area under the ROC curve (rank statistic; ties get half credit)."""
    score = np.asarray(score, np.float64)
    label = np.asarray(label).astype(bool)
    npos, nneg = label.sum(), (~label).sum()
    if npos == 0 or nneg == 0:
        return float("nan")
    order = np.argsort(score, kind="mergesort")
    s = score[order]
    ranks = np.empty(len(s), np.float64)
    i = 0
    while i < len(s):
        j = i
        while j + 1 < len(s) and s[j + 1] == s[i]:
            j += 1
        ranks[order[i:j + 1]] = 0.5 * (i + j) + 1
        i = j + 1
    return float((ranks[label].sum() - npos * (npos + 1) / 2) / (npos * nneg))
