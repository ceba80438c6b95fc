"""Surrogate MLP kernels (BF16 MFMA forward/backward, Adam) against the numpy
oracle (oracle/mlp.py, pinned by finite differences), and its training on
labels of the exact checker."""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_and_gradients_match_oracle():
    import torch
    import gpuprm_amd
    rng = np.random.default_rng(0)
    s = gpuprm_amd.Surrogate(ndof=7, batch=1024, seed=1)
    p = s.get_params()
    assert p.size == sum(mlp.SIZES)
    W1 = p[:mlp.H * mlp.IN_PAD].reshape(mlp.H, mlp.IN_PAD)
    assert np.all(W1[:, 7:] == 0)
    n = 777  # not a multiple of the 64-row tile
    Q = rng.uniform(-np.pi, np.pi, size=(7, n)).astype(np.float32)
    codes = np.where(rng.uniform(size=n) < 0.4, -1, 0).astype(np.int32)
    q, c = torch.from_numpy(Q).cuda(), torch.from_numpy(codes).cuda()
    prob = torch.zeros(n, dtype=torch.float32, device="cuda")
    s.predict(q, prob)
    X = mlp.inputs(Q)
    want_p, _ = mlp.forward(p, X, "bf16")
    assert np.max(np.abs(prob.cpu().numpy() - want_p)) < 2e-2
    g, L = s.gradients(q, c)
    want_g, want_L = mlp.gradients(p, X, (codes != 0).astype(np.float32), "bf16")
    assert abs(L - want_L) < 2e-3 * max(1.0, abs(want_L))
    offs = np.cumsum([0] + mlp.SIZES)
    for t in range(8):
        a, b = g[offs[t]:offs[t + 1]], want_g[offs[t]:offs[t + 1]]
        assert _rel(a, b) < 3e-2, (t, _rel(a, b))
    # round trip of the parameters
    s.set_params(p * 0.5)
    assert np.array_equal(s.get_params(), (p * 0.5).astype(np.float32))


def test_training_on_exact_labels(workload_dir):
    import torch
    import gpuprm_amd
    d = workload_dir(3)
    space = gpuprm_amd.RobotConfigspace(d, 0.01, 1 << 20)
    space.init(0, 1)
    rng = np.random.default_rng(7)
    n_train, n_test = 1 << 19, 1 << 15
    Q = torch.from_numpy(rng.uniform(-np.pi, np.pi, size=(7, n_train + n_test)).astype(np.float32)).cuda()
    codes = torch.zeros(n_train + n_test, dtype=torch.int32, device="cuda")
    space.indicator_dev(Q, codes)
    torch.cuda.synchronize()
    lab = (codes.cpu().numpy() != 0)
    s = gpuprm_amd.Surrogate(ndof=7, batch=4096, seed=3)
    B = 4096
    losses = []
    perm = torch.from_numpy(rng.permutation(n_train)).cuda()
    for it in range(1200):
        idx = perm[(it * B) % n_train:(it * B) % n_train + B]
        qb = Q[:, idx].contiguous()
        cb = codes[idx].contiguous()
        losses.append(s.train_step(qb, cb, lr=2e-3))
    qt = Q[:, n_train:].contiguous()
    pt = torch.zeros(n_test, dtype=torch.float32, device="cuda")
    s.predict(qt, pt)
    a = mlp.auc(pt.cpu().numpy(), lab[n_train:])
    print("loss %.4f -> %.4f, test AUC %.4f, positive rate %.3f" % (np.mean(losses[:20]), np.mean(losses[-50:]), a,
                                                                 lab.mean()))
    assert np.mean(losses[-50:]) < 0.8 * np.mean(losses[:20])
    assert a > 0.85
