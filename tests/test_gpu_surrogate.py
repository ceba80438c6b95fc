"""Surrogate MLP kernels (BF16 MFMA forward/backward, Adam) against the numpy
oracle (oracle/mlp.py, pinned by finite differences), and its training on
labels of the exact checker.

Two references: the oracle's "bf16" mode rounds where the kernels round
(inputs, weight copies, activations, SiLU derivatives, activation gradients)
and leaves only accumulation order and the hardware exponential/reciprocal
as differences; its "f64" mode is the exact mathematics, so the distance to
it is the whole BF16 error of the kernels.  Tolerances (stated per check):
probabilities within 5e-4 (bf16) and 1e-3 (f64) absolute; gradients per
tensor within 2e-3 (bf16) and 2e-2 (f64) relative L2; loss within 1e-4
(bf16) and 5e-4 (f64) absolute.  Measured on MI355X at these seeds: 2e-5 /
6e-5 (probabilities), 1.1e-4 / 4.3e-3 (gradients), 1e-6 / 5e-6 (loss)."""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _check(p, Q, codes, prob, g, L):
    X = mlp.inputs(Q)
    y = (codes != 0).astype(np.float32)
    offs = np.cumsum([0] + mlp.SIZES)
    for mode, tp, tg, tl in (("bf16", 5e-4, 2e-3, 1e-4), ("f64", 1e-3, 2e-2, 5e-4)):
        want_p, _ = mlp.forward(p, X, mode)
        want_g, want_L = mlp.gradients(p, X, y, mode)
        errs = [_rel(g[offs[t]:offs[t + 1]], want_g[offs[t]:offs[t + 1]]) for t in range(8)]
        dp = float(np.max(np.abs(prob - want_p)))
        print("%s: max |dp| %.2e, loss %.6f vs %.6f, gradient rel errors %s" %
              (mode, dp, L, want_L, " ".join("%.1e" % e for e in errs)))
        assert dp < tp, (mode, dp)
        assert abs(L - want_L) < tl * max(1.0, abs(want_L)), (mode, L, want_L)
        assert max(errs) < tg, (mode, errs)


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
    _check(p, Q, codes, prob.cpu().numpy(), *s.gradients(q, c))
    # round trip of the parameters
    s.set_params(p * 0.5)
    assert np.array_equal(s.get_params(), (p * 0.5).astype(np.float32))


def test_many_tiles_strided_input():
    """40,000 rows: every block of the persistent kernels takes several
    tiles, the weight gradients take 32 splits of K, the batch is a strided
    window of a wider SoA array (offset 50,021), and the last tile is ragged."""
    import torch
    import gpuprm_amd
    rng = np.random.default_rng(4)
    n, off = 40000, 50021
    s = gpuprm_amd.Surrogate(ndof=7, batch=n, seed=5)
    p = s.get_params()
    Qw = rng.uniform(-np.pi, np.pi, size=(7, off)).astype(np.float32)
    codes = np.where(rng.uniform(size=n) < 0.3, 3, 0).astype(np.int32)
    qw = torch.from_numpy(Qw).cuda()
    c = torch.from_numpy(codes).cuda()
    prob = torch.zeros(n, dtype=torch.float32, device="cuda")
    s.predict(qw, prob, offset=off)
    g, L = s.gradients(qw, c, offset=off)
    _check(p, np.ascontiguousarray(Qw[:, :n]), codes, prob.cpu().numpy(), g, L)
    # the same rows, contiguous: identical results (same kernels, same sums)
    q2 = torch.from_numpy(np.ascontiguousarray(Qw[:, :n])).cuda()
    prob2 = torch.zeros(n, dtype=torch.float32, device="cuda")
    s.predict(q2, prob2)
    assert torch.equal(prob, prob2)
    g2, L2 = s.gradients(q2, c)
    assert np.array_equal(g, g2) and L == L2


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
