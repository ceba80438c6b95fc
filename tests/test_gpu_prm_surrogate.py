"""What the trained surrogate changes in the planner (gprm_prm_solve with a
model): on the 6-DOF edge-checking scene, a model trained by the library's
own kernels on exact labels ranks held-out configurations well (AUC), and
the planner that uses it for biased sampling and edge ranking checks fewer
edge samples exactly than the plain planner while every path of both is
re-verified by the oracle.  Times are printed, not asserted."""
import os

import numpy as np
import pytest

from oracle import mlp
from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu


def _ok(o, path):
    import gpuprm_amd
    assert path is not None
    S, E = gpuprm_amd.canonical_edges(path[:-1], path[1:])
    res, _ = o.indicator2(S.T.copy(), E.T.copy(), nthreads=8)
    return bool(np.all(res == 0) and np.all(o.indicator_batch(path.T) == 0))


def test_surrogate_saves_exact_checks_at_equal_validity(tmp_path):
    import gpuprm_amd
    import torch
    d = str(tmp_path / "wl2")
    gpuprm_amd.generate_scene(2, 2024, d)
    sp = gpuprm_amd.RobotConfigspace(d, 0.01, 1 << 20)
    sp.init(0, 1)
    o = OracleScene(d)
    rng = np.random.default_rng(17)
    N, B = 1 << 20, 8192
    Q = torch.from_numpy(rng.uniform(-np.pi, np.pi, size=(6, N)).astype(np.float32)).cuda()
    codes = torch.zeros(N, dtype=torch.int32, device="cuda")
    sp.indicator_dev(Q, codes)
    model = gpuprm_amd.Surrogate(ndof=6, batch=B, seed=5)
    ntr = N - (1 << 15)
    for it in range(3 * (ntr // B)):
        k = it % (ntr // B)
        model.train_step(Q[:, k * B:(k + 1) * B], codes[k * B:(k + 1) * B], lr=2e-3, offset=N)
    pt = torch.zeros(N - ntr, dtype=torch.float32, device="cuda")
    model.predict(Q[:, ntr:].contiguous(), pt)
    auc = mlp.auc(pt.cpu().numpy(), codes[ntr:].cpu().numpy() != 0)
    print("AUC %.4f" % auc)
    assert auc > 0.85
    tasks, r2 = [], np.random.default_rng(5)
    while len(tasks) < 3:
        a = r2.uniform(-np.pi, np.pi, 6).astype(np.float32)
        b = r2.uniform(-np.pi, np.pi, 6).astype(np.float32)
        if o.indicator1(a) == 0 and o.indicator1(b) == 0 and np.linalg.norm(a - b) > 4.0:
            tasks.append((a, b))
    tot = {"plain": 0, "surrogate": 0}
    for i, (a, b) in enumerate(tasks):
        rp, pp = sp.prm_solve(a, b, num=1024, D=1.0, max_nodes=1 << 18, max_steps=300, seed=i)
        rs, ps = sp.prm_solve(a, b, num=1024, D=1.0, max_nodes=1 << 18, max_steps=300, seed=i, surrogate=model,
                              order_edges=True, bias_sampling=True, edge_budget=4)
        print("task %d: plain %.1f ms %d samples, surrogate %.1f ms %d samples (model %.1f ms)"
              % (i, rp.ms_total, rp.edge_samples, rs.ms_total, rs.edge_samples, rs.ms_model))
        assert rp.found == 1 and rs.found == 1
        assert _ok(o, pp) and _ok(o, ps)
        assert rs.model_predictions > 0 and rs.edges_skipped > 0
        tot["plain"] += rp.edge_samples
        tot["surrogate"] += rs.edge_samples
    assert tot["surrogate"] < 0.9 * tot["plain"], tot
