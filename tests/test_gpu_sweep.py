"""Parity sweep over generated scenes: for every benchmark workload and six
scene seeds, point-check codes and edge decisions of the HIP path equal the
oracle's (random configurations a little beyond the joint bounds, edges of
0-1.2 rad)."""
import tempfile

import numpy as np
import pytest

from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu

NDOF = {1: 4, 2: 6, 3: 7, 5: 7}


@pytest.mark.parametrize("workload", [1, 2, 3, 5])
@pytest.mark.parametrize("seed", range(1, 7))
def test_generated_scene_parity(workload, seed):
    import torch

    import gpuprm_amd
    nd = NDOF[workload]
    d = tempfile.mkdtemp(prefix="gprm_sweep_")
    gpuprm_amd.generate_scene(workload, 5000 + 17 * seed, d)
    sp = gpuprm_amd.RobotConfigspace(d, 0.01, 1 << 16)
    sp.init(0, 1)
    o = OracleScene(d)
    rng = np.random.default_rng(seed * 31 + workload)
    n = 20000
    Q = rng.uniform(-3.3, 3.3, size=(nd, n)).astype(np.float32)
    r = torch.zeros(n, dtype=torch.int32, device="cuda")
    sp.indicator_dev(torch.from_numpy(Q).cuda(), r)
    assert np.array_equal(r.cpu().numpy(), o.indicator_batch(Q, nthreads=8))
    ne = 1500
    QS = rng.uniform(-3.1, 3.1, size=(nd, ne)).astype(np.float32)
    dirn = rng.normal(size=(nd, ne))
    dirn /= np.linalg.norm(dirn, axis=0, keepdims=True)
    QE = (QS + dirn * rng.uniform(0, 1.2, size=ne)).astype(np.float32)
    re = torch.zeros(ne, dtype=torch.int32, device="cuda")
    sp.indicator2_dev(torch.from_numpy(QS).cuda(), torch.from_numpy(QE).cuda(), re)
    want, _ = o.indicator2(QS, QE, nthreads=8)
    assert np.array_equal(re.cpu().numpy() != 0, want != 0)
