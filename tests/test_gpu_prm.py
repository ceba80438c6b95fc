"""Single-process PRM (gprm_prm_solve) on the GPU: the path it returns is
re-verified node by node and edge by edge by the oracle (the exact checker
of record), starts at qstart, ends at qgoal, and every step is shorter than
the connection radius."""
import numpy as np
import pytest

from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu


def _space(d):
    import gpuprm_amd
    s = gpuprm_amd.RobotConfigspace(d, 0.01, 1 << 20)
    s.init(0, 1)
    return s


def _verify(o, path, qs, qg, D):
    assert np.array_equal(path[0], qs) and np.array_equal(path[-1], qg)
    assert np.all(o.indicator_batch(path.T) == 0)
    steps = np.linalg.norm(np.diff(path.astype(np.float64), axis=0), axis=1)
    assert np.all(steps < D * (1 + 1e-6))
    res, _ = o.indicator2(path[:-1].T.copy(), path[1:].T.copy(), nthreads=8)
    assert np.all(res == 0), np.flatnonzero(res)


def test_prm_reference_task_config1(golden_dir):
    # prm/src/main.cpp:128-133: start (0,0,0,0), goal (3,3,3,3), D = 1.0
    z, d = golden_dir("config1")
    s = _space(d)
    o = OracleScene(d)
    qs = np.zeros(4, np.float32)
    qg = np.full(4, 3.0, np.float32)
    r, path = s.prm_solve(qs, qg, num=1024, D=1.0, max_nodes=1 << 17, max_steps=300, seed=3)
    print(r.as_dict())
    assert r.found == 1 and path is not None
    _verify(o, path, qs, qg, 1.0)


def test_prm_7dof_workload(workload_dir):
    d = workload_dir(3)
    s = _space(d)
    o = OracleScene(d)
    rng = np.random.default_rng(5)
    free = []
    while len(free) < 2:
        q = rng.uniform(-np.pi, np.pi, 7).astype(np.float32)
        if o.indicator1(q) == 0:
            free.append(q)
    r, path = s.prm_solve(free[0], free[1], num=2048, D=0.5, max_nodes=1 << 18, max_steps=400, seed=9)
    print(r.as_dict())
    assert r.found == 1
    _verify(o, path, free[0], free[1], 0.5)
