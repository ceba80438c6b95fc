"""A scene whose image does not fit in LDS (generator workload 6: a 7-DOF arm
among 256 obstacles of 32 vertices, about 265 KB of image against 160 KB of
LDS): the separating-vector kernel reads it from the global image instead of
staging it.  Point codes and edge decisions equal the oracle's, under every
launch variant, and the planner runs on it."""
import numpy as np
import pytest

from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big(tmp_path_factory):
    import gpuprm_amd
    import scenes
    d = str(tmp_path_factory.mktemp("wl6"))
    gpuprm_amd.generate_scene(6, 7, d)
    s = scenes.read_dir(d)
    # vertex records alone (16 B each) and the adjacency (2 B per entry)
    # already exceed the 160 KB of a CU's LDS
    assert 16 * int(np.sum(s["n"])) + 2 * int(np.sum(s["m"])) > 160 * 1024
    sp = gpuprm_amd.RobotConfigspace(d, 0.01, 1 << 20)
    sp.init(0, 1)
    return d, sp, OracleScene(d)


def test_points_match_oracle_every_variant(big):
    import torch
    d, sp, o = big
    rng = np.random.default_rng(3)
    N = 4099
    Q = rng.uniform(-np.pi * 1.02, np.pi * 1.02, size=(7, N)).astype(np.float32)
    want = o.indicator_batch(Q, nthreads=8)
    assert 0.2 < (want != 0).mean() < 0.95
    q = torch.from_numpy(Q).cuda()
    for block, waves in ((256, 5), (256, 4), (1024, 4)):
        print("large scene points: shape (%d, %d)" % (block, waves), flush=True)
        sp.set_launch_shape(block, waves, 128)
        r = torch.zeros(N, dtype=torch.int32, device="cuda")
        sp.indicator_dev(q, r)
        assert np.array_equal(r.cpu().numpy(), want), (block, waves)
    # host entry points too (the small-call kernel needs the scene in LDS, so
    # single configurations take the main path)
    assert sp.indicator(np.ascontiguousarray(Q[:, 0])) == want[0]
    assert np.array_equal(sp.indicator_batch(np.ascontiguousarray(Q[:, :37])), want[:37])


def test_edges_match_oracle(big):
    import torch
    d, sp, o = big
    rng = np.random.default_rng(4)
    N = 400
    S = rng.uniform(-np.pi, np.pi, size=(7, N)).astype(np.float32)
    E = np.clip(S + rng.normal(0, 0.15, size=S.shape), -np.pi, np.pi).astype(np.float32)
    want, _ = o.indicator2(S, E, nthreads=8)
    assert 0 < (want != 0).mean() < 1
    print("large scene edges: %d" % N, flush=True)
    r = torch.zeros(N, dtype=torch.int32, device="cuda")
    sp.indicator2_dev(torch.from_numpy(S).cuda(), torch.from_numpy(E).cuda(), r)
    assert np.array_equal(r.cpu().numpy() != 0, want != 0)


def test_planner_runs(big):
    import gpuprm_amd
    d, sp, o = big
    rng = np.random.default_rng(5)
    free = []
    while len(free) < 2:
        q = rng.uniform(-np.pi, np.pi, 7).astype(np.float32)
        if o.indicator1(q) == 0:
            free.append(q)
    print("large scene planner", flush=True)
    r, path = sp.prm_solve(free[0], free[1], num=2048, D=0.5, max_nodes=1 << 18, max_steps=60, seed=1)
    print(r.as_dict())
    assert r.steps > 0 and r.nodes_left + r.nodes_right > 2
    if r.found:
        S, E = gpuprm_amd.canonical_edges(path[:-1], path[1:])
        res, _ = o.indicator2(S.T.copy(), E.T.copy(), nthreads=8)
        assert np.all(res == 0) and np.all(o.indicator_batch(path.T) == 0)


def test_batch_refuses_scene_beyond_lds(big):
    import gpuprm_amd
    d, sp, o = big
    print("large scene in a multi-scene batch", flush=True)
    with pytest.raises(gpuprm_amd.GprmError) as ei:
        gpuprm_amd.SceneBatch([sp], tuning=None)
    assert ei.value.code == -4 and "beyond LDS" in str(ei.value)
