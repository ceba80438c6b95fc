"""Multi-scene batches (gprm_batch_*): every scene's codes equal the
single-scene path's and the oracle's, also when the launch sequence is
replayed from a captured hipGraph and after the autotuner changed the
shapes."""
import numpy as np
import pytest

from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu

NSCENES = 12


@pytest.fixture(scope="module")
def scenes(tmp_path_factory):
    import gpuprm_amd
    root = tmp_path_factory.mktemp("batch")
    dirs, spaces = [], []
    for s in range(NSCENES):
        d = str(root / ("s%d" % s))
        gpuprm_amd.generate_scene(5, 100 + s, d)
        sp = gpuprm_amd.RobotConfigspace(d, 0.01, 1024)
        sp.init(0, 1)
        dirs.append(d)
        spaces.append(sp)
    return dirs, spaces


def _inputs(rng, sizes):
    import torch
    Q = [rng.uniform(-3.3, 3.3, size=(7, n)).astype(np.float32) for n in sizes]  # a few out of bounds
    q = [torch.from_numpy(x).cuda() for x in Q]
    r = [torch.full((n,), 99, dtype=torch.int32, device="cuda") for n in sizes]
    return Q, q, r


def _single(spaces, q):
    import torch
    out = []
    for sp, x in zip(spaces, q):
        r = torch.zeros(x.shape[1], dtype=torch.int32, device="cuda")
        if x.shape[1]:
            sp.indicator_dev(x, r)
        out.append(r.cpu().numpy())
    return out


def test_batch_matches_single_scene_and_oracle(scenes):
    import gpuprm_amd
    dirs, spaces = scenes
    rng = np.random.default_rng(1)
    sizes = [int(x) for x in rng.integers(1000, 6000, NSCENES)]
    sizes[3] = 0  # an empty scene in the middle
    Q, q, r = _inputs(rng, sizes)
    b = gpuprm_amd.SceneBatch(spaces)
    info = b.bucket_info()
    assert sum(x["scenes"] for x in info) == NSCENES
    done = b.indicator_dev(q, r)
    want = _single(spaces, q)
    for s in range(NSCENES):
        got = r[s].cpu().numpy()
        assert np.array_equal(got, want[s]), s
        if sizes[s]:
            o = OracleScene(dirs[s])
            assert np.array_equal(got[:300], o.indicator_batch(Q[s][:, :300].copy(), nthreads=8)), s
            assert done[s] > 0
        else:
            assert done[s] == 0


def test_batch_graph_replays(scenes):
    """The captured sequence is replayed with new contents in the same arrays;
    eager work (torch fills, single-scene checks) runs between replays, and
    every replay is checked against the eager single-scene result."""
    import gpuprm_amd
    import torch
    dirs, spaces = scenes
    rng = np.random.default_rng(2)
    sizes = [int(x) for x in rng.integers(500, 3000, NSCENES)]
    Q, q, r = _inputs(rng, sizes)
    b = gpuprm_amd.SceneBatch(spaces)
    for rep in range(4):
        for s in range(NSCENES):
            q[s].copy_(torch.from_numpy(rng.uniform(-3.3, 3.3, size=(7, sizes[s])).astype(np.float32)))
            r[s].fill_(99)
        torch.cuda.synchronize()
        b.indicator_dev(q, r, use_graph=True)
        want = _single(spaces, q)
        for s in range(NSCENES):
            assert np.array_equal(r[s].cpu().numpy(), want[s]), (rep, s)


def test_batch_autotune_and_shapes(scenes):
    import gpuprm_amd
    dirs, spaces = scenes
    rng = np.random.default_rng(3)
    sizes = [2000] * NSCENES
    Q, q, r = _inputs(rng, sizes)
    b = gpuprm_amd.SceneBatch(spaces)
    table = b.autotune(q, r, reps=1)
    assert table and all(row["pairs_per_unit"] in (32, 64, 128, 256) for row in table)
    for row in table:
        assert (row["block"], row["waves"]) in ((256, 5), (256, 4), (1024, 4))
    b2 = gpuprm_amd.SceneBatch(spaces)
    b2.apply_tuning(table)
    assert b2.tuning() == table
    b2.indicator_dev(q, r, use_graph=True)
    want = _single(spaces, q)
    for s in range(NSCENES):
        assert np.array_equal(r[s].cpu().numpy(), want[s]), s
    used = [i for i, x in enumerate(b2.bucket_info()) if x["scenes"]]
    with pytest.raises(gpuprm_amd.GprmError):
        b2.set_bucket_shape(used[0], 512, 4, 32)
