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


def _edge_inputs(rng, sizes, nd=7):
    """Edges of lengths 0..2 rad (testnum 1..~200 at dq = 0.01), a few of
    them shorter than dq (the testnum == 1 case), start points slightly
    beyond the joint bounds now and then."""
    import torch
    QS, QE = [], []
    for n in sizes:
        a = rng.uniform(-3.3, 3.3, size=(nd, n)).astype(np.float32)
        d = rng.normal(size=(nd, n))
        d /= np.linalg.norm(d, axis=0, keepdims=True) + 1e-12
        length = rng.uniform(0.0, 2.0, size=n)
        length[: max(1, n // 50)] = rng.uniform(0.0, 0.009, size=max(1, n // 50))  # shorter than dq
        QS.append(a)
        QE.append((a + d * length).astype(np.float32))
    qs = [torch.from_numpy(x).cuda() for x in QS]
    qe = [torch.from_numpy(x).cuda() for x in QE]
    r = [torch.full((n,), 99, dtype=torch.int32, device="cuda") for n in sizes]
    return QS, QE, qs, qe, r


def _single_edges(spaces, qs, qe):
    import torch
    out = []
    for sp, a, b in zip(spaces, qs, qe):
        r = torch.zeros(a.shape[1], dtype=torch.int32, device="cuda")
        if a.shape[1]:
            sp.indicator2_dev(a, b, r)
        out.append(r.cpu().numpy())
    return out


def test_batch_edges_match_single_scene_and_oracle(scenes):
    """Edge decisions (res != 0) of every scene equal the single-scene path's
    and the oracle's host loop; the sample total is the reference's."""
    import gpuprm_amd
    dirs, spaces = scenes
    rng = np.random.default_rng(11)
    sizes = [int(x) for x in rng.integers(300, 3000, NSCENES)]
    sizes[5] = 0  # an empty scene
    QS, QE, qs, qe, r = _edge_inputs(rng, sizes)
    b = gpuprm_amd.SceneBatch(spaces)
    done, total = b.indicator2_dev(qs, qe, r)
    want = _single_edges(spaces, qs, qe)
    tot = 0
    for s in range(NSCENES):
        got = r[s].cpu().numpy()
        assert np.array_equal(got != 0, want[s] != 0), s
        if sizes[s]:
            o = OracleScene(dirs[s])
            k = 60
            ores, _ = o.indicator2(QS[s][:, :k].copy(), QE[s][:, :k].copy(), nthreads=8)
            assert np.array_equal(got[:k] != 0, ores != 0), s
            t, n = o.edge_testnum(QS[s], QE[s])
            tot += int(n)
            assert done[s] > 0
        else:
            assert done[s] == 0
    assert total == tot


def test_batch_edges_graph_replays(scenes):
    """The captured steps are replayed with new edges in the same arrays
    (re-captured when a scene's largest testnum changes), with eager work
    between replays; every replay equals the single-scene results."""
    import gpuprm_amd
    import torch
    dirs, spaces = scenes
    rng = np.random.default_rng(12)
    sizes = [int(x) for x in rng.integers(200, 1500, NSCENES)]
    QS, QE, qs, qe, r = _edge_inputs(rng, sizes)
    b = gpuprm_amd.SceneBatch(spaces)
    for rep in range(4):
        QS2, QE2, _, _, _ = _edge_inputs(rng, sizes)
        for s in range(NSCENES):
            qs[s].copy_(torch.from_numpy(QS2[s]))
            qe[s].copy_(torch.from_numpy(QE2[s]))
            r[s].fill_(99)
        torch.cuda.synchronize()
        b.indicator2_dev(qs, qe, r, use_graph=True)
        want = _single_edges(spaces, qs, qe)
        for s in range(NSCENES):
            assert np.array_equal(r[s].cpu().numpy() != 0, want[s] != 0), (rep, s)


def test_batch_edges_every_shape(scenes):
    """Every bucket shape that fits gives the same edge decisions."""
    import gpuprm_amd
    dirs, spaces = scenes
    rng = np.random.default_rng(13)
    sizes = [int(x) for x in rng.integers(100, 800, NSCENES)]
    QS, QE, qs, qe, r = _edge_inputs(rng, sizes)
    want = _single_edges(spaces, qs, qe)
    b = gpuprm_amd.SceneBatch(spaces, tuning=None)
    used = [i for i, x in enumerate(b.bucket_info()) if x["scenes"]]
    for block, waves, ppu in ((256, 5, 16), (256, 4, 64), (1024, 4, 256)):
        for i in used:
            try:
                b.set_bucket_shape(i, block, waves, ppu)
            except gpuprm_amd.GprmError:
                pass  # the shape does not fit this bucket's scenes; it keeps its shape
        b.indicator2_dev(qs, qe, r, use_graph=True)
        for s in range(NSCENES):
            assert np.array_equal(r[s].cpu().numpy() != 0, want[s] != 0), (block, waves, ppu, s)


def test_batch_edges_several_steps_per_window(scenes):
    """Long edges on large scenes: a window's samples exceed a scene's
    per-step budget (2^19), so the window runs as several steps over ranges
    of the scene's edges; decisions still equal the single-scene path's."""
    import gpuprm_amd
    import torch
    dirs, spaces = scenes
    rng = np.random.default_rng(14)
    sub = [spaces[0], spaces[1], spaces[2]]
    sizes = [20000, 0, 15000]
    QS, QE = [], []
    for n in sizes:
        a = rng.uniform(-3.0, 3.0, size=(7, n)).astype(np.float32)
        d = rng.normal(size=(7, n))
        d /= np.linalg.norm(d, axis=0, keepdims=True) + 1e-12
        QS.append(a)
        QE.append((a + d * rng.uniform(1.5, 2.5, size=n)).astype(np.float32))
    qs = [torch.from_numpy(x).cuda() for x in QS]
    qe = [torch.from_numpy(x).cuda() for x in QE]
    r = [torch.full((n,), 99, dtype=torch.int32, device="cuda") for n in sizes]
    b = gpuprm_amd.SceneBatch(sub)
    done, total = b.indicator2_dev(qs, qe, r, use_graph=True)
    want = _single_edges(sub, qs, qe)
    for s in range(3):
        assert np.array_equal(r[s].cpu().numpy() != 0, want[s] != 0), s
    assert done[1] == 0 and done[0] > 0 and done[2] > 0
    assert total == sum(int(OracleScene(dirs[s]).edge_testnum(QS[s], QE[s])[1]) for s in (0, 2))
