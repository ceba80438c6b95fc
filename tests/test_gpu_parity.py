"""HIP path (libgprm.so through the C ABI) against the oracle and the golden
vectors.  Point checks: codes bit-exact.  Edge checks: decisions (code != 0)
exact; the blocked code of an edge is whichever colliding sample the early
exit finds, so only its non-zero-ness is compared."""
import numpy as np
import pytest

from conftest import GOLDEN_NAMES
from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu


def _space(d, dq=0.01):
    import gpuprm_amd
    s = gpuprm_amd.RobotConfigspace(d, dq, 2048)
    s.init(0, 1)
    return s


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_points_golden(golden_dir, name):
    z, d = golden_dir(name)
    s = _space(d, float(z["dq"]))
    got = s.indicator_batch(z["Q"])
    bad = np.flatnonzero(got != z["codes"])
    assert bad.size == 0, (name, bad[:8], got[bad[:8]], z["codes"][bad[:8]])


def test_point_single_and_known_answers(golden_dir):
    z, d = golden_dir("config1")
    s = _space(d)
    assert s.indicator(np.zeros(4, np.float32)) == 0
    assert s.indicator(np.full(4, 3.0, np.float32)) == 0
    for i in range(2, 40):
        assert s.indicator(z["Q"][:, i]) == z["codes"][i]


@pytest.mark.parametrize("wl,n", [(1, 65536), (2, 16384), (3, 8192), (5, 8192)])
def test_points_workloads_vs_oracle(workload_dir, wl, n):
    d = workload_dir(wl)
    o = OracleScene(d)
    rng = np.random.default_rng(1000 + wl)
    Q = rng.uniform(-np.pi, np.pi, size=(o.ndof, n)).astype(np.float32)
    want = o.indicator_batch(Q, nthreads=8)
    got = _space(d).indicator_batch(Q)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (wl, bad.size, bad[:8], got[bad[:8]], want[bad[:8]])


def test_points_soa_offset_and_empty(workload_dir):
    d = workload_dir(1)
    o = OracleScene(d)
    s = _space(d)
    rng = np.random.default_rng(5)
    N, off = 1000, 1337  # offset larger than N, as the reference's buffers allow
    buf = rng.uniform(-np.pi, np.pi, size=off * (o.ndof - 1) + N).astype(np.float32)
    res = np.zeros(N, np.int32)
    s.indicator(buf, res, N, off)
    Q = np.stack([buf[i * off:i * off + N] for i in range(o.ndof)])
    assert np.array_equal(res, o.indicator_batch(Q))
    empty = np.zeros(1, np.int32)
    assert s.indicator(buf, empty, 0, off) == 0


def test_points_out_of_bounds(workload_dir):
    d = workload_dir(1)
    s = _space(d)
    Q = np.zeros((4, 3), np.float32)
    Q[2, 1] = 4.0   # above maxs = pi
    Q[0, 2] = -3.5  # below mins
    got = s.indicator_batch(Q)
    assert got[1] == 1 and got[2] == 1


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_edges_golden(golden_dir, name):
    z, d = golden_dir(name)
    s = _space(d, float(z["dq"]))
    N = z["QS"].shape[1]
    res = np.zeros(N, np.int32)
    before = s.get_numthreads_all()
    s.indicator2(z["QS"].ravel(), z["QE"].ravel(), res, N, N)
    assert s.get_numthreads_all() - before == int(z["numthreads"])
    assert np.array_equal(res != 0, z["edge_res"] != 0), np.flatnonzero((res != 0) != (z["edge_res"] != 0))


def _knn_edges(V, k, rng, nmax):
    """Undirected k-nearest-neighbour edges of the columns of V (brute force)."""
    X = V.T.astype(np.float64)
    d2 = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    nn = np.argsort(d2, axis=1)[:, :k]
    e = {(min(i, j), max(i, j)) for i in range(X.shape[0]) for j in nn[i]}
    e = sorted(e)
    sel = rng.permutation(len(e))[:nmax]
    I = np.array([e[t][0] for t in sel])
    J = np.array([e[t][1] for t in sel])
    return V[:, I].copy(), V[:, J].copy()


@pytest.mark.parametrize("wl,nv,ne", [(1, 600, 3000), (2, 500, 1500), (3, 400, 600)])
def test_edges_workloads_vs_oracle(workload_dir, wl, nv, ne):
    d = workload_dir(wl)
    o = OracleScene(d)
    rng = np.random.default_rng(2000 + wl)
    V = rng.uniform(-np.pi, np.pi, size=(o.ndof, nv)).astype(np.float32)
    QS, QE = _knn_edges(V, 10, rng, ne)
    want, tot = o.indicator2(QS, QE, nthreads=8)
    s = _space(d)
    N = QS.shape[1]
    res = np.zeros(N, np.int32)
    s.indicator2(QS.ravel(), QE.ravel(), res, N, N)
    assert np.array_equal(res != 0, want != 0), np.flatnonzero((res != 0) != (want != 0))[:10]


def test_edges_async_overlap(golden_dir):
    z, d = golden_dir("wl2")
    s = _space(d)
    N = z["QS"].shape[1]
    ra, rb = np.zeros(N, np.int32), np.zeros(N, np.int32)
    # two requests in flight (the reference's two-worker overlap)
    a = s.indicator2_async(z["QS"].ravel(), z["QE"].ravel(), ra, N, N)
    b = s.indicator2_async(z["QE"].ravel(), z["QS"].ravel(), rb, N, N)
    s.indicator2_async_wait(b)
    s.indicator2_async_wait(a)
    want = z["edge_res"] != 0
    assert np.array_equal(ra != 0, want)
    # reversed edges visit the same sample set except for the c1 rounding:
    # compare against the oracle run on the reversed edges
    o = OracleScene(d)
    rev, _ = o.indicator2(z["QE"], z["QS"])
    assert np.array_equal(rb != 0, rev != 0)


def test_compressed_indicator2(golden_dir):
    z, d = golden_dir("wl1")
    s = _space(d)
    o = OracleScene(d)
    QS = z["QS"][:, :10]
    QE = z["QE"][:, :40]
    posqe = np.array([0, 4, 8, 12, 16, 20, 24, 28, 32, 36], np.int32)
    numqe = np.array([4, 4, 4, 4, 4, 4, 4, 4, 4, 4], np.int32)
    res = np.zeros(40, np.int32)
    # the reference's signature has one component stride for qs and qe
    # (configspace.hpp:32): lay qs (10 items) out with the stride of qe (40)
    qe_pad = np.zeros((4, 40), np.float32)
    qs_pad = np.zeros((4, 40), np.float32)
    qs_pad[:, :10] = QS
    qe_pad[:, :40] = QE
    s.indicator2_compressed(qs_pad.ravel(), 10, qe_pad.ravel(), res, posqe, numqe, 40)
    S = np.repeat(QS, 4, axis=1)
    want, _ = o.indicator2(S, QE)
    assert np.array_equal(res != 0, want != 0)


def test_device_variants(workload_dir):
    import torch
    d = workload_dir(2)
    o = OracleScene(d)
    s = _space(d)
    rng = np.random.default_rng(77)
    Q = rng.uniform(-np.pi, np.pi, size=(6, 5000)).astype(np.float32)
    q = torch.from_numpy(Q).cuda()
    r = torch.zeros(5000, dtype=torch.int32, device="cuda")
    s.indicator_dev(q, r)
    torch.cuda.synchronize()
    assert np.array_equal(r.cpu().numpy(), o.indicator_batch(Q, nthreads=8))
    V = rng.uniform(-np.pi, np.pi, size=(6, 300)).astype(np.float32)
    QS, QE = _knn_edges(V, 10, rng, 800)
    qs, qe = torch.from_numpy(QS).cuda(), torch.from_numpy(QE).cuda()
    re = torch.zeros(QS.shape[1], dtype=torch.int32, device="cuda")
    st = s.indicator2_dev(qs, qe, re)
    want, tot = o.indicator2(QS, QE, nthreads=8)
    assert st.samples_total == tot
    assert 0 < st.samples_scheduled <= tot
    assert np.array_equal(re.cpu().numpy() != 0, want != 0)


def test_edge_rounds_overlapped_repeatable(workload_dir):
    """Long edges (several hundred samples: four rounds, the later ones
    compacted while the previous round still blocks edges) mixed with short
    ones: every repeated call gives the oracle's decisions, and the scheduled
    samples never exceed the reference's."""
    import torch
    d = workload_dir(2)
    o = OracleScene(d)
    rng = np.random.default_rng(4242)
    A = rng.uniform(-np.pi, np.pi, size=(6, 400)).astype(np.float32)
    B = rng.uniform(-np.pi, np.pi, size=(6, 400)).astype(np.float32)
    V = rng.uniform(-np.pi, np.pi, size=(6, 300)).astype(np.float32)
    KS, KE = _knn_edges(V, 10, rng, 800)
    QS = np.ascontiguousarray(np.concatenate([A, KS], axis=1))
    QE = np.ascontiguousarray(np.concatenate([B, KE], axis=1))
    want, tot = o.indicator2(QS, QE, nthreads=8)
    assert 0 < (want != 0).sum() < QS.shape[1]
    s = _space(d)
    qs, qe = torch.from_numpy(QS).cuda(), torch.from_numpy(QE).cuda()
    for rep in range(4):
        re = torch.full((QS.shape[1],), 7, dtype=torch.int32, device="cuda")
        st = s.indicator2_dev(qs, qe, re)
        assert st.samples_total == tot
        assert st.rounds >= 4
        assert 0 < st.samples_scheduled <= tot
        got = re.cpu().numpy()
        assert np.array_equal(got != 0, want != 0), (rep, np.flatnonzero((got != 0) != (want != 0))[:10])


@pytest.mark.parametrize("block,waves", [(256, 5), (256, 4), (1024, 4)])
def test_launch_shapes_agree(workload_dir, block, waves):
    """Every kernel variant and work-unit size gives the oracle's codes and
    decisions (units of 1 and 7 pairs split the pair list at odd places)."""
    import torch
    d = workload_dir(2)
    o = OracleScene(d)
    s = _space(d)
    rng = np.random.default_rng(block + waves)
    Q = rng.uniform(-np.pi, np.pi, size=(6, 3000)).astype(np.float32)
    want = o.indicator_batch(Q, nthreads=8)
    V = rng.uniform(-np.pi, np.pi, size=(6, 200)).astype(np.float32)
    QS, QE = _knn_edges(V, 10, rng, 500)
    wante, _ = o.indicator2(QS, QE, nthreads=8)
    q = torch.from_numpy(Q).cuda()
    qs, qe = torch.from_numpy(QS).cuda(), torch.from_numpy(QE).cuda()
    for ppu in (1, 7, 64, 256):
        sh = s.set_launch_shape(block, waves, ppu)
        assert (sh.block, sh.waves, sh.pairs_per_unit) == (block, waves, ppu) and sh.grid >= 256
        r = torch.zeros(Q.shape[1], dtype=torch.int32, device="cuda")
        s.indicator_dev(q, r)
        assert np.array_equal(r.cpu().numpy(), want), ppu
        re = torch.zeros(QS.shape[1], dtype=torch.int32, device="cuda")
        s.indicator2_dev(qs, qe, re)
        assert np.array_equal(re.cpu().numpy() != 0, wante != 0), ppu


def test_autotune_keeps_results(workload_dir):
    import gpuprm_amd
    import torch
    d = workload_dir(2)
    o = OracleScene(d)
    s = _space(d)
    rng = np.random.default_rng(5)
    Q = rng.uniform(-np.pi, np.pi, size=(6, 20000)).astype(np.float32)
    q = torch.from_numpy(Q).cuda()
    sh, ms = s.autotune(q, reps=2)
    assert sh.pairs_per_unit in (32, 64, 128, 256) and (sh.block, sh.waves) in ((256, 5), (256, 4), (1024, 4))
    assert ms > 0 and s.get_launch_shape().as_dict() == sh.as_dict()
    r = torch.zeros(Q.shape[1], dtype=torch.int32, device="cuda")
    s.indicator_dev(q, r)
    assert np.array_equal(r.cpu().numpy(), o.indicator_batch(Q, nthreads=8))
    with pytest.raises(gpuprm_amd.GprmError):
        s.set_launch_shape(512, 4, 32)
    with pytest.raises(gpuprm_amd.GprmError):
        s.set_launch_shape(256, 5, 0)
