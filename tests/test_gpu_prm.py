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
    import gpuprm_amd
    assert np.array_equal(path[0], qs) and np.array_equal(path[-1], qg)
    assert np.all(o.indicator_batch(path.T) == 0)
    steps = np.linalg.norm(np.diff(path.astype(np.float64), axis=0), axis=1)
    assert np.all(steps < D * (1 + 1e-6))
    # each edge in the orientation the planner checked it
    S, E = gpuprm_amd.canonical_edges(path[:-1], path[1:])
    res, _ = o.indicator2(S.T.copy(), E.T.copy(), nthreads=8)
    assert np.all(res == 0), np.flatnonzero(res)


def test_prm_reference_task_config1(golden_dir, tmp_path):
    # prm/src/main.cpp:128-133: start (0,0,0,0), goal (3,3,3,3), D = 1.0
    import gpuprm_amd
    z, d = golden_dir("config1")
    s = _space(d)
    o = OracleScene(d)
    qs = np.zeros(4, np.float32)
    qg = np.full(4, 3.0, np.float32)
    out = str(tmp_path / "prmoutput")
    r, path = s.prm_solve(qs, qg, num=1024, D=1.0, max_nodes=1 << 17, max_steps=300, seed=3, output_dir=out)
    print(r.as_dict())
    assert r.found == 1 and path is not None
    _verify(o, path, qs, qg, 1.0)
    # the output directory, read back as prm/matlab/lib/prmoutputread.m reads it
    g = gpuprm_amd.read_prm_output(out, 4)
    assert g["connected"] == 1
    assert g["Ql"].shape == (r.nodes_left, 4) and g["Qr"].shape == (r.nodes_right, 4)
    assert np.array_equal(g["Ql"][0], qs) and np.array_equal(g["Qr"][0], qg)
    assert np.array_equal(g["Ql"][g["conl"]], g["Qr"][g["conr"]])
    # Qpath runs goal -> connection (twice: once per tree) -> start
    Qp = g["Qpath"]
    assert len(Qp) == len(path) + 1
    assert np.array_equal(np.delete(Qp, len(g["pathr"]), axis=0)[::-1], path)
    for side in ("l", "r"):
        Q, E = g["Q" + side], g["E" + side]
        assert np.all(np.linalg.norm(Q[E[:, 0]] - Q[E[:, 1]], axis=1) < 1.0)
        raw_f = np.fromfile(out + "/graph%s/edgesfrom.bin" % side, np.int32)
        raw_t = np.fromfile(out + "/graph%s/edgesto.bin" % side, np.int32)
        assert sorted(zip(raw_f, raw_t)) == sorted(zip(raw_t, raw_f))  # both directions stored
        surr = np.fromfile(out + "/graph%s/surrnum.bin" % side, np.int32)
        assert np.array_equal(surr, np.bincount(raw_f, minlength=len(Q)))
        w = np.fromfile(out + "/graph%s/edgesweight.bin" % side, np.int32)
        dist = np.linalg.norm(Q[raw_f].astype(np.float64) - Q[raw_t], axis=1)
        assert np.array_equal(w, dist.astype(np.int32))  # truncated like the reference (all 0 below D = 1)
        pth = g["path" + side]
        edges = set(zip(raw_f.tolist(), raw_t.tolist()))
        assert all((int(a), int(b)) in edges for a, b in zip(pth[:-1], pth[1:]))


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


@pytest.mark.parametrize("workers", [1, 2])
def test_prm_every_tree_node_and_edge_verified(golden_dir, tmp_path, workers):
    """Every node and every edge of both trees (not only the path) is
    re-decided by the oracle; the trees are connected and consistent
    (surrnum = degree, edges shorter than D); the returned path is a shortest
    path of the written graph (scipy Dijkstra in float64 over the edges'
    lengths)."""
    import gpuprm_amd
    csgraph = pytest.importorskip("scipy.sparse.csgraph")
    sparse = pytest.importorskip("scipy.sparse")
    z, d = golden_dir("config1")
    s = _space(d)
    o = OracleScene(d)
    qs = np.zeros(4, np.float32)
    qg = np.full(4, 3.0, np.float32)
    out = str(tmp_path / "trees")
    D = 1.0
    print("planner: config1, workers %d" % workers, flush=True)
    r, path = s.prm_solve(qs, qg, num=512, D=D, max_nodes=1 << 16, max_steps=300, seed=11, output_dir=out,
                          workers=workers)
    print(r.as_dict())
    assert r.found == 1
    g = gpuprm_amd.read_prm_output(out, 4)
    lengths = {}
    for side, root in (("l", 0), ("r", 0)):
        Q, E = g["Q" + side], g["E" + side]
        n = len(Q)
        # nodes: all free by the exact checker of record
        assert np.all(o.indicator_batch(Q.T.copy(), nthreads=8) == 0)
        # edges: each once (from < to), shorter than D, free by the oracle's
        # host loop of indicator2 in the orientation the planner checked them
        assert len(E) > 0
        L = np.linalg.norm(Q[E[:, 0]].astype(np.float64) - Q[E[:, 1]], axis=1)
        assert np.all(L < D * (1 + 1e-6))
        A, B = gpuprm_amd.canonical_edges(Q[E[:, 0]], Q[E[:, 1]])
        res, _ = o.indicator2(A.T.copy(), B.T.copy(), nthreads=8)
        bad = np.flatnonzero(res != 0)
        assert bad.size == 0, (side, bad[:8])
        surr = np.fromfile(out + "/graph%s/surrnum.bin" % side, np.int32)
        deg = np.bincount(E.ravel(), minlength=n)
        assert np.array_equal(surr, deg)
        # connected from the root, and the path is a shortest path
        A = sparse.coo_matrix((L, (E[:, 0], E[:, 1])), shape=(n, n)).tocsr()
        dist = csgraph.dijkstra(A, directed=False, indices=root)
        assert np.all(np.isfinite(dist))
        lengths[side] = dist
    for side, conn, pth in (("l", g["conl"], g["pathl"]), ("r", g["conr"], g["pathr"])):
        Q = g["Q" + side]
        plen = np.linalg.norm(np.diff(Q[pth].astype(np.float64), axis=0), axis=1).sum()
        # float32 sums in the planner's Dijkstra: equal up to their rounding
        assert plen == pytest.approx(lengths[side][conn], rel=1e-4, abs=1e-5), (side, plen, lengths[side][conn])
    _verify(o, path, qs, qg, D)
    assert r.nodes_left == len(g["Ql"]) and r.nodes_right == len(g["Qr"])
    # free edges of the last step's nodes after the connection are not inserted
    assert r.edges_free >= len(g["El"]) + len(g["Er"])


def _edge_set(E):
    return E[np.lexsort((E[:, 1], E[:, 0]))]


@pytest.mark.parametrize("budget", [0, 4])
def test_prm_radius_search_placement_independent(golden_dir, tmp_path, budget):
    """The radius search places groups in the order its blocks finish, so the
    step's edge arrays differ between runs; the trees must not.  Without a
    budget: two runs.  With an edge budget (shortest first, no model) a group
    keeps its min(512, nbuf / groups) shortest matches, cut by a sort in LDS
    when it has more (config1's groups hold hundreds): nbuf 2^20 (keep 512)
    against 2^17 (keep 128) must pick the same budget edges.  Identical node
    arrays, the same edge sets and counts, every tree edge free by the
    oracle."""
    import gpuprm_amd
    z, d = golden_dir("config1")
    s = _space(d)
    o = OracleScene(d)
    qs = np.zeros(4, np.float32)
    qg = np.full(4, 3.0, np.float32)
    runs = []
    for nbuf in ((1 << 20, 1 << 20) if budget == 0 else (1 << 20, 1 << 17)):
        out = str(tmp_path / ("trees_%d_%d" % (nbuf, len(runs))))
        print("planner: config1, budget %d, nbuf %d" % (budget, nbuf), flush=True)
        r, path = s.prm_solve(qs, qg, num=512, D=1.0, max_nodes=1 << 16, max_steps=300, seed=11, output_dir=out,
                              edge_budget=budget, nbuf=nbuf)
        print(r.as_dict())
        runs.append((r, path, gpuprm_amd.read_prm_output(out, 4)))
    (r0, p0, g0), (r1, p1, g1) = runs
    for k in ("found", "steps", "nodes_left", "nodes_right", "point_checks", "edges_checked", "edges_free",
              "edge_samples"):
        assert getattr(r0, k) == getattr(r1, k), k
    if budget:
        assert r0.edges_skipped > r1.edges_skipped > 0  # the smaller keep stores fewer
    for side in ("l", "r"):
        assert np.array_equal(g0["Q" + side], g1["Q" + side])
        assert np.array_equal(_edge_set(g0["E" + side]), _edge_set(g1["E" + side]))
        Q, E = g0["Q" + side], g0["E" + side]
        A, B = gpuprm_amd.canonical_edges(Q[E[:, 0]], Q[E[:, 1]])
        res, _ = o.indicator2(A.T.copy(), B.T.copy(), nthreads=8)
        assert np.all(res == 0)
    if r0.found:
        assert np.array_equal(p0, p1)
        _verify(o, p0, qs, qg, 1.0)


@pytest.mark.parametrize("budget", [0, 4])
def test_prm_radius_search_complete_for_large_groups(golden_dir, tmp_path, budget):
    """Groups with more matches than the radius search's LDS stage holds
    (config1 at D = 5 after 11 steps: over a thousand earlier nodes within D
    of a new node), so a group's matches are stored in several ranges, or
    (with an edge budget) sorted and cut in the stage.  The planner is
    deterministic, so k + 1 steps repeat the k-step trees and add one step's
    nodes; each of those must have an edge to exactly the earlier nodes of its
    tree within D (the planner's float32 sum in coordinate order, by brute
    force here) whose edge the exact checker finds free, and to no other node;
    with a budget of b, to the free ones among its b nearest (a node whose
    b-th and (b+1)-th distances tie is not checked: the planner breaks the
    tie by its traversal order)."""
    import gpuprm_amd
    z, d = golden_dir("config1")
    s = _space(d)
    qs = np.zeros(4, np.float32)
    qg = np.full(4, 3.0, np.float32)
    D, num, k = 5.0, 512, 11
    runs = []
    for steps in (k, k + 1):
        out = str(tmp_path / ("trees_%d" % steps))
        print("planner: config1, D %.1f, %d steps, budget %d" % (D, steps, budget), flush=True)
        r, _ = s.prm_solve(qs, qg, num=num, D=D, max_nodes=1 << 15, max_steps=steps, seed=5, output_dir=out,
                           nbuf=1 << 22, edge_budget=budget)
        print(r.as_dict())
        assert r.found == 0 and r.steps == steps
        runs.append(gpuprm_amd.read_prm_output(out, 4))
    ga, gb = runs
    largest = 0
    for side in ("l", "r"):
        Qa, Qb, Eb = ga["Q" + side], gb["Q" + side], gb["E" + side]
        na = len(Qa)
        assert np.array_equal(Qb[:na], Qa)
        assert np.array_equal(_edge_set(Eb[Eb[:, 1] < na]), _edge_set(ga["E" + side]))
        got = Eb[Eb[:, 1] >= na]
        assert np.all(got[:, 0] < na)  # the step's nodes connect to earlier ones only
        if len(Qb) == na:
            continue
        new = Qb[na:]
        diff = Qa[None, :, :] - new[:, None, :]
        s2 = diff[..., 0] * diff[..., 0]
        for i in range(1, 4):
            s2 = s2 + diff[..., i] * diff[..., i]
        within = s2 < np.float32(D * D)
        largest = max(largest, int(within.sum(axis=1).max()))
        checked = np.ones(len(new), bool)
        if budget:
            # the b nearest within D; a tie across the cut leaves the node unchecked
            order = np.argsort(np.where(within, s2, np.inf), axis=1, kind="stable")
            keep = np.zeros_like(within)
            for j in range(len(new)):
                c = int(within[j].sum())
                if c > budget and s2[j, order[j, budget - 1]] == s2[j, order[j, budget]]:
                    checked[j] = False
                keep[j, order[j, :min(c, budget)]] = True
            within &= keep
        cand = np.argwhere(within)  # (new node, earlier node)
        A, B = gpuprm_amd.canonical_edges(new[cand[:, 0]], Qa[cand[:, 1]])
        n = len(cand)
        res = np.zeros(n, np.int32)
        s.indicator2(A.T.copy(), B.T.copy(), res, n, n)
        free = cand[(res == 0) & checked[cand[:, 0]]]
        want = np.stack([free[:, 1], free[:, 0] + na], axis=1)
        got = got[checked[got[:, 1] - na]]
        print(side, "earlier nodes", na, "new", len(new), "candidates", n, "free", len(want), "edges", len(got),
              "unchecked (ties)", int((~checked).sum()))
        assert np.array_equal(_edge_set(got), _edge_set(want))
    print("largest group", largest)
    assert largest > 1024  # more than two chunks' matches: the stage was stored or cut mid-group
