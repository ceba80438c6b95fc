"""GPU radius search (gprm_radius_neighbours, the planner's get_near_vertices,
prm/src/lib/prm/prmsolver.cpp:2532-2574) against brute force on the CPU:
the same pairs, the same float32 squared distances (summed in coordinate
order, as the reference's norm loop does)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _brute(nodes, queries, D):
    s = np.zeros((queries.shape[1], nodes.shape[1]), np.float32)
    for i in range(nodes.shape[0]):
        diff = nodes[i][None, :] - queries[i][:, None]
        s = s + diff * diff
    q, n = np.nonzero(s < np.float32(D) * np.float32(D))
    return q.astype(np.int32), n.astype(np.int32), s[q, n]


@pytest.mark.parametrize("ndof,D,n", [(7, 0.5, 20000), (7, 1.3, 5000), (4, 0.3, 20000), (2, 4.0, 3000),
                                      (6, 0.05, 1)])
def test_radius_neighbours_match_brute_force(ndof, D, n):
    import gpuprm_amd
    import torch
    rng = np.random.default_rng(ndof * 100 + n)
    lo = np.full(ndof, -np.pi, np.float32)
    hi = np.full(ndof, np.pi, np.float32)
    # clustered nodes, like a tree grown by local sampling, plus uniform ones
    centres = rng.uniform(-2.5, 2.5, size=(ndof, 8))
    nodes = np.concatenate([centres[:, rng.integers(0, 8, n // 2)] + rng.normal(0, 0.4, size=(ndof, n // 2)),
                            rng.uniform(-np.pi, np.pi, size=(ndof, n - n // 2))], axis=1)
    nodes = np.clip(nodes, -np.pi, np.pi).astype(np.float32)
    queries = np.clip(nodes[:, rng.integers(0, n, 400)] + rng.normal(0, D / 2, size=(ndof, 400)),
                      -np.pi, np.pi).astype(np.float32)
    want = _brute(nodes, queries, D)
    gq, gn, gd = gpuprm_amd.radius_neighbours(torch.from_numpy(nodes).cuda(), torch.from_numpy(queries).cuda(),
                                              lo, hi, D, cap=max(16, 2 * len(want[0])))
    o = np.lexsort((want[1], want[0]))
    assert np.array_equal(gq, want[0][o]) and np.array_equal(gn, want[1][o])
    assert np.array_equal(gd.view(np.int32), want[2][o].view(np.int32))


def test_radius_neighbours_capacity_and_errors():
    import gpuprm_amd
    import torch
    nodes = torch.zeros((3, 50), dtype=torch.float32, device="cuda")
    lo = np.full(3, -1, np.float32)
    hi = np.full(3, 1, np.float32)
    with pytest.raises(gpuprm_amd.GprmError):
        gpuprm_amd.radius_neighbours(nodes, nodes[:, :10].contiguous(), lo, hi, 0.5, cap=100)  # 500 pairs
    with pytest.raises(gpuprm_amd.GprmError):
        gpuprm_amd.radius_neighbours(nodes, nodes, lo, hi, -1.0)


@pytest.mark.parametrize("ndof,D", [(7, 0.5), (4, 0.3), (5, 1.0)])
def test_radius_neighbours_at_cell_faces(ndof, D):
    """Cells whose faces lie D or more from a query are skipped: queries on
    and beside cell faces, nodes at distance D (1 -/+ 1e-6) from them in
    random directions, across one or several faces at once; the matches
    equal brute force."""
    import gpuprm_amd
    import torch
    rng = np.random.default_rng(ndof * 7 + 1)
    lo = np.full(ndof, -np.pi, np.float32)
    hi = np.full(ndof, np.pi, np.float32)
    cells = max(1, int(np.floor(2 * np.pi / D)))
    w = 2 * np.pi / cells
    nq = 300
    k = rng.integers(1, cells, size=(ndof, nq))
    queries = -np.pi + k * w + rng.choice([-1e-6, 0.0, 1e-6, 1e-3], size=(ndof, nq)) * w
    queries = np.clip(queries, -np.pi, np.pi).astype(np.float32)
    nodes = []
    for scale in (1 - 1e-6, 1 + 1e-6, 0.999, 0.7):
        u = rng.normal(size=(ndof, nq))
        # mostly along a few axes, so that nodes sit across one or two faces
        u[rng.random(size=u.shape) < 0.6] = 0.0
        u[0, np.all(u == 0, axis=0)] = 1.0
        u /= np.linalg.norm(u, axis=0)
        nodes.append(queries + D * scale * u)
    nodes = np.clip(np.concatenate(nodes, axis=1), -np.pi, np.pi).astype(np.float32)
    want = _brute(nodes, queries, D)
    assert len(want[0]) > nq
    gq, gn, gd = gpuprm_amd.radius_neighbours(torch.from_numpy(nodes).cuda(), torch.from_numpy(queries).cuda(),
                                              lo, hi, D, cap=max(16, 2 * len(want[0])))
    o = np.lexsort((want[1], want[0]))
    assert np.array_equal(gq, want[0][o]) and np.array_equal(gn, want[1][o])
    assert np.array_equal(gd.view(np.int32), want[2][o].view(np.int32))
