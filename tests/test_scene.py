"""Scene layer: generators, hull adjacency convention, directory format."""
import os

import numpy as np
import pytest

import scenes

REF_CONFIG1 = "/root/reference/prm/config1"


def _rows(s, k):
    """Adjacency rows of polytope k as a list of lists."""
    vo = int(np.sum(s["n"][:k]))
    eo = int(np.sum(s["m"][:k]))
    n = int(s["n"][k])
    dsp, cnt = s["dsp"][vo:vo + n], s["cnt"][vo:vo + n]
    dest = s["dest"][eo:eo + int(s["m"][k])]
    return [list(dest[dsp[i]:dsp[i] + cnt[i]]) for i in range(n)], s["vertices"][3 * vo:3 * (vo + n)].reshape(n, 3)


@pytest.mark.parametrize("wl", [1, 2, 3, 5])
def test_generated_scene_structure(workload_dir, wl):
    s = scenes.read_dir(workload_dir(wl))
    nd = int(s["ndof"])
    assert nd == {1: 4, 2: 6, 3: 7, 5: 7}[wl]
    N = len(s["n"])
    nobs = int(np.sum(s["sys"] == 0))
    assert list(s["sys"]) == [0] * nobs + list(range(1, nd + 1))
    # pairs: obstacle x link, then non-adjacent link pairs; sorted by from
    want = [(e, nobs + l) for e in range(nobs) for l in range(nd)]
    want += [(nobs + i, nobs + j) for i in range(nd) for j in range(i + 2, nd)]
    assert list(zip(s["from"].tolist(), s["to"].tolist())) == want
    assert np.all(np.diff(s["from"]) >= 0)
    for k in range(N):
        rows, V = _rows(s, k)
        n = len(rows)
        E = sum(len(r) for r in rows)
        # triangulated sphere: 3n-6 undirected edges, symmetric, rows ascending
        assert E == 2 * (3 * n - 6), (k, n, E)
        for i, r in enumerate(rows):
            assert r == sorted(r) and len(set(r)) == len(r) and i not in r
            for j in r:
                assert i in rows[j]


def test_generator_is_deterministic(tmp_path):
    import gpuprm_amd
    a = scenes.read_dir(gpuprm_amd.generate_scene(2, 99, str(tmp_path / "a")))
    b = scenes.read_dir(gpuprm_amd.generate_scene(2, 99, str(tmp_path / "b")))
    c = scenes.read_dir(gpuprm_amd.generate_scene(2, 100, str(tmp_path / "c")))
    for k in scenes.SCENE_KEYS:
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["vertices"], c["vertices"])


def test_workload_parameters(workload_dir):
    s = scenes.read_dir(workload_dir(1))
    assert np.allclose(s["a"], 1.0) and np.all(s["d"] == 0)
    assert np.allclose(s["alpha"], [0, np.pi / 2, 0, np.pi / 2])
    assert np.all(s["mins"] == np.float32(-np.pi)) and np.all(s["maxs"] == np.float32(np.pi))
    # link boxes 1.0 x 0.2 x 0.2 spanning x in [0, 1]
    rows, V = _rows(s, 8)
    assert np.allclose(V.min(0), [0, -0.1, -0.1]) and np.allclose(V.max(0), [1.0, 0.1, 0.1])
    # obstacle boxes: 8 vertices, half-extents in [0.2, 0.6], centres |c| >= 0.5 in [-3,3]^3
    for k in range(8):
        rows, V = _rows(s, k)
        c = V.mean(0)
        assert np.linalg.norm(c) >= 0.5 - 1e-5 and np.all(np.abs(c) <= 3 + 1e-5)
        e = np.sort([np.linalg.norm(V[1] - V[0]), np.linalg.norm(V[2] - V[0]), np.linalg.norm(V[4] - V[0])])
        assert np.all(e >= 0.4 - 1e-5) and np.all(e <= 1.2 + 1e-5)
    s2 = scenes.read_dir(workload_dir(2))
    assert np.allclose(s2["alpha"], [0, np.pi / 2, -np.pi / 2] * 2) and np.allclose(s2["a"], 0.8)


def test_hull_matches_scipy(workload_dir):
    """Hull adjacency of generated obstacles equals Qhull's triangulation
    (the reference's scene tooling uses Qhull through MATLAB convhulln,
    prm/matlab/lib/polygen2.m:4-34)."""
    ConvexHull = pytest.importorskip("scipy.spatial").ConvexHull
    s = scenes.read_dir(workload_dir(2))
    for k in range(16):
        rows, V = _rows(s, k)
        h = ConvexHull(V.astype(np.float64))
        assert len(h.vertices) == len(V)  # interior points were dropped
        edges = set()
        for f in h.simplices:
            for a in range(3):
                i, j = sorted((int(f[a]), int(f[(a + 1) % 3])))
                edges.add((i, j))
        mine = {(i, j) for i, r in enumerate(rows) for j in r if i < j}
        assert mine == edges


def test_config1_round_trip(tmp_path):
    """numpy reader/writer and the C++ loader agree on the reference's scene."""
    if not os.path.isdir(REF_CONFIG1):
        pytest.skip("reference scene not present")
    s = scenes.read_dir(REF_CONFIG1)
    d = scenes.write_dir(s, str(tmp_path / "c1"))
    s2 = scenes.read_dir(d)
    for k in scenes.SCENE_KEYS:
        assert np.array_equal(np.asarray(s[k]), np.asarray(s2[k])), k
    # byte-identical files
    for root, _, files in os.walk(REF_CONFIG1):
        for f in files:
            a = os.path.join(root, f)
            b = os.path.join(d, os.path.relpath(a, REF_CONFIG1))
            assert open(a, "rb").read() == open(b, "rb").read(), a


def test_golden_config1_fixture_is_the_reference_scene(golden_dir):
    if not os.path.isdir(REF_CONFIG1):
        pytest.skip("reference scene not present")
    z, d = golden_dir("config1")
    s = scenes.read_dir(REF_CONFIG1)
    for k in scenes.SCENE_KEYS:
        assert np.array_equal(np.asarray(s[k]), np.asarray(z["scene_" + k])), k
