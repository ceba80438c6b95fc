"""The oracle pinned by mathematics, independently of the reference build:

  - forward kinematics against a float64 product of modified-DH matrices
    (prm/src/lib/robotspace/kinematics.h:27-37, geo4.h:194-203);
  - separating-vector decisions against an exact intersection test (linear
    programming in double precision) on random convex hulls, outside a +-5%
    scale band around contact;
  - decision invariance under swapping P and Q and under rigid motions;
  - the edge sample count and the edge result as the OR of its samples.
"""
import numpy as np
import pytest

import scenes
from oracle.oracle import OracleScene

linprog = pytest.importorskip("scipy.optimize").linprog
ConvexHull = pytest.importorskip("scipy.spatial").ConvexHull


def dh(a, alpha, q, d):
    ca, sa, cq, sq = np.cos(alpha), np.sin(alpha), np.cos(q), np.sin(q)
    return np.array([[cq, -sq, 0, a], [sq * ca, cq * ca, -sa, -sa * d], [sq * sa, cq * sa, ca, ca * d],
                     [0, 0, 0, 1]], np.float64)


@pytest.mark.parametrize("name", ["config1", "wl2", "wl3"])
def test_fk_matches_float64_dh_product(golden_dir, name):
    z, d = golden_dir(name)
    o = OracleScene(d)
    a, al, d0, th = (z["scene_" + k].astype(np.float64) for k in ("a", "alpha", "d", "q"))
    rng = np.random.default_rng(3)
    for _ in range(50):
        q = rng.uniform(-np.pi, np.pi, o.ndof).astype(np.float32)
        T = o.fk(q)
        M = np.eye(4)
        assert np.allclose(T[0, :, :3].T, np.eye(4)[:3], atol=0)
        for i in range(o.ndof):
            M = M @ dh(a[i], al[i], float(q[i]), d0[i])
            got = T[i + 1, :, :3].T  # rows x,y,z; columns 0..3
            assert np.allclose(got, M[:3], atol=2e-5 * (1 + np.abs(M[:3]).max())), (i, got, M[:3])


def test_fk_closed_form_at_zero(golden_dir):
    # config1 at q = 0: frame translations follow from a, d and alpha alone
    z, d = golden_dir("config1")
    o = OracleScene(d)
    T = o.fk(np.zeros(4, np.float32))
    M = np.eye(4)
    for i in range(4):
        M = M @ dh(z["scene_a"][i], z["scene_alpha"][i], 0.0, z["scene_d"][i])
        assert np.allclose(T[i + 1, 3, :3], M[:3, 3], atol=1e-5)


def _hull_arrays(P):
    h = ConvexHull(P)
    keep = sorted(set(h.vertices.tolist()))
    idx = {v: i for i, v in enumerate(keep)}
    nb = [set() for _ in keep]
    for f in h.simplices:
        for a in range(3):
            i, j = idx[int(f[a])], idx[int(f[(a + 1) % 3])]
            nb[i].add(j)
            nb[j].add(i)
    V = P[keep].astype(np.float32)
    dsp, cnt, dest = [], [], []
    for r in nb:
        dsp.append(len(dest))
        cnt.append(len(r))
        dest.extend(sorted(r))
    return V, np.array(dsp, np.int32), np.array(cnt, np.int32), np.array(dest, np.int32)


def _pair_scene(A, B, path):
    """Two world-frame hulls and one pair (0, 1) behind a 1-DOF dummy arm."""
    va, da, ca, ea = _hull_arrays(A)
    vb, db, cb, eb = _hull_arrays(B)
    s = {"ndof": np.int32(1), "a": np.zeros(1, np.float32), "alpha": np.zeros(1, np.float32),
         "q": np.zeros(1, np.float32), "d": np.zeros(1, np.float32), "types": np.zeros(1, np.int32),
         "mins": np.full(1, -10, np.float32), "maxs": np.full(1, 10, np.float32),
         "sys": np.zeros(2, np.int32), "from": np.array([0], np.int32), "to": np.array([1], np.int32),
         "n": np.array([len(va), len(vb)], np.int32), "m": np.array([len(ea), len(eb)], np.int32),
         "vertices": np.concatenate([va.ravel(), vb.ravel()]), "dsp": np.concatenate([da, db]),
         "cnt": np.concatenate([ca, cb]), "dest": np.concatenate([ea, eb])}
    scenes.write_dir(s, path)
    return va.astype(np.float64), vb.astype(np.float64)


def _intersect(A, B):
    """Exact-in-double feasibility: is there x = A^T lam = B^T mu (simplex weights)?"""
    na, nb = len(A), len(B)
    Aeq = np.zeros((5, na + nb))
    Aeq[:3, :na] = A.T
    Aeq[:3, na:] = -B.T
    Aeq[3, :na] = 1
    Aeq[4, na:] = 1
    r = linprog(np.zeros(na + nb), A_eq=Aeq, b_eq=[0, 0, 0, 1, 1], bounds=(0, None), method="highs")
    return r.status == 0


def _scaled(P, f):
    c = P.mean(0)
    return c + f * (P - c)


def _truth(A, B, band=0.05):
    if not _intersect(_scaled(A, 1 + band), _scaled(B, 1 + band)):
        return 0          # separated with margin
    if _intersect(_scaled(A, 1 - band), _scaled(B, 1 - band)):
        return 1          # overlapping with margin
    return None           # within the band around contact


def test_sva_decisions_match_exact_intersection(tmp_path):
    rng = np.random.default_rng(11)
    decided = agree = 0
    classes = [0, 0]
    for t in range(150):
        na, nb = rng.integers(6, 16, size=2)
        A = rng.normal(size=(na, 3)) * rng.uniform(0.2, 0.6)
        B = rng.normal(size=(nb, 3)) * rng.uniform(0.2, 0.6) + rng.normal(size=3) * rng.uniform(0.0, 1.5)
        d = str(tmp_path / ("p%d" % t))
        Ah, Bh = _pair_scene(A, B, d)
        truth = _truth(Ah, Bh)
        if truth is None:
            continue
        o = OracleScene(d)
        code, _ = o.check_config(np.zeros(1, np.float32))
        decided += 1
        classes[truth] += 1
        agree += int((code != 0) == bool(truth))
    assert decided > 100 and min(classes) > 30, classes
    assert agree == decided, (agree, decided)


def test_sva_invariances(tmp_path):
    """Swapping P/Q and rotating/translating both keep the decision (outside the band)."""
    rng = np.random.default_rng(12)
    checked = 0
    for t in range(60):
        A = rng.normal(size=(10, 3)) * 0.4
        B = rng.normal(size=(10, 3)) * 0.4 + rng.normal(size=3) * 0.8
        Ah, Bh = _pair_scene(A, B, str(tmp_path / ("a%d" % t)))
        truth = _truth(Ah, Bh)
        if truth is None:
            continue
        codes = []
        _pair_scene(B, A, str(tmp_path / ("b%d" % t)))
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        sh = rng.normal(size=3)
        _pair_scene(A @ Q.T + sh, B @ Q.T + sh, str(tmp_path / ("c%d" % t)))
        for tag in "abc":
            codes.append(OracleScene(str(tmp_path / ("%s%d" % (tag, t)))).check_config(np.zeros(1, np.float32))[0])
        assert len({c != 0 for c in codes}) == 1, codes
        checked += 1
    assert checked > 30


def test_testnum_formula(golden_dir):
    z, d = golden_dir("wl2")
    o = OracleScene(d, float(z["dq"]))
    tn, tot = o.edge_testnum(z["QS"], z["QE"])
    diff = (z["QE"] - z["QS"]).astype(np.float32)
    norm = np.zeros(diff.shape[1], np.float32)
    for i in range(diff.shape[0]):
        norm = (norm + diff[i] * diff[i]).astype(np.float32)
    want = (np.sqrt(norm) / np.float32(z["dq"])).astype(np.float32).astype(np.int64) + 1
    assert np.array_equal(tn, want) and tot == want.sum()


def test_edge_result_is_or_of_samples(golden_dir):
    z, d = golden_dir("wl1")
    o = OracleScene(d, float(z["dq"]))
    QS, QE = z["QS"][:, :40], z["QE"][:, :40]
    res, _ = o.indicator2(QS, QE)
    tn, _ = o.edge_testnum(QS, QE)
    for k in range(QS.shape[1]):
        codes = [o.check_config(o.edge_sample(QS, QE, k, j, int(tn[k])))[0] for j in range(tn[k])]
        blocked = [c for c in codes if c != 0]
        assert (res[k] != 0) == bool(blocked)
        if blocked:
            assert res[k] == blocked[-1]  # last colliding sample's code (host loop order)
