"""The oracle's counting build (liboracle_count.so), which gives the
algorithmic FP32 operations behind the benchmark's roofline, pinned by a
count made by hand.

A separated pair whose start vertices are already the support vertices ends
in one iteration of the separating-vector algorithm (collision4.h:227-300).
Counting add, sub, mul, div and sqrt (negations and compares excluded), the
reference executes:

  start      2 x trafo4::apply (3 rows x (3 mul + 3 add))          36
             difference of the two points                             3
             normalize (3 mul, 2 add, sqrt, 1.0/x, 3 mul)            10
  iteration  2 x trafo4::apply_rot_inv (3 rows x (3 mul + 2 add))    30
             climb of P: sprod at the start vertex, one per neighbour
                         of its row, none improving: 5 (1 + deg P0)
             climb of Q: 5 (1 + deg Q0)
             2 x trafo4::apply of the support vertices               36
             difference, normalize, sprod with S              3 + 10 + 5

  total = 143 + 5 (deg P0 + deg Q0)
"""
import numpy as np
import pytest

import scenes
from oracle.oracle import OracleScene

ConvexHull = pytest.importorskip("scipy.spatial").ConvexHull


def _box(lo, hi, first):
    """Corners of the axis-aligned box [lo, hi]^3, the corner `first` as
    vertex 0, with the hull adjacency in the reference's convention (rows
    ascending, symmetric)."""
    C = np.array([[x, y, z] for x in (lo, hi) for y in (lo, hi) for z in (lo, hi)], np.float64)
    i0 = int(np.flatnonzero(np.all(C == np.asarray(first, np.float64), axis=1))[0])
    order = [i0] + [i for i in range(8) if i != i0]
    C = C[order]
    nb = [set() for _ in range(8)]
    for f in ConvexHull(C).simplices:
        for a in f:
            for b in f:
                if a != b:
                    nb[a].add(int(b))
    rows = [sorted(x) for x in nb]
    cnt = np.array([len(r) for r in rows], np.int32)
    dsp = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32)
    dest = np.array([v for r in rows for v in r], np.int32)
    return C.astype(np.float32), dsp, cnt, dest


def _scene(tmp_path, P, Q):
    """A one-joint arm at q = 0 (its frame is the world frame) carrying box
    P, and the world-frame obstacle Q; one pair (P, Q)."""
    s = {"ndof": 1, "a": np.zeros(1), "alpha": np.zeros(1), "q": np.zeros(1), "d": np.zeros(1),
         "types": np.zeros(1, np.int32), "mins": np.full(1, -np.pi), "maxs": np.full(1, np.pi),
         "sys": np.array([1, 0], np.int32), "from": np.array([0], np.int32), "to": np.array([1], np.int32),
         "n": np.array([8, 8], np.int32), "m": np.array([len(P[3]), len(Q[3])], np.int32),
         "vertices": np.concatenate([P[0].ravel(), Q[0].ravel()]).astype(np.float32),
         "dsp": np.concatenate([P[1], Q[1]]), "cnt": np.concatenate([P[2], Q[2]]),
         "dest": np.concatenate([P[3], Q[3]])}
    return scenes.write_dir(s, str(tmp_path / "pair"))


@pytest.mark.parametrize("gap", [2.0, 0.5, 7.25])
@pytest.mark.parametrize("pfirst", [(1, 1, 1), (1, 1, 0)])
def test_one_iteration_separated_pair(tmp_path, gap, pfirst):
    # P = [0,1]^3 on the arm, Q = [1+gap, 2+gap]^3: the corners facing each
    # other along the start direction are both supports of it
    lo, hi = 1.0 + gap, 2.0 + gap
    P = _box(0.0, 1.0, pfirst)
    Q = _box(lo, hi, (lo, lo, lo))
    if pfirst == (1, 1, 0):
        # mirrored through z = 0 (same adjacency), so that P's corner (1,1,0)
        # faces Q's corner (lo, lo, -lo); P's vertex 0 has another degree
        Q = (Q[0] * np.array([1, 1, -1], np.float32),) + Q[1:]
    d = _scene(tmp_path, P, Q)
    o = OracleScene(d, 0.01, counting=True)
    codes, flops = o.pair_profile(np.zeros(1, np.float32))
    assert codes.tolist() == [0]
    degP, degQ = int(P[2][0]), int(Q[2][0])
    # both start vertices must be the supports for the hand count to hold
    S = Q[0][0].astype(np.float64) - P[0][0].astype(np.float64)
    assert np.argmax(P[0] @ S) == 0 and np.argmax(-(Q[0] @ S)) == 0
    assert int(flops[0]) == 143 + 5 * (degP + degQ), (int(flops[0]), degP, degQ)


def test_climb_costs_five_per_neighbour_visited(tmp_path):
    """Starting P's climb one vertex away from its support adds the
    operations of the neighbours visited: 5 per dot product."""
    gap = 2.0
    lo, hi = 1.0 + gap, 2.0 + gap
    Q = _box(lo, hi, (lo, lo, lo))
    P_at_support = _box(0.0, 1.0, (1, 1, 1))
    base = OracleScene(_scene(tmp_path / "a", P_at_support, Q), 0.01, counting=True).pair_profile(
        np.zeros(1, np.float32))[1][0]
    # P's vertex 0 at (0,1,1): the climb scans the row of (0,1,1) up to its
    # first better neighbour, which here is the support (1,1,1), and moves once
    P = _box(0.0, 1.0, (0, 1, 1))
    codes, flops = OracleScene(_scene(tmp_path / "b", P, Q), 0.01, counting=True).pair_profile(
        np.zeros(1, np.float32))
    assert codes.tolist() == [0]
    V = P[0]
    row0 = P[3][P[1][0]:P[1][0] + P[2][0]]
    tgt = int(np.flatnonzero(np.all(V == np.float32(1.0), axis=1))[0])
    S = Q[0][0].astype(np.float64) - V[0].astype(np.float64)
    better = [int(v) for v in row0 if V[v] @ S > V[0] @ S]
    j = row0.tolist().index(better[0])
    assert better[0] == tgt and np.argmax(V @ S) == tgt
    deg_support = int(P[2][tgt])
    # Q's start vertex stays its support, so its climb costs 5 (1 + deg Q0)
    assert np.argmax(-(Q[0] @ S)) == 0
    # climb of P: 5 at the start + 5 (j + 1) up to the move + 5 deg(support)
    # scanning the support's row without improvement; everything else as in
    # the closed form
    expect = 143 + 5 * (int(Q[2][0])) + 5 * (j + 1) + 5 * deg_support
    assert int(flops[0]) == expect, (int(flops[0]), expect, base)
