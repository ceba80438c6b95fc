"""Write the golden vectors of tests/golden/*.npz.

Runs the reference's own CPU path (compiled from its sources by
`make -C oracle ref` into oracle/_ref/libref.so) on seeded inputs and stores
inputs and outputs only:

  config1.npz   the reference's shipped scene (prm/config1), 4-DOF
  wl1.npz       generated 4-DOF box scene (workload 1)
  wl2.npz       generated 6-DOF hull scene (workload 2)
  wl3.npz       generated 7-DOF hull scene (workload 3)

Per scene: configurations Q and the reference's indicator(q) codes
(robotconfigspace.cpp:312-346); forward-kinematics frames (kinematics.h:27-37)
of the first configurations; (code, iterations) of
seperating_vector_algorithm (collision4.h:227-332) for every pair of a few
configurations; edges with the reference's indicator2 host-loop results
(robotconfigspace.cpp:389-465,694-705) and its sample count.

Usage (where the reference sources are, under /root/reference):
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle.oracle import RefSpace  # noqa: E402
import scenes  # noqa: E402

CONFIG1 = "/root/reference/prm/config1"
DQ = 0.01


def make(name, scene_dir, seed, nq, nfk, npair_q, nedges, spread):
    rng = np.random.default_rng(seed)
    s = scenes.read_dir(scene_dir)
    nd = int(s["ndof"])
    lo, hi = s["mins"].astype(np.float64), s["maxs"].astype(np.float64)
    span = hi - lo
    # a few percent outside the bounds so check_boundaries is exercised
    Q = rng.uniform(lo - 0.03 * span, hi + 0.03 * span, size=(nq, nd)).astype(np.float32)
    if name == "config1":
        Q[0] = 0.0
        Q[1] = 3.0  # prm/src/main.cpp:128-138 known answers
    ref = RefSpace(scene_dir, DQ, nbuf=1 << 20)
    codes = np.array([ref.indicator1(q) for q in Q], np.int32)
    fk = np.stack([ref.fk(Q[i]) for i in range(nfk)]).astype(np.float32)  # (nfk, nd+1, 4, 3)
    P = len(s["from"])
    pair_code = np.zeros((npair_q, P), np.int32)
    pair_iter = np.zeros((npair_q, P), np.int32)
    for i in range(npair_q):
        for j in range(P):
            c, it = ref.sva(Q[i], int(s["from"][j]), int(s["to"][j]))
            pair_code[i, j] = c
            pair_iter[i, j] = it
    # edges: long random edges, short local edges, and edges shorter than dq
    # (testnum == 1: c1 = 0/0, robotconfigspace.cpp:411-413)
    inb = lambda n: rng.uniform(lo, hi, size=(n, nd)).astype(np.float32)
    n_long = nedges // 3
    n_short = nedges // 3
    n_tiny = nedges - n_long - n_short
    qs1, qe1 = inb(n_long), inb(n_long)
    qs2 = inb(n_short)
    qe2 = np.clip(qs2 + rng.normal(0, spread, size=qs2.shape), lo, hi).astype(np.float32)
    qs3 = inb(n_tiny)
    qe3 = (qs3 + rng.uniform(-0.3, 0.3, size=qs3.shape) * DQ / np.sqrt(nd)).astype(np.float32)
    qe3[: max(1, n_tiny // 4)] = qs3[: max(1, n_tiny // 4)]  # zero-length edges
    QS = np.concatenate([qs1, qs2, qs3]).T.copy()
    QE = np.concatenate([qe1, qe2, qe3]).T.copy()
    before = ref.numthreads_all()
    edge_res = ref.indicator2(QS, QE)
    numthreads = ref.numthreads_all() - before
    out = dict(scenes.scene_to_npz_dict(s))
    out.update(Q=Q.T.copy(), codes=codes, fk=fk, pair_code=pair_code, pair_iter=pair_iter,
               QS=QS, QE=QE, edge_res=edge_res, numthreads=np.int64(numthreads), dq=np.float32(DQ))
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("%s: %d configs (%.1f%% blocked), %d pair checks, %d edges (%d samples) -> %s (%d bytes)"
          % (name, nq, 100 * (codes != 0).mean(), pair_code.size, QS.shape[1], numthreads, path,
             os.path.getsize(path)))


def main():
    import gpuprm_amd

    make("config1", CONFIG1, 11, 4000, 32, 24, 1200, 0.3)
    tmp = tempfile.mkdtemp(prefix="gprm_golden_")
    for wl, nq, npq, ne in ((1, 4000, 24, 1200), (2, 3000, 12, 900), (3, 2000, 8, 600)):
        d = os.path.join(tmp, "wl%d" % wl)
        gpuprm_amd.generate_scene(wl, 2024, d)
        make("wl%d" % wl, d, 100 + wl, nq, 32, npq, ne, 0.2)


if __name__ == "__main__":
    main()
