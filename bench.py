"""Benchmark: exact edge collision checks on one MI355X.

Workload (the edge-checking configuration): 6-DOF arm (modified DH, a=0.8,
alpha cycling 0,+pi/2,-pi/2; box links 0.8x0.15x0.15), 16 obstacles that are
convex hulls of 12 points in balls of radius U[0.2,0.5]; 16,384 collision-free
roadmap vertices drawn uniformly in [-pi,pi]^6; the undirected k=10
nearest-neighbour edges; every edge discretised with the reference's rule
testnum = (int)(|qe-qs|_2/dq)+1, dq=0.01 (prm/src/lib/robotspace/
robotconfigspace.cpp:660-673).

One step = deciding every edge (gprm_indicator2_dev, inputs resident in HBM).
`value` counts reference-equivalent checks: the sum of testnum, i.e. the
configuration checks the reference's indicator2 performs for the same job.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "exact collision checks/s on 1 MI355X; PRM time-to-path, surrogate vs plain"
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
SEED = 2024
NDOF = 6
M_PAIRS = 0
# HBM bytes per sva_kernel<1> launch of this workload (FETCH_SIZE + WRITE_SIZE,
# each scaled by the same counter's reading for a known-size copy kernel),
# from profiles/sva_kernel_edges_wl2_traffic.md; None until measured
TRAFFIC_BYTES_PER_LAUNCH = None


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def free_vertices(space, n, seed):
    """n collision-free configurations, uniform in [-pi,pi]^6 by rejection
    (decided by the exact checker)."""
    import torch

    rng = np.random.default_rng(seed)
    out, have = [], 0
    while have < n:
        C = rng.uniform(-np.pi, np.pi, size=(NDOF, 32768)).astype(np.float32)
        q = torch.from_numpy(C).cuda()
        r = torch.zeros(C.shape[1], dtype=torch.int32, device="cuda")
        space.indicator_dev(q, r)
        keep = (r == 0).cpu().numpy()
        out.append(C[:, keep])
        have += int(keep.sum())
    return np.concatenate(out, axis=1)[:, :n].copy()


def knn_edges(V, k):
    """Undirected k-nearest-neighbour edge set (i<j, unique) of the columns of
    V, exact squared distances in float64 on the GPU."""
    import torch

    X = torch.from_numpy(V.T.astype(np.float64)).cuda()
    n = X.shape[0]
    nbr = []
    for i0 in range(0, n, 2048):
        xb = X[i0:i0 + 2048]
        d2 = ((xb[:, None, :] - X[None, :, :]) ** 2).sum(-1)
        idx = torch.arange(xb.shape[0], device="cuda")
        d2[idx, idx + i0] = float("inf")
        nbr.append(torch.topk(d2, k, dim=1, largest=False).indices.cpu().numpy())
    nbr = np.concatenate(nbr)
    I = np.repeat(np.arange(n), k)
    J = nbr.ravel()
    a, b = np.minimum(I, J), np.maximum(I, J)
    key = np.unique(a.astype(np.int64) * n + b)
    return (key // n).astype(np.int64), (key % n).astype(np.int64)


def flops_per_pair_test(scene_dir, QS, QE, testnum, rng, nsamp=3000):
    """Mean algorithmic FP32 operations of one separating-vector pair test,
    split by outcome, from the oracle's counting build over the pair tests the
    reference runs for random samples of this job."""
    from oracle.oracle import OracleScene

    o = OracleScene(scene_dir, 0.01, counting=True)
    sep, col = [], []
    for k in rng.integers(0, QS.shape[1], size=nsamp):
        j = int(rng.integers(0, testnum[k]))
        q = o.edge_sample(QS, QE, int(k), j, int(testnum[k]))
        codes, flops = o.pair_profile(q)
        sep.extend(flops[codes == 0].tolist())
        col.extend(flops[codes != 0].tolist())
    return {
        "separated": float(np.mean(sep)) if sep else 0.0,
        "colliding": float(np.mean(col)) if col else 0.0,
        "n_separated": len(sep),
        "n_colliding": len(col),
    }


def cpu_baseline(scene_dir, QS, QE, testnum, target_s, rng):
    """The reference's own CPU indicator2 (oracle/_ref, its host loop), or the
    oracle when the reference build is absent, on a bounded sample of edges."""
    from oracle import oracle as orc

    nthreads = max(1, min(16, os.cpu_count() or 1))
    if orc.ref_available():
        pool = orc.RefPool(scene_dir, nthreads)
        run = pool.indicator2
        kind = "reference"
    else:
        o = orc.OracleScene(scene_dir)
        run = lambda a, b: o.indicator2(a, b, nthreads=nthreads)[0]
        kind = "port"
    perm = rng.permutation(QS.shape[1])
    cal = perm[:max(64, 16 * nthreads)]
    t0 = time.perf_counter()
    run(QS[:, cal].copy(), QE[:, cal].copy())
    dt = time.perf_counter() - t0
    rate = testnum[cal].sum() / max(dt, 1e-6)
    # grow the sample to about target_s seconds of CPU work
    cum = np.cumsum(testnum[perm])
    n = int(np.searchsorted(cum, rate * target_s)) + 1
    n = max(len(cal), min(n, len(perm)))
    sub = perm[:n]
    t0 = time.perf_counter()
    res = run(QS[:, sub].copy(), QE[:, sub].copy())
    dt = time.perf_counter() - t0
    samples = int(testnum[sub].sum())
    return {
        "value": samples / dt,
        "unit": "checks/s",
        "cores": nthreads,
        "kind": kind,
        "sample": "%d of the %d edges (%d reference samples), reference indicator2 host loop, %.1f s"
                  % (n, QS.shape[1], samples, dt),
    }, sub, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--verify-edges", type=int, default=3000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the timed steps, run a plain elementwise kernel that reads and writes "
                         "MIB MiB once each (the known-size kernel that scales HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    import gpuprm_amd

    rng = np.random.default_rng(SEED)
    scene_dir = tempfile.mkdtemp(prefix="gprm_bench_")
    gpuprm_amd.generate_scene(2, SEED, scene_dir)
    space = gpuprm_amd.RobotConfigspace(scene_dir, 0.01, 1 << 20)
    space.init(0, 1)
    log("scene ready; sampling 16384 free roadmap vertices")
    V = free_vertices(space, 16384, SEED)
    I, J = knn_edges(V, 10)
    QS = V[:, I].copy()
    QE = V[:, J].copy()
    E = QS.shape[1]
    global M_PAIRS
    M_PAIRS = int(np.fromfile(os.path.join(scene_dir, "pairs/M.bin"), np.int32)[0])
    qs = torch.from_numpy(QS).cuda()
    qe = torch.from_numpy(QE).cuda()
    res = torch.zeros(E, dtype=torch.int32, device="cuda")
    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log("copy calibration: neg kernel over %d MiB read + %d MiB written"
            % (args.copy_calibration, args.copy_calibration))
        del src, dst
    log("%d edges; warmup %d" % (E, args.warmup))
    for _ in range(args.warmup):
        st = space.indicator2_dev(qs, qe, res)
    torch.cuda.synchronize()
    space.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = space.indicator2_dev(qs, qe, res)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = space.profile_read()
    space.profile_enable(False)
    samples = int(st.samples_total)
    ms_per_step = 1000.0 * dt / args.steps
    value = samples * args.steps / dt
    gres = res.cpu().numpy()
    log("timed %d steps: %.3f ms/step, %.3g reference-equivalent checks/s" % (args.steps, ms_per_step, value))

    from oracle.oracle import OracleScene

    o = OracleScene(scene_dir, 0.01)
    testnum, tot = o.edge_testnum(QS, QE)
    assert int(tot) == samples, (tot, samples)

    # decisions of a random subset of edges against the oracle
    vsub = rng.choice(E, size=min(args.verify_edges, E), replace=False)
    want, _ = o.indicator2(QS[:, vsub].copy(), QE[:, vsub].copy(), nthreads=min(16, os.cpu_count() or 1))
    mismatches = int(((gres[vsub] != 0) != (want != 0)).sum())
    log("verified %d edges against the oracle: %d mismatches" % (len(vsub), mismatches))

    # roofline of the dominant kernel (separating-vector): algorithmic FP32
    # operations per launch over its HIP-event duration
    fpp = flops_per_pair_test(scene_dir, QS, QE, testnum, rng)
    flops = prof.pairs_separated * fpp["separated"] + prof.pairs_colliding * fpp["colliding"]
    sva_s = prof.sva_ms / 1000.0
    achieved = flops / sva_s / 1e12 if sva_s > 0 else 0.0
    roofline = {
        # compute-bound on the FP32 pipe; the test is scalar geometry, so the
        # FP32 vector ALU (same 157.3 TFLOP/s published peak) does the work
        "bound": "mfma",
        "pipe": "fp32 valu",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "traffic": TRAFFIC_BYTES_PER_LAUNCH,
        "kernel": "sva_kernel<1> (separating-vector, edge mode)",
        "launches_per_step": prof.sva_launches / args.steps,
        "avg_launch_ms": prof.sva_ms / max(prof.sva_launches, 1),
        "flops_per_launch": flops / max(prof.sva_launches, 1),
        "flops_per_pair_test": {"separated": round(fpp["separated"], 1),
                                "colliding": round(fpp["colliding"], 1)},
        "pair_tests_per_step": (prof.pairs_separated + prof.pairs_colliding) / args.steps,
        "iterations_per_step": prof.iterations / args.steps,
        "bound_decided_per_step": prof.bound_separated / args.steps,
    }

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (about %.0f s)" % args.cpu_seconds)
        cpu, csub, cres = cpu_baseline(scene_dir, QS, QE, testnum, args.cpu_seconds, rng)
        mismatches += int(((gres[csub] != 0) != (cres != 0)).sum())
        log("cpu baseline %.3g checks/s (%s, %d threads)" % (cpu["value"], cpu["kind"], cpu["cores"]))

    # configuration checks run to a decision: all pairs free, or a collision found
    evaluated = (prof.pairs_separated / max(M_PAIRS, 1) + prof.pairs_colliding) / max(args.steps, 1)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "edge checks: 6-DOF arm, 16 hulls of 12 points, 16384 free roadmap vertices, "
                        "k=10 NN undirected edges, dq=0.01 (reference L2 sample rule)",
            "edges": E,
            "samples_per_step": samples,
            "max_testnum": int(st.max_testnum),
        },
        "evaluated_checks_per_s": evaluated * args.steps / dt,
        "edges_per_s": E * args.steps / dt,
        "blocked_edges": int((gres != 0).sum()),
        "decision_mismatches": mismatches,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
