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
configuration checks the reference's indicator2 performs for the same job;
`evaluated_checks_per_s` counts the configurations actually decided (after
early exit).

Further legs (--extras): point checks of the 4-DOF / 8-box configuration
with their roofline and the latency of small host-array calls; 64 scenes in
one graph-replayed batch (the committed tuning table, a fresh tuning run
beside it) with roofline and CPU baseline; the surrogate MLP's training and
inference throughput against the BF16 MFMA peak; the planner's time-to-path on the
7-DOF scene, the 7-DOF configuration as written (200,000 uniform samples),
and plain vs surrogate over 16 tasks of the 6-DOF scene.

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
# the FP32 operations of a wave64 VALU instruction without FMA or packed math
# (the reference's rounding admits neither): 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_NOFMA_TOPS = 39.3
SEED = 2024
NDOF = 6
M_PAIRS = 0
# Counter figures of the dominant kernel on this workload (HBM bytes per
# launch from FETCH_SIZE / WRITE_SIZE, scaled by a known-size copy; VALU
# instruction counts), measured by rocprofv3 on a build of the sources whose
# hash the file records; see profiles/sva_kernel_edges_wl2_traffic.md.
PROFILE_SIDECAR = os.path.join(ROOT, "profiles", "sva_kernel_edges_wl2.json")
PROFILED_SOURCES = ("gpuprm_amd/csrc/kernels.hip", "gpuprm_amd/csrc/gprm_api.cpp", "gpuprm_amd/csrc/runtime.h")


def sources_sha16():
    """Hash of the sources that define the separating-vector kernel and its
    launches (the counter sidecar holds the hash of the build it measured)."""
    import hashlib

    h = hashlib.sha256()
    for f in PROFILED_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def profile_sidecar():
    """The counter figures if they were measured on these sources, else None."""
    try:
        with open(PROFILE_SIDECAR) as f:
            p = json.load(f)
    except (OSError, ValueError):
        return None
    return p if p.get("sources_sha16") == sources_sha16() else None


def host_cpu():
    """Model, sockets, cores and threads of the host (lscpu), for the CPU baseline."""
    import subprocess

    info = {"logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    return info


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


def _cpu_pool(scene_dir, nthreads):
    from oracle import oracle as orc

    if orc.ref_available():
        return orc.RefPool(scene_dir, nthreads), "reference"
    o = orc.OracleScene(scene_dir)

    class _Port:
        def indicator_batch(self, Q):
            return o.indicator_batch(Q, nthreads=nthreads)

    return _Port(), "port"


def flops_per_check(scene_dir, Q, nsamp, rng):
    """Mean algorithmic FP32 operations of one configuration check (the
    reference's pair loop to the first collision, separating-vector tests
    only), from the oracle's counting build on nsamp of the configurations."""
    from oracle.oracle import OracleScene

    o = OracleScene(scene_dir, 0.01, counting=True)
    lo = np.fromfile(os.path.join(scene_dir, "mins.bin"), np.float32)
    hi = np.fromfile(os.path.join(scene_dir, "maxs.bin"), np.float32)
    idx = rng.choice(Q.shape[1], size=min(nsamp, Q.shape[1]), replace=False)
    tot = 0.0
    for k in idx:
        q = np.ascontiguousarray(Q[:, k])
        if np.any(q < lo) or np.any(q > hi):
            continue  # out of bounds: no pair test
        tot += float(o.pair_profile(q)[1].sum())
    return tot / len(idx)


def latency(sp, Q, QS, QE):
    """Microseconds per host-array call of the calls a reference-style planner
    makes: indicator(q) of one configuration (prmsolver.cpp:2611), indicator
    of 8, and indicator2 of 8 edges (through ctypes, about 1 us of it)."""
    q1 = np.ascontiguousarray(Q[:, 0])
    Q8 = np.ascontiguousarray(Q[:, :8])
    S8, E8 = np.ascontiguousarray(QS[:, :8]), np.ascontiguousarray(QE[:, :8])
    r8 = np.zeros(8, np.int32)
    out = {}
    for name, fn, reps in (("us_indicator1", lambda: sp.indicator(q1), 400),
                           ("us_indicator_8", lambda: sp.indicator(Q8.ravel(), r8, 8, 8), 200),
                           ("us_indicator2_8_edges", lambda: sp.indicator2(S8.ravel(), E8.ravel(), r8, 8, 8), 100)):
        for _ in range(10):
            fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        out[name] = 1e6 * (time.perf_counter() - t0) / reps
    return out


def leg_points(steps):
    """Point checks: 4-DOF arm, 8 boxes, 65,536 configurations (one
    gprm_indicator_dev call per step), with the CPU path on a bounded sample;
    the roofline of the call (reference FP32 operations per check x
    checks/s) and the latency of small host-array calls."""
    import torch

    import gpuprm_amd
    from oracle.oracle import OracleScene

    d = tempfile.mkdtemp(prefix="gprm_pts_")
    gpuprm_amd.generate_scene(1, SEED, d)
    sp = gpuprm_amd.RobotConfigspace(d, 0.01, 1 << 16)
    sp.init(0, 1)
    rng = np.random.default_rng(SEED + 1)
    Q = rng.uniform(-np.pi, np.pi, size=(4, 65536)).astype(np.float32)
    q = torch.from_numpy(Q).cuda()
    r = torch.zeros(65536, dtype=torch.int32, device="cuda")
    for _ in range(3):
        sp.indicator_dev(q, r)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sp.indicator_dev(q, r)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    # the same calls, HIP events around the separating-vector launches
    sp.profile_enable(True)
    for _ in range(steps):
        sp.indicator_dev(q, r)
    torch.cuda.synchronize()
    prof = sp.profile_read()
    sp.profile_enable(False)
    got = r.cpu().numpy()
    sub = rng.choice(65536, 8192, replace=False)
    mism = int((got[sub] != OracleScene(d).indicator_batch(Q[:, sub].copy(), nthreads=min(16, os.cpu_count() or 1))).sum())
    nthreads = max(1, min(16, os.cpu_count() or 1))
    pool, kind = _cpu_pool(d, nthreads)
    want = pool.indicator_batch(Q)
    mism += int((got != want).sum())
    # the CPU path repeats the same 65536 configurations for about 2 s
    reps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        pool.indicator_batch(Q)
        reps += 1
    cdt = (time.perf_counter() - t0) / reps
    fpc = flops_per_check(d, Q, 4000, rng)
    achieved = 65536 * fpc / dt / 1e12
    QE = np.clip(Q + rng.normal(0, 0.2, size=Q.shape), -np.pi, np.pi).astype(np.float32)
    return {"workload": "4-DOF arm, 8 obstacle boxes, 65536 configurations", "checks_per_s": 65536 / dt,
            "ms_per_call": 1000 * dt, "code_mismatches": mism,
            "sva_ms_per_call": prof.sva_ms / max(1, prof.sva_launches),
            "roofline": {"bound": "valu", "pipe": "fp32 valu", "unit": "TFLOP/s", "peak": PEAK_FP32_TFLOPS,
                         "achieved": round(achieved, 3), "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                         "flops_per_check": round(fpc, 1), "traffic": None,
                         "note": "whole call (kinematics, separating-vector kernel, finalize) over the reference's "
                                 "FP32 operations per check from the oracle's counting build"},
            "latency": latency(sp, Q, Q, QE),
            "cpu": {"value": 65536 / cdt, "unit": "checks/s", "cores": nthreads, "kind": kind,
                    "sample": "the 65536 configurations %d times, %.2f s each" % (reps, cdt)},
            "speedup_vs_cpu": cdt / dt}


def leg_multi_scene(reps):
    """64 scenes (7-DOF arm, 8..128 hulls of 8..24 points), 1k-64k uniform
    configurations each, all checked by one graph-replayed batch (one
    kinematics launch, one persistent launch per vertex-count bucket), shapes
    from the autotuner (run first, untimed)."""
    import torch

    import gpuprm_amd
    from gpuprm_amd.tune import multi_scene_workload

    # the workload the committed tuning table was made on (python -m
    # gpuprm_amd.tune)
    spaces, dirs, sizes, qs, rs = multi_scene_workload(SEED)
    # the committed MI355X table (SceneBatch's default); a fresh tuning run
    # is reported beside it, not applied
    b = gpuprm_amd.SceneBatch(spaces)
    table = b.tuning()
    b.indicator_dev(qs, rs, use_graph=True)  # capture
    torch.cuda.synchronize()
    times, lat = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        done = b.indicator_dev(qs, rs, use_graph=True)
        times.append(time.perf_counter() - t0)
        lat.append(done)
    # exactness of the batch against the single-scene path on four scenes
    mism = 0
    for s in (0, 21, 42, 63):
        r1 = torch.zeros(sizes[s], dtype=torch.int32, device="cuda")
        spaces[s].indicator_dev(qs[s], r1)
        mism += int((r1 != rs[s]).sum().item())
    dt = float(np.median(times))
    L = np.median(np.stack(lat), axis=0)
    info = b.bucket_info()
    # algorithmic work: per scene the reference's FP32 operations per check
    # (oracle counting build, 48 configurations each) x its configurations
    frng = np.random.default_rng(SEED + 6)
    flops = sum(sizes[s] * flops_per_check(dirs[s], qs[s][:, :2048].cpu().numpy(), 48, frng) for s in range(64))
    achieved = flops / dt / 1e12
    # the reference's CPU path on a bounded sample: configurations of 8 of
    # the scenes, all host cores, about 4 s in all (sized from a first
    # 2000-configuration call)
    nthreads = max(1, min(16, os.cpu_count() or 1))
    csum, ct = 0, 0.0
    kind = None
    per = 2000
    for s in range(0, 64, 8):
        pool, kind = _cpu_pool(dirs[s], nthreads)
        Qs = qs[s][:, :per].cpu().numpy().copy()
        t0 = time.perf_counter()
        pool.indicator_batch(Qs)
        ct += time.perf_counter() - t0
        csum += Qs.shape[1]
        if s == 0:
            per = int(min(sizes[s], max(2000, 0.5 * per / max(ct, 1e-6))))
    retune = b.autotune(qs, rs, reps=3)
    out = {"workload": "64 scenes: 7-DOF arm, 8..128 hulls of 8..24 points, 1k-64k configurations each",
           "configs": int(sum(sizes)), "checks_per_s": sum(sizes) / dt, "ms_per_batch": 1000 * dt,
           "scene_done_ms_p50": float(np.percentile(L, 50)), "scene_done_ms_p99": float(np.percentile(L, 99)),
           "buckets": [{"vertices": [x["min_vertices"], x["max_vertices"]], "scenes": x["scenes"],
                        "shape": x["shape"]} for x in info if x["scenes"]],
           "tuning": table, "tuning_source": "gpuprm_amd/tuning_mi355x.json (committed)",
           "retune": {"table": retune, "reps": 3, "statistic": "median", "same_as_committed": retune == table},
           "roofline": {"bound": "valu", "pipe": "fp32 valu", "unit": "TFLOP/s", "peak": PEAK_FP32_TFLOPS,
                        "achieved": round(achieved, 3), "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                        "flops_per_check": round(flops / sum(sizes), 1), "traffic": None,
                        "note": "graph replay (kinematics, one separating-vector launch per bucket, finalize) "
                                "over the reference's FP32 operations per check, oracle counting build"},
           "cpu_baseline": {"value": csum / ct, "unit": "checks/s", "cores": nthreads, "kind": kind,
                            "sample": "%d configurations of 8 of the 64 scenes, %.1f s" % (csum, ct)},
           "speedup_vs_cpu": (sum(sizes) / dt) / (csum / ct),
           "mismatches_vs_single_scene": mism}
    b.apply_tuning(table)
    # edge sets: per scene the k=10 nearest-neighbour edges among its first
    # 2048 configurations, all scenes in one graph-replayed edge batch
    QS, QE = [], []
    for s in range(64):
        V = qs[s][:, :2048].cpu().numpy()
        I, J = knn_edges(V, 10)
        QS.append(torch.from_numpy(V[:, I].copy()).cuda())
        QE.append(torch.from_numpy(V[:, J].copy()).cuda())
    er = [torch.zeros(x.shape[1], dtype=torch.int32, device="cuda") for x in QS]
    b.indicator2_dev(QS, QE, er, use_graph=True)  # plan and capture
    torch.cuda.synchronize()
    times, lat = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        done, samples = b.indicator2_dev(QS, QE, er, use_graph=True)
        times.append(time.perf_counter() - t0)
        lat.append(done)
    emism = 0
    for s in (0, 21, 42, 63):
        r1 = torch.zeros(QS[s].shape[1], dtype=torch.int32, device="cuda")
        spaces[s].indicator2_dev(QS[s], QE[s], r1)
        emism += int(((r1 != 0) != (er[s] != 0)).sum().item())
    # the same edges scene by scene (64 single-scene calls), for comparison
    r1 = [torch.zeros(x.shape[1], dtype=torch.int32, device="cuda") for x in QS]
    for s in range(64):
        spaces[s].indicator2_dev(QS[s], QE[s], r1[s])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(64):
        spaces[s].indicator2_dev(QS[s], QE[s], r1[s])
    torch.cuda.synchronize()
    single_ms = 1000 * (time.perf_counter() - t0)
    dt = float(np.median(times))
    L = np.median(np.stack(lat), axis=0)
    nedges = int(sum(x.shape[1] for x in QS))
    out["edges"] = {"workload": "k=10 nearest-neighbour edges among each scene's first 2048 configurations, "
                                "dq=0.01 (reference L2 sample rule)",
                    "edges": nedges, "samples": int(samples), "ms_per_batch": 1000 * dt,
                    "checks_per_s": samples / dt, "edges_per_s": nedges / dt,
                    "blocked_edges": int(sum(int((x != 0).sum().item()) for x in er)),
                    "scene_done_ms_p50": float(np.percentile(L, 50)),
                    "scene_done_ms_p99": float(np.percentile(L, 99)),
                    "decision_mismatches_vs_single_scene": emism,
                    "scene_by_scene_ms": single_ms}
    return out


def _tasks(o, nd, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        a = rng.uniform(-np.pi, np.pi, nd).astype(np.float32)
        b = rng.uniform(-np.pi, np.pi, nd).astype(np.float32)
        if o.indicator1(a) == 0 and o.indicator1(b) == 0 and np.linalg.norm(a - b) > 4.0:
            out.append((a, b))
    return out


def _path_ok(o, path):
    """The path re-decided by the oracle: every node, and every edge in the
    orientation the planner checked it (gpuprm_amd.canonical_edges)."""
    import gpuprm_amd

    if path is None:
        return False
    S, E = gpuprm_amd.canonical_edges(path[:-1], path[1:])
    res, _ = o.indicator2(S.T.copy(), E.T.copy(), nthreads=min(16, os.cpu_count() or 1))
    return bool(np.all(res == 0) and np.all(o.indicator_batch(path.T) == 0))


def _prm_run(sp, o, a, b, verify=True, **kw):
    r, path = sp.prm_solve(a, b, **kw)
    return {"found": r.found, "valid": _path_ok(o, path) if verify else None, "ms": r.ms_total, "steps": r.steps,
            "nodes": r.nodes_left + r.nodes_right, "edges_checked": r.edges_checked,
            "edge_samples": r.edge_samples, "point_checks": r.point_checks,
            # reference-equivalent checks (edge samples + point checks) per second of the whole run
            "checks_per_s": (r.edge_samples + r.point_checks) / max(1e-9, r.ms_total / 1000),
            "ms_device": {"sampling": round(r.ms_sampling, 2), "near": round(r.ms_near, 2),
                          "edges": round(r.ms_edges, 2), "insert": round(r.ms_insert, 2),
                          "model": round(r.ms_model, 2)},
            "ms_wait": round(r.ms_wait, 2), "ms_final": round(r.ms_final, 2)}


def leg_uniform_prm(sp3, o3, start, goal):
    """The 7-DOF configuration as written: 200,000 uniform samples in
    [-pi, pi]^7, every pair closer than 0.5 (L2) an edge candidate, start and
    goal added; all decided by the exact checks, then the components."""
    import torch

    import gpuprm_amd

    from scipy import sparse
    from scipy.sparse import csgraph

    rng = np.random.default_rng(SEED + 17)
    N, D = 200000, 0.5
    t0 = time.perf_counter()
    Q = torch.from_numpy(rng.uniform(-np.pi, np.pi, size=(7, N)).astype(np.float32)).cuda()
    codes = torch.zeros(N, dtype=torch.int32, device="cuda")
    sp3.indicator_dev(Q, codes)
    free = torch.nonzero(codes == 0).flatten()
    V = torch.cat([Q[:, free], torch.from_numpy(np.stack([start, goal], axis=1)).cuda()], dim=1).contiguous()
    n = V.shape[1]
    torch.cuda.synchronize()
    t_points = time.perf_counter() - t0
    t0 = time.perf_counter()
    lo = np.full(7, -np.pi, np.float32)
    qi, ni, _ = gpuprm_amd.radius_neighbours(V, V, lo, -lo, D, cap=1 << 24)
    keep = qi < ni
    I, J = qi[keep], ni[keep]
    t_near = time.perf_counter() - t0
    t0 = time.perf_counter()
    Vh = V.cpu().numpy()
    S, E = gpuprm_amd.canonical_edges(Vh[:, I].T, Vh[:, J].T)
    res = torch.zeros(max(1, len(I)), dtype=torch.int32, device="cuda")
    samples = 0
    if len(I):
        st = sp3.indicator2_dev(torch.from_numpy(S.T.copy()).cuda(), torch.from_numpy(E.T.copy()).cuda(), res)
        samples = int(st.samples_total)
    okE = res[:len(I)].cpu().numpy() == 0
    t_edges = time.perf_counter() - t0
    A = sparse.coo_matrix((np.ones(int(okE.sum())), (I[okE], J[okE])), shape=(n, n))
    ncomp, lab = csgraph.connected_components(A, directed=False)
    sizes = np.bincount(lab)
    return {"workload": "7-DOF arm, 32 hulls of 16 points: 200000 uniform samples, radius 0.5 (L2)",
            "samples": N, "free": int(free.numel()), "candidate_edges": int(len(I)), "free_edges": int(okE.sum()),
            "edge_samples": samples, "components": int(ncomp), "largest_component": int(sizes.max()),
            "start_goal_connected": bool(lab[n - 2] == lab[n - 1]),
            "expected_neighbours_per_sample": round(float(len(I) * 2 / max(1, n)), 4),
            "ms": {"points": 1000 * t_points, "near": 1000 * t_near, "edges": 1000 * t_edges}}


PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def leg_mlp():
    """The surrogate MLP's kernels on the 6-DOF scene: training (the four
    launches of a step, batch 32,768, Adam) over 8 epochs of 4M exact labels,
    then inference over the 4M configurations; HIP events on the stream the
    kernels run on (the null stream), algorithmic FLOPs (forward
    2 (6 x 256 + 2 x 256^2 + 256) per configuration, a training step three
    times that), held-out AUC."""
    import torch

    import gpuprm_amd
    from oracle import mlp

    d = tempfile.mkdtemp(prefix="gprm_mlp_")
    gpuprm_amd.generate_scene(2, SEED, d)
    sp = gpuprm_amd.RobotConfigspace(d, 0.01, 1 << 20)
    sp.init(0, 1)
    rng = np.random.default_rng(SEED + 11)
    N, ntest, B, epochs = 1 << 22, 1 << 16, 32768, 8
    Q = torch.from_numpy(rng.uniform(-np.pi, np.pi, size=(6, N)).astype(np.float32)).cuda()
    codes = torch.zeros(N, dtype=torch.int32, device="cuda")
    sp.indicator_dev(Q, codes)
    model = gpuprm_amd.Surrogate(ndof=6, batch=B, seed=3)
    ntr = N - ntest
    flop = 2 * (6 * 256 + 2 * 256 * 256 + 256)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    steps = 0
    for ep in range(epochs):
        for it in range(ntr // B):
            model.train_step(Q[:, it * B:(it + 1) * B], codes[it * B:(it + 1) * B],
                             lr=4e-3 if ep < epochs - 1 else 1e-3, offset=N, want_loss=False)
            steps += 1
    e1.record()
    torch.cuda.synchronize()
    t_train = e0.elapsed_time(e1) / 1000.0
    pt = torch.zeros(ntest, dtype=torch.float32, device="cuda")
    model.predict(Q[:, ntr:].contiguous(), pt)
    auc = mlp.auc(pt.cpu().numpy(), codes[ntr:].cpu().numpy() != 0)
    P = torch.zeros(N, dtype=torch.float32, device="cuda")
    model.predict(Q, P)
    ts = []
    for _ in range(5):
        e0.record()
        model.predict(Q, P)
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) / 1000.0)
    t_pred = float(np.median(ts))
    train_tf = steps * B * 3 * flop / t_train / 1e12
    pred_tf = N * flop / t_pred / 1e12
    return {"workload": "6-DOF arm scene: 4,194,304 exact labels; model 6 -> 256 -> 256 -> 256 -> 1, BF16 MFMA",
            "train": {"batch": B, "steps": steps, "ms_per_step": 1000 * t_train / steps,
                      "roofline": {"bound": "mfma", "pipe": "bf16 mfma", "unit": "TFLOP/s", "peak": PEAK_BF16_TFLOPS,
                                   "achieved": round(train_tf, 2), "frac": round(train_tf / PEAK_BF16_TFLOPS, 4),
                                   "flops_per_sample": 3 * flop}},
            "predict": {"configs": N, "ms": 1000 * t_pred, "configs_per_s": N / t_pred,
                        "roofline": {"bound": "mfma", "pipe": "bf16 mfma", "unit": "TFLOP/s",
                                     "peak": PEAK_BF16_TFLOPS, "achieved": round(pred_tf, 2),
                                     "frac": round(pred_tf / PEAK_BF16_TFLOPS, 4), "flops_per_config": flop}},
            "auc_heldout": auc}


def leg_prm():
    """Time-to-path of the single-GPU planner on the 7-DOF scene (32 hulls of
    16 points, radius 0.5, 8192 new nodes per step): the three tasks of the
    earlier measurements and six more, one worker (steps in issue order) and
    the reference's two-worker scheme; the configuration as written (200,000
    uniform samples); plain vs surrogate on the 7-DOF scene (a model trained
    on 4M exact labels of it, the same 9 tasks); and plain vs surrogate on the
    6-DOF edge-checking scene with a model trained on 4M exact labels, 16
    tasks.  Every path is re-verified by the oracle."""
    import torch

    import gpuprm_amd
    from oracle import mlp
    from oracle.oracle import OracleScene

    out = {}
    d3 = tempfile.mkdtemp(prefix="gprm_prm3_")
    gpuprm_amd.generate_scene(3, SEED, d3)
    sp3 = gpuprm_amd.RobotConfigspace(d3, 0.01, 1 << 20)
    sp3.init(0, 1)
    o3 = OracleScene(d3)
    tasks = _tasks(o3, 7, 9, SEED + 7)
    kw3 = dict(num=8192, D=0.5, max_nodes=1 << 20, max_steps=400)
    runs = [_prm_run(sp3, o3, a, b, seed=i, workers=1, **kw3) for i, (a, b) in enumerate(tasks)]
    runs2 = [_prm_run(sp3, o3, a, b, seed=i, workers=2, **kw3) for i, (a, b) in enumerate(tasks[:3])]
    out["7dof_time_to_path"] = {
        "workload": "7-DOF arm, 32 hulls of 16 points, D=0.5, 8192 new nodes/step",
        "ms_median": float(np.median([x["ms"] for x in runs[:3]])),
        "ms_median_9": float(np.median([x["ms"] for x in runs])),
        "ms_median_workers2": float(np.median([x["ms"] for x in runs2])),
        "all_paths_valid": all(x["valid"] for x in runs + runs2),
        "runs": runs, "runs_workers2": runs2}
    out["7dof_uniform_as_written"] = leg_uniform_prm(sp3, o3, *tasks[0])

    # the model on the 7-DOF scene too: 4M exact labels, the same 9 tasks
    rng7 = np.random.default_rng(SEED + 19)
    N7, ntest7, B7, ep7 = 1 << 22, 1 << 16, 32768, 8
    Q7 = torch.from_numpy(rng7.uniform(-np.pi, np.pi, size=(7, N7)).astype(np.float32)).cuda()
    codes7 = torch.zeros(N7, dtype=torch.int32, device="cuda")
    sp3.indicator_dev(Q7, codes7)
    model7 = gpuprm_amd.Surrogate(ndof=7, batch=B7, seed=3)
    ntr7 = N7 - ntest7
    for ep in range(ep7):
        for it in range(ntr7 // B7):
            model7.train_step(Q7[:, it * B7:(it + 1) * B7], codes7[it * B7:(it + 1) * B7],
                              lr=4e-3 if ep < ep7 - 1 else 1e-3, offset=N7, want_loss=False)
    pt7 = torch.zeros(ntest7, dtype=torch.float32, device="cuda")
    model7.predict(Q7[:, ntr7:].contiguous(), pt7)
    auc7 = mlp.auc(pt7.cpu().numpy(), codes7[ntr7:].cpu().numpy() != 0)
    modes7 = {"surrogate": {"surrogate": model7, "order_edges": True, "bias_sampling": True, "edge_budget": 4},
              "budget_shortest_no_model": {"edge_budget": 4}}
    res7 = {"plain": runs}
    for k, kw in modes7.items():
        res7[k] = [_prm_run(sp3, o3, a, b, seed=i, workers=1, **kw3, **kw) for i, (a, b) in enumerate(tasks)]
    med7 = {k: float(np.median([x["ms"] for x in v])) for k, v in res7.items()}
    out["surrogate_vs_plain_7dof"] = {
        "workload": "7-DOF arm, 32 hulls of 16 points, D=0.5, 8192 new nodes/step, the 9 tasks above",
        "auc": auc7, "collision_rate": float((codes7 != 0).float().mean().item()),
        "ms_median": med7, "speedup": med7["plain"] / med7["surrogate"],
        "speedup_budget_no_model": med7["plain"] / med7["budget_shortest_no_model"],
        "edge_samples_saved": 1 - sum(x["edge_samples"] for x in res7["surrogate"]) /
        max(1, sum(x["edge_samples"] for x in res7["plain"])),
        "all_paths_valid": all(x["valid"] for v in res7.values() for x in v),
        "runs": {k: v for k, v in res7.items() if k != "plain"}}

    d2 = tempfile.mkdtemp(prefix="gprm_prm2_")
    gpuprm_amd.generate_scene(2, SEED, d2)
    sp2 = gpuprm_amd.RobotConfigspace(d2, 0.01, 1 << 20)
    sp2.init(0, 1)
    o2 = OracleScene(d2)
    rng = np.random.default_rng(SEED + 11)
    N, ntest, B = 1 << 22, 1 << 16, 32768
    Q = torch.from_numpy(rng.uniform(-np.pi, np.pi, size=(6, N)).astype(np.float32)).cuda()
    codes = torch.zeros(N, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sp2.indicator_dev(Q, codes)
    torch.cuda.synchronize()
    t_label = time.perf_counter() - t0
    model = gpuprm_amd.Surrogate(ndof=6, batch=B, seed=3)
    ntr = N - ntest
    epochs = 8
    t0 = time.perf_counter()
    for ep in range(epochs):
        for it in range(ntr // B):
            model.train_step(Q[:, it * B:(it + 1) * B], codes[it * B:(it + 1) * B],
                             lr=4e-3 if ep < epochs - 1 else 1e-3, offset=N, want_loss=False)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    pt = torch.zeros(ntest, dtype=torch.float32, device="cuda")
    model.predict(Q[:, ntr:].contiguous(), pt)
    auc = mlp.auc(pt.cpu().numpy(), codes[ntr:].cpu().numpy() != 0)
    P = torch.zeros(N, dtype=torch.float32, device="cuda")
    model.predict(Q, P)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model.predict(Q, P)
    torch.cuda.synchronize()
    t_pred = time.perf_counter() - t0
    flop = 2 * (6 * 256 + 2 * 256 * 256 + 256)  # one 6-DOF forward, algorithmic (no input padding)
    modes = {"plain": {},
             "surrogate": {"surrogate": model, "order_edges": True, "bias_sampling": True, "edge_budget": 4},
             "budget_shortest_no_model": {"edge_budget": 4}}
    tasks2 = _tasks(o2, 6, 16, SEED + 13)
    sweep = {}
    for num in (1024, 2048, 4096):
        res = {k: [] for k in modes}
        for i, (a, b) in enumerate(tasks2):
            for k, kw in modes.items():
                res[k].append(_prm_run(sp2, o2, a, b, num=num, D=1.0, max_nodes=1 << 18,
                                       max_steps=300, seed=i, **kw))
        med = {k: float(np.median([x["ms"] for x in v])) for k, v in res.items()}
        phases = {k: {ph: float(np.median([x["ms_device"][ph] for x in v])) for ph in v[0]["ms_device"]}
                  for k, v in res.items()}
        sweep[num] = {"ms_median": med, "ms_device_median": phases, "speedup": med["plain"] / med["surrogate"],
                      "speedup_budget_no_model": med["plain"] / med["budget_shortest_no_model"],
                      "edge_samples_saved": 1 - sum(x["edge_samples"] for x in res["surrogate"]) /
                      max(1, sum(x["edge_samples"] for x in res["plain"])),
                      "point_checks_saved": 1 - sum(x["point_checks"] for x in res["surrogate"]) /
                      max(1, sum(x["point_checks"] for x in res["plain"])),
                      "all_found": all(x["found"] for v in res.values() for x in v),
                      "all_paths_valid": all(x["valid"] for v in res.values() for x in v),
                      "runs": res}
    best = {k: min(sweep[n]["ms_median"][k] for n in sweep) for k in modes}
    out["surrogate_vs_plain"] = {
        "workload": "6-DOF arm, 16 hulls of 12 points (edge-checking scene), D=1.0, 16 tasks, "
                    "1024, 2048 and 4096 new nodes/step",
        "labels": N, "label_ms": 1000 * t_label, "train_epochs": epochs, "train_batch": B, "train_s": t_train,
        "train_tflops": epochs * (ntr // B) * B * 3 * flop / t_train / 1e12, "auc": auc,
        "predict_configs_per_s": N / t_pred, "predict_tflops": N * flop / t_pred / 1e12,
        "predict_mfma_frac": N * flop / t_pred / 2.5e15,
        "ms_median": sweep[1024]["ms_median"], "speedup": sweep[1024]["speedup"],
        "speedup_2048": sweep[2048]["speedup"], "speedup_4096": sweep[4096]["speedup"],
        "best_of_num_ms": best, "speedup_best_of_num": best["plain"] / best["surrogate"],
        "edge_samples_saved": sweep[1024]["edge_samples_saved"],
        "point_checks_saved": sweep[1024]["point_checks_saved"],
        "all_paths_valid": all(sweep[n]["all_paths_valid"] for n in sweep), "by_num": sweep}
    out["config1_reference_task"] = leg_config1()
    return out


def leg_config1(seeds: int = 10):
    """The reference's own planner task (prm/src/main.cpp:122-133): the shipped
    4-DOF scene config1, start (0,0,0,0), goal (3,3,3,3), D = 1.0, dq = 0.01,
    at the reference's default of 32 new nodes per step and at 1,024; mean and
    median time-to-path over 10 seeds (the paper's rows are means of 10
    seeds), every path re-verified by the oracle."""
    import gpuprm_amd
    from oracle.oracle import OracleScene

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scenes

    d = tempfile.mkdtemp(prefix="gprm_config1_")
    z = dict(np.load(os.path.join(ROOT, "tests", "golden", "config1.npz")))
    scenes.write_dir(scenes.scene_from_npz(z), d)
    sp = gpuprm_amd.RobotConfigspace(d, 0.01, 1 << 20)
    sp.init(0, 1)
    o = OracleScene(d)
    qs, qg = np.zeros(4, np.float32), np.full(4, 3.0, np.float32)
    res = {}
    for num in (32, 1024):
        runs = [_prm_run(sp, o, qs, qg, num=num, D=1.0, max_nodes=1 << 18, max_steps=10000, seed=s)
                for s in range(seeds)]
        res[num] = {"ms_mean": float(np.mean([x["ms"] for x in runs])),
                    "ms_median": float(np.median([x["ms"] for x in runs])),
                    "checks_mean": float(np.mean([x["edge_samples"] + x["point_checks"] for x in runs])),
                    "steps_mean": float(np.mean([x["steps"] for x in runs])),
                    "all_found": all(x["found"] for x in runs), "all_paths_valid": all(x["valid"] for x in runs),
                    "runs": runs}
    return {"workload": "config1 (4-DOF, the reference's shipped scene), start 0, goal 3, D=1.0, 10 seeds",
            "by_num": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--verify-edges", type=int, default=3000)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the point-check, multi-scene, MLP and planner legs")
    ap.add_argument("--extras", default="points,multi,mlp,prm",
                    help="comma-separated legs after the headline: points, multi, mlp, prm")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the timed steps, run a plain elementwise kernel that reads and writes "
                         "MIB MiB once each (the known-size kernel that scales HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (asked for --gpus %d)"
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
    side = profile_sidecar()
    avg_launch_s = sva_s / max(prof.sva_launches, 1)
    hw = None
    if side:
        # what the hardware executed, from the counters of the same sources
        hw = {"valu_busy": side["valu_busy"], "lane_utilisation": side["lane_utilisation"],
              "executed_fp32_tflops": round(side["fp32_ops_per_launch"] / avg_launch_s / 1e12, 3),
              "executed_frac_of_nofma_valu": round(side["fp32_ops_per_launch"] / avg_launch_s / 1e12
                                                   / VALU_NOFMA_TOPS, 4),
              "hbm_tb_s": round(side["bytes_per_launch"] / avg_launch_s / 1e12, 3)}
    roofline = {
        # compute-bound on the FP32 vector pipe: scalar geometry without FMA
        # (the reference's rounding), so VALU, not the matrix cores
        "bound": "valu",
        "pipe": "fp32 valu",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        # HBM bytes per launch measured by rocprofv3 on these very sources
        # (profiles/sva_kernel_edges_wl2_traffic.md), else unknown
        "traffic": side["bytes_per_launch"] if side else None,
        "traffic_sources_sha16": sources_sha16(),
        "hw": hw,
        "frac_note": "achieved counts the reference algorithm's FP32 operations for the pair tests run "
                     "(oracle counting build); the bound test decides most pairs in far fewer executed "
                     "operations, so it can exceed what the VALU executes (hw.executed_fp32_tflops)",
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
        cpu["host"] = host_cpu()
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
    extras = [] if args.no_extras else [x.strip() for x in args.extras.split(",") if x.strip()]
    if "points" in extras:
        log("point checks")
        line["points"] = leg_points(20)
    if "multi" in extras:
        log("multi-scene batch")
        line["multi_scene"] = leg_multi_scene(10)
    if "mlp" in extras:
        log("surrogate MLP kernels")
        line["mlp"] = leg_mlp()
    if "prm" in extras:
        log("planner")
        line["prm"] = leg_prm()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
