"""Regenerate the committed multi-scene tuning table (tuning_mi355x.json).

    python -m gpuprm_amd.tune [--reps 5] [--out gpuprm_amd/tuning_mi355x.json]

Runs on one GPU.  The workload is the multi-scene benchmark's: 64 seeded
scenes (7-DOF arm, 8..128 hulls of 8..24 points), 1k-64k uniform
configurations each.  gprm_batch_autotune times every kernel variant x
pairs-per-unit per vertex-count bucket (HIP events around the bucket's own
launch, one warm-up, the median of `reps` timed runs); the tuning runs twice,
independently, and the file records both passes and whether they agree.  The
first pass is the table SceneBatch applies by default.
"""
from __future__ import annotations

import argparse
import json
import os
import tempfile

import numpy as np

WORKLOAD = "64 scenes: 7-DOF arm, 8..128 hulls of 8..24 points, 1k-64k configurations each"


def multi_scene_workload(seed: int = 2024):
    """(spaces, scene dirs, sizes, device configuration tensors, result
    tensors) of the 64-scene workload, scene s seeded 1000 + s and the
    configuration counts and values from numpy's generator at seed + 5."""
    import torch

    import gpuprm_amd

    rng = np.random.default_rng(seed + 5)
    spaces, dirs = [], []
    for s in range(64):
        d = tempfile.mkdtemp(prefix="gprm_ms_")
        gpuprm_amd.generate_scene(5, 1000 + s, d)
        sp = gpuprm_amd.RobotConfigspace(d, 0.01, 1024)
        sp.init(0, 1)
        spaces.append(sp)
        dirs.append(d)
    sizes = [int(x) for x in rng.integers(1000, 65537, 64)]
    qs = [torch.from_numpy(rng.uniform(-np.pi, np.pi, size=(7, n)).astype(np.float32)).cuda() for n in sizes]
    rs = [torch.zeros(n, dtype=torch.int32, device="cuda") for n in sizes]
    return spaces, dirs, sizes, qs, rs


def main() -> None:
    import gpuprm_amd

    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=gpuprm_amd.SceneBatch.TUNING_FILE)
    args = ap.parse_args()
    spaces, _, _, qs, rs = multi_scene_workload()
    passes = []
    for _ in range(2):
        b = gpuprm_amd.SceneBatch(spaces, tuning=None)
        passes.append(b.autotune(qs, rs, reps=args.reps))
        b.close()
    import torch

    doc = {
        "device": torch.cuda.get_device_name(0) + " (gfx950)",
        "workload": WORKLOAD,
        "method": "gprm_batch_autotune: every variant (256x5, 256x4, 1024x4 threads x waves/SIMD) x "
                  "pairs_per_unit in {32, 64, 128, 256} per vertex-count bucket, HIP events around the "
                  "bucket's own launch, one warm-up, median of %d timed runs; two independent passes" % args.reps,
        "command": "python -m gpuprm_amd.tune --reps %d" % args.reps,
        "buckets": passes[0],
        "second_pass": passes[1],
        "passes_agree": passes[0] == passes[1],
    }
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print(json.dumps(doc, indent=1))


if __name__ == "__main__":
    main()
