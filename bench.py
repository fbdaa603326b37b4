"""Benchmark of the MI355X mesh stage.

Workload: the synthetic 58k-atom assembly (14 copies of a seeded 4.1k-atom
random-walk globule in two stacked 7-fold rings, ring radius 55 A, rings
40 A apart), SES probe 1.4 A, grid spacing 0.25 A, brick 10, exact mode.
A step is one whole mesh stage on the device (classification, gather,
sampling, crossing edges, f64 refinement, cell pass and assembly) with the
arrangement resident in HBM; the arrangement build is excluded and
reported separately.

Prints one JSON line: metric "s per SES mesh (58k atoms, 0.25 A, exact)".
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "s per SES mesh (58k atoms, 0.25 Å, exact) + batched molecules/s vs plain GPU"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the mesh stage shards by "
              "brick ranges across GPUs only in design)", file=sys.stderr)
        sys.exit(2)
    import sesmesh_amd

    xyzr = sesmesh_amd.synth_rings(seed=2)
    probe, spacing, brick = 1.4, 0.25, 10
    t0 = time.perf_counter()
    arr = sesmesh_amd.Arrangement(xyzr, probe)
    t_arr = time.perf_counter() - t0
    dev = sesmesh_amd.Device(0)
    res = sesmesh_amd.Resident(dev, arr)
    for _ in range(args.warmup):
        m = res.mesh(spacing, brick)
        del m
    times = []
    phase = []
    for _ in range(args.steps):
        t = time.perf_counter()
        m = res.mesh(spacing, brick)
        times.append(time.perf_counter() - t)
        phase.append(m.times.as_dict())
        info = m.info
        del m
    total = sum(times)
    s_per_mesh = total / args.steps
    out = {
        "metric": METRIC, "value": s_per_mesh, "unit": "s/mesh", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1000.0 * s_per_mesh, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 sampling + f64 refinement",
        "data": "synthetic (seeded random-walk globules)",
        "config": {"workload": "58k-atom two-ring assembly, SES probe 1.4 A, spacing 0.25 A, brick 10, exact",
                   "atoms": int(len(xyzr)), "surface_bricks": int(info.surface_bricks),
                   "vertices": int(info.vertices), "triangles": int(info.triangles)},
        "best_s": min(times), "median_s": float(np.median(times)),
        "phases_ms": {k: float(np.median([p[k] for p in phase])) for k in phase[0]},
        "arrangement_s": t_arr,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
