"""Benchmark of the Caelix hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lighttree]

lighttree  full light-tree rebuild (bounds, Morton, 4 onesweep passes, Karras, aggregate, depth scans,
           4-wide collapse, header) over 4,194,304 emissive brick leaves: half uniform in an 8192^3
           box, half in 64 Gaussian clusters (sigma 32 voxels), emission log-uniform in [0.01, 100].
           One step = one rebuild from leaves resident in HBM. value = Mprims/s (leaves per second).

Prints one JSON line. The roofline object is for the dominant kernel of a step (largest share of
the step's time), from HIP events around each launch on the stream it runs on; its algorithmic bytes
per leaf are listed in STAGE_BYTES. cpu_baseline times the oracle (the sequential C++ restatement of
the reference's kernels) on a bounded sample of the same workload on the host, single-threaded.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
for _p in (REPO, REPO / "caelix-renderer_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per leaf of each light-tree stage (reads + writes the stage must make):
#   bounds 24 (box lo/hi read), morton 32 (centre read 12 + key/value write 8 + counter/cover clears 8
#   + ...), sort pass 16 (key/value read + write), tree level 1 = Karras 20 + aggregate 148 (the
#   fused bottom-up build does both; the upper levels finish the few nodes that span level-1 blocks),
#   gather (leaf summaries into sorted order: value 4 + leaf 48 + write 48 = 100, outside the 520
#   canonical bytes), cover+flags scans 16, wide 212 (224-byte nodes for n/2 nodes = 112, slot summaries 96, leaf
#   parent 4).
STAGE_BYTES = {
    "lt_bounds": 24.0,
    "lt_morton": 32.0,
    "sort_pass": 16.0,
    "lt_tree_l1": 168.0,
    "lt_cover_scan": 8.0,
    "lt_flags_scan": 8.0,
    "lt_wide": 212.0,
}
BUILD_BYTES_PER_LEAF = 520.0


def lighttree_bench(args) -> dict:
    import numpy as np
    import torch

    import caelix_amd
    from caelix_amd import synthetic

    n = 1 << 22
    leaves = synthetic.light_leaves(n, 2024, extent=8192.0, clusters=64, sigma=32.0)
    builder = caelix_amd.LightBvhBuilder(n)
    builder.upload(leaves)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        builder.build(stream)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        builder.build(stream)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.steps
    out = builder.read()
    assert int(out["header"]["leaf_count"]) == n and int(out["state"]["node_overflow"]) == 0

    # Per-kernel durations (HIP events around each launch), after the timed region.
    caelix_amd.profile_enable(True)
    caelix_amd.profile_reset()
    profile_steps = max(3, min(args.steps, 10))
    for _ in range(profile_steps):
        builder.build(stream)
    torch.cuda.synchronize()
    stages = caelix_amd.profile_read()
    caelix_amd.profile_enable(False)
    per_stage = {name: {"avg_ms": total / launches, "launches_per_step": launches / profile_steps}
                 for name, (total, launches) in stages.items()}
    dominant = max(stages, key=lambda k: stages[k][0])
    avg_ms = stages[dominant][0] / stages[dominant][1]
    achieved = STAGE_BYTES.get(dominant, 0.0) * n / (avg_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": dominant,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
    }
    pipeline_gbs = BUILD_BYTES_PER_LEAF * n / (ms * 1e-3) / 1e9
    value = n / (ms * 1e-3) / 1e6
    result = {
        "metric": "light-tree build Mprims/s (4,194,304 emissive leaves, full rebuild)",
        "value": round(value, 2),
        "unit": "Mprims/s",
        "ms_per_step": round(ms, 4),
        "dtype": "f32+u32",
        "config": {"workload": "light-tree rebuild, 4,194,304 leaves (half uniform in 8192^3, half in 64 "
                               "clusters sigma 32), u32 Morton keys", "leaves": n},
        "roofline": roofline,
        "pipeline_roofline": {"bytes_per_leaf": BUILD_BYTES_PER_LEAF, "achieved": round(pipeline_gbs, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(pipeline_gbs / HBM_PEAK_GBS, 4)},
        "stages": {k: {kk: round(vv, 5) for kk, vv in v.items()} for k, v in per_stage.items()},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = lighttree_cpu_baseline(leaves)
    return result


def lighttree_cpu_baseline(leaves) -> dict:
    import numpy as np

    import oracle
    from caelix_amd import layout

    sample = 1 << 20
    sub = leaves[:sample]
    state = np.zeros(1, layout.STATE)
    state["building"] = 1
    state["leaves_found"] = sample
    node_capacity = layout.node_capacity_for(sample)
    oracle.load()
    runs = 0
    t0 = time.perf_counter()
    while True:
        oracle.light_bvh_build(sub.view(np.uint8), sample, node_capacity, state.view(np.uint8))
        runs += 1
        if time.perf_counter() - t0 >= 10.0:
            break
    seconds = (time.perf_counter() - t0) / runs
    return {"value": round(sample / seconds / 1e6, 3), "unit": "Mprims/s", "cores": 1, "kind": "port",
            "sample": f"oracle light-tree build of the first {sample} leaves of the same workload, "
                      f"{runs} runs, single thread"}


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--workload", default="lighttree", choices=["lighttree"])
    parser.add_argument("--no-cpu-baseline", action="store_true")
    args = parser.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the hot path does not span GPUs (see DESIGN.md on sharding); "
              "run it with --gpus 1", file=sys.stderr)
        return 2
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    result = lighttree_bench(args)
    line = {
        "metric": result["metric"],
        "value": result["value"],
        "unit": result["unit"],
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": result["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": result["dtype"],
        "data": "synthetic",
        "config": result["config"],
        "roofline": result["roofline"],
    }
    for key in ("cpu_baseline", "pipeline_roofline", "stages"):
        if key in result:
            line[key] = result[key]
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
