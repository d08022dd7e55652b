"""Benchmark of the Caelix hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload trace|lighttree|copy|cxw]

trace      (default) DDA + NEE at 1920 x 1080: 1,048,576 resident 8^3 bricks in a 4096^3 world (8-octave
           value-noise terrain, 4 surface bricks per column), bulk hash table at load 0.7, 65,536
           emissive bricks with emission log-uniform in [0.01, 100], 1 spp, 2 bounces, NEE through the
           light tree at the primary and bounce-1 vertices, float32 throughout. One step = one frame
           (the frame seed changes per step). value = Mrays/s: primary + bounce + shadow rays traced
           per second. The light-tree rebuild below, the differentiable form (adjoint / forward
           time at 512x512, 4 spp, 3 bounces, 262,144 bricks) and the batched form (256 chunk scenes,
           batched vs launch-per-scene) are measured in the same run.
lighttree  full light-tree rebuild over 4,194,304 emissive brick leaves (half uniform in an 8192^3
           box, half in 64 Gaussian clusters, sigma 32 voxels, emission log-uniform in [0.01, 100]).
           One step = one rebuild from leaves resident in HBM. value = Mprims/s.
copy       a plain 1 GiB device copy: the calibration run for the HBM byte counters of rocprofv3.
cxw        the 1080p terrain world loaded from a .cxw file: open + raw-deflate decode of every sector
           on the host cores (value = bricks/s), then the GPU scene build from the decoded bricks.

Prints one JSON line. `roofline` is for the dominant kernel of a step, timed with HIP events around
its launches on the stream it runs on. Its algorithmic bytes: for the trace, the demand bytes per ray
(every global load the kernels make, at its size) counted by the oracle's instrumented traversal on a
band of the same frame's pixels, times the rays of the frame; for the light tree, the per-leaf bytes
of STAGE_BYTES. `cpu_baseline` times the oracle (the CPU restatement of the same algorithms) on a
bounded sample of the same workload on the host.
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
#   bounds 24 (box lo/hi read), morton 32 (centre read 12 + key/value write 8 + counter/cover clears
#   8 + partials), sort pass 16 (key/value read + write), tree level 1 = Karras 20 + aggregate 148
#   (the fused bottom-up build does both; the upper levels finish the few nodes that span level-1
#   blocks), cover + flags scans 16, wide 212 (224-byte nodes for n/2 nodes = 112, slot summaries
#   96, leaf parent 4).
STAGE_BYTES = {
    "lt_bounds": 24.0,
    "lt_morton": 32.0,
    "sort_pass": 16.0,
    "sort64_pass": 16.0,
    "lt_tree_l1": 168.0,
    "lt_cover_scan": 8.0,
    "lt_flags_scan": 8.0,
    "lt_wide": 212.0,
}
BUILD_BYTES_PER_LEAF = 520.0
BUILD_BYTES_PER_LEAF_KEYS64 = 580.0  # 7 key-only passes of 16 B + 8 B wider keys in place of 4 pair passes

TRACE_BOUNCES = 2
TRACE_SPP = 1

# HBM bytes per launch of the dominant kernels, measured with rocprofv3 --pmc (FETCH_SIZE, then
# WRITE_SIZE) and scaled by the same counters on a 1 GiB copy; null when not measured.
TRAFFIC_FILE = REPO / "profiles" / "hbm_traffic.json"


def _traffic(kernel: str):
    try:
        entry = json.loads(TRAFFIC_FILE.read_text()).get(kernel)
    except (OSError, ValueError):
        return None
    return int(entry["bytes"]) if entry else None


def _profile(run, steps: int) -> dict:
    """{stage: (total ms, launches)} over `steps` calls of run()."""
    import torch

    import caelix_amd

    caelix_amd.profile_enable(True)
    caelix_amd.profile_reset()
    for i in range(steps):
        run(i)
    torch.cuda.synchronize()
    stages = caelix_amd.profile_read()
    caelix_amd.profile_enable(False)
    return stages


def _time_steps(run, steps: int, warmup: int, stream) -> float:
    import torch

    for i in range(warmup):
        run(i)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for i in range(steps):
        run(warmup + i)
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


# ------------------------------------------------------------------------------------------------
# DDA + NEE at 1080p


def trace_bench(args) -> dict:
    import numpy as np
    import torch

    import caelix_amd
    from caelix_amd import scene

    t0 = time.perf_counter()
    data = scene.hd_scene()
    gen_s = time.perf_counter() - t0
    ts = caelix_amd.TraceScene(data, load_factor=0.7)
    cam = data.camera
    w, h = cam["width"], cam["height"]
    stream = torch.cuda.current_stream()
    color = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
    radiance = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
    attrib = torch.empty(w * h, dtype=torch.int32, device="cuda")
    counters = torch.zeros(8, dtype=torch.int64, device="cuda")

    def frame(i):
        ts.trace_frame(cam, TRACE_BOUNCES, 1000 + i, TRACE_SPP, color, radiance, attrib, counters, stream)

    for i in range(args.warmup):
        frame(i)
    torch.cuda.synchronize()
    before = counters.cpu().numpy().copy()
    ms = _time_steps(lambda i: frame(args.warmup + i), args.steps, 0, stream)
    after = counters.cpu().numpy()
    per_frame = (after - before) / args.steps
    rays_per_frame = float(per_frame[0] + per_frame[1])
    value = rays_per_frame / (ms * 1e-3) / 1e6

    profile_frames = 3
    stages = _profile(frame, profile_frames)

    # Demand bytes per ray and stage and a parity check, from the oracle on a band of rows of the
    # last timed frame.
    import oracle

    seed = 1000 + args.warmup + args.steps - 1
    frame(args.warmup + args.steps - 1)  # the last timed frame again, into the output buffers
    torch.cuda.synchronize()
    o = oracle.OracleScene.from_gpu_arrays(ts.host_arrays(), data)
    band = (h // 2 - 24, h // 2 + 24)
    o_color, o_rad, o_attrib, o_counters = o.trace_frame(cam, TRACE_BOUNCES, seed, TRACE_SPP, rows=band)
    lo, hi = band[0] * w, band[1] * w
    g_attrib = attrib.cpu().numpy().view(np.uint32)[lo:hi]
    g_rad = radiance.cpu().numpy().reshape(-1, 4)[lo:hi, :3].astype(np.float64)
    rel = np.abs(g_rad - o_rad[:, :3]) / (1.0 + np.abs(o_rad[:, :3]))
    band_pixels = float((band[1] - band[0]) * w)
    band_bounce = float(o_counters[0]) - band_pixels
    band_shadow = float(o_counters[1])
    pixels = float(w * h)
    frame_bounce = float(per_frame[0]) - pixels
    frame_shadow = float(per_frame[1])
    bytes_per_ray = float(o_counters[5]) / float(o_counters[0] + o_counters[1])
    # Algorithmic bytes of each stage per frame: the scene data it loads (the oracle's demand bytes of
    # that kind of work, per ray of that kind, times the frame's rays of that kind) plus the path
    # state the wavefront moves (per ray: primary hit 16; shade reads hit 16 and writes radiance,
    # albedo, emission, throughput, state 16 each, ray 32, queue 4, shadow entry 48; extend reads
    # queue 4 + ray 32 and writes the hit 16; shadow reads the entry 48 and updates radiance 32;
    # bounce shade reads hit, state, ray, throughput, radiance 96 and writes 116; finish 84).
    per_kind = [float(o_counters[6 + k]) for k in range(5)]
    algorithmic = {
        "trace_primary": per_kind[0] / band_pixels * pixels + 16.0 * pixels,
        "trace_extend": per_kind[1] / max(band_bounce, 1.0) * frame_bounce + 52.0 * frame_bounce,
        "trace_shadow": per_kind[2] / max(band_shadow, 1.0) * frame_shadow + 80.0 * frame_shadow,
        "trace_shade_primary": per_kind[3] / band_pixels * pixels + 196.0 * pixels,
        "trace_shade_bounce": per_kind[4] / max(band_bounce, 1.0) * frame_bounce + 260.0 * frame_bounce,
        "trace_finish": 84.0 * pixels,
    }
    stage_table = {}
    for name, (total_ms, launches) in stages.items():
        per_frame_ms = total_ms / profile_frames
        gb = algorithmic.get(name, 0.0) / 1e9
        stage_table[name] = {"ms_per_frame": round(per_frame_ms, 4), "launches_per_frame": launches / profile_frames,
                             "algorithmic_GB": round(gb, 3),
                             "GB_per_s": round(gb / (per_frame_ms * 1e-3), 1) if per_frame_ms > 0 else 0.0}
    dominant = max(stages, key=lambda k: stages[k][0])
    kernel_ms = stages[dominant][0] / stages[dominant][1]
    launch_bytes = algorithmic.get(dominant, 0.0) / (stages[dominant][1] / profile_frames)
    achieved = launch_bytes / (kernel_ms * 1e-3) / 1e9
    frame_bytes = sum(algorithmic.values())
    roofline = {
        "bound": "hbm",
        "kernel": dominant,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": _traffic(dominant),
        "kernel_ms": round(kernel_ms, 4),
        "algorithmic_bytes": int(launch_bytes),
        "frame_algorithmic_bytes": int(frame_bytes),
        "frame_GB_per_s": round(frame_bytes / (ms * 1e-3) / 1e9, 1),
        "demand_bytes_per_ray": round(bytes_per_ray, 1),
    }
    result = {
        "metric": "Mrays/s DDA+NEE @1080p (primary + bounce + shadow rays)",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "ms_per_step": round(ms, 4),
        "dtype": "f32",
        "config": {
            "workload": "DDA+NEE 1920x1080, 1 spp, 2 bounces; 1,048,576 bricks in 4096^3 (8-octave value-noise "
                        "terrain), hash load 0.7, 65,536 emissive bricks, emission log-uniform [0.01, 100]",
            "bricks": int(len(data.keys)),
            "light_bricks": ts.light_bricks,
            "light_faces": ts.light_faces,
            "slots": ts.hash.slot_count,
            "rays_per_frame": int(rays_per_frame),
            "shadow_rays_per_frame": int(per_frame[1]),
            "scene_generation_s": round(gen_s, 1),
        },
        "roofline": roofline,
        "parity": {
            "rows": list(band),
            "attrib_exact": bool(np.array_equal(g_attrib, o_attrib)),
            "max_rel_radiance": float(rel.max()) if len(rel) else 0.0,
        },
        "stages": stage_table,
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = trace_cpu_baseline(o, cam, seed)
    return result


def host_cpu() -> dict:
    """The host's CPU model (the first "model name" of /proc/cpuinfo, as lscpu reports it) and the
    hardware threads this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "host_threads": usable}


def trace_cpu_baseline(o, cam, seed) -> dict:
    """The oracle's path tracer (OpenMP over rows) on bands of 64 rows of the same frame, repeated
    over the next bands until at least 10 s have passed."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rays = 0
    t0 = time.perf_counter()
    rows = 0
    r0 = 0
    while True:
        band = (r0, min(r0 + 64, cam["height"]))
        _, _, _, c = o.trace_frame(cam, TRACE_BOUNCES, seed, TRACE_SPP, rows=band)
        rays += int(c[0] + c[1])
        rows += band[1] - band[0]
        r0 = band[1] % cam["height"]
        if time.perf_counter() - t0 >= 10.0:
            break
    seconds = time.perf_counter() - t0
    return {"value": round(rays / seconds / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"oracle path tracer on {rows} rows of the same 1080p frame ({rays} rays), "
                      f"{threads} OpenMP threads", **host_cpu()}


# ------------------------------------------------------------------------------------------------
# Light-tree rebuild


def lighttree_bench(args, with_cpu: bool) -> dict:
    import torch

    import caelix_amd
    from caelix_amd import synthetic

    n = 1 << 22
    leaves = synthetic.light_leaves(n, 2024, extent=8192.0, clusters=64, sigma=32.0)
    builder = caelix_amd.LightBvhBuilder(n)
    builder.upload(leaves)
    stream = torch.cuda.current_stream()

    def build(_):
        builder.build(stream)

    # the sort's digit width and tile, picked for this workload (speed only: the tree is the same)
    from caelix_amd import autotune

    sort_tuning = autotune.tune_build("lighttree-4m", lambda: build(0), reps=3, use_cache=False)

    ms = _time_steps(build, args.steps, args.warmup, stream)
    out = builder.read()
    assert int(out["header"]["leaf_count"]) == n and int(out["state"]["node_overflow"]) == 0
    profile_steps = max(3, min(args.steps, 10))
    stages = _profile(build, profile_steps)
    per_stage = {name: {"avg_ms": round(total / launches, 5), "launches_per_step": launches / profile_steps}
                 for name, (total, launches) in stages.items()}
    dominant = max(stages, key=lambda k: stages[k][0])
    avg_ms = stages[dominant][0] / stages[dominant][1]
    achieved = STAGE_BYTES.get(dominant, 0.0) * n / (avg_ms * 1e-3) / 1e9
    pipeline_gbs = BUILD_BYTES_PER_LEAF * n / (ms * 1e-3) / 1e9
    result = {
        "metric": "light-tree build Mprims/s (4,194,304 emissive leaves, full rebuild)",
        "value": round(n / (ms * 1e-3) / 1e6, 2),
        "unit": "Mprims/s",
        "ms_per_step": round(ms, 4),
        "dtype": "f32+u32",
        "config": {"workload": "light-tree rebuild, 4,194,304 leaves (half uniform in 8192^3, half in 64 "
                               "clusters sigma 32), u32 Morton keys", "leaves": n},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _traffic(dominant),
                     "algorithmic_bytes": int(STAGE_BYTES.get(dominant, 0.0) * n)},
        "pipeline_roofline": {"bytes_per_leaf": BUILD_BYTES_PER_LEAF, "achieved": round(pipeline_gbs, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(pipeline_gbs / HBM_PEAK_GBS, 4)},
        "stages": per_stage,
        "sort_tuning": {"digit_bits": sort_tuning["digit_bits"], "tile": sort_tuning["tile"],
                        "trials": [{k: t[k] for k in ("digit_bits", "tile", "ms")} for t in sort_tuning["trials"]]},
    }
    # The 64-bit key form of the same rebuild (code << 32 | leaf sorted as one key): the same tree.
    def build64(_):
        builder.build(stream, keys64=True)

    ms64 = _time_steps(build64, args.steps, args.warmup, stream)
    out64 = builder.read()
    assert out64["header"].tobytes() == out["header"].tobytes()
    stages64 = _profile(build64, profile_steps)
    gbs64 = BUILD_BYTES_PER_LEAF_KEYS64 * n / (ms64 * 1e-3) / 1e9
    result["keys64"] = {
        "value": round(n / (ms64 * 1e-3) / 1e6, 2), "unit": "Mprims/s", "ms_per_step": round(ms64, 4),
        "pipeline_roofline": {"bytes_per_leaf": BUILD_BYTES_PER_LEAF_KEYS64, "achieved": round(gbs64, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs64 / HBM_PEAK_GBS, 4)},
        "stages": {name: {"avg_ms": round(total / launches, 5), "launches_per_step": launches / profile_steps}
                   for name, (total, launches) in stages64.items() if name in ("lt_morton", "sort64_pass")},
    }
    if with_cpu:
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
        if time.perf_counter() - t0 >= 8.0:
            break
    seconds = (time.perf_counter() - t0) / runs
    # The oracle's build is the reference's kernel sequence restated as one sequential pass (each kernel
    # a loop over its lanes), so it runs on one thread; the trace baseline shows the multi-core oracle.
    return {"value": round(sample / seconds / 1e6, 3), "unit": "Mprims/s", "cores": 1, "kind": "port",
            "sample": f"oracle light-tree build of the first {sample} leaves of the same workload, "
                      f"{runs} runs, single thread", **host_cpu()}


def diff_bench(args) -> dict:
    """The differentiable form: 512x512, 4 spp, 3 bounces on the terrain scaled to 262,144 bricks
    (16,384 emitters). Loss = squared error to a target rendered with the parameters perturbed by
    seeded N(0, 0.05); gradients of the 768 albedos and 16,384 emission scales by the path-replay
    adjoint. Reports forward and adjoint time (HIP events) and a central finite-difference check
    (h = 1e-3, roulette off) of the two largest albedo and emission-scale gradients."""
    import numpy as np
    import torch

    import caelix_amd
    from caelix_amd import scene
    from caelix_amd.diff import DifferentiableRenderer, TRACE_NO_RUSSIAN_ROULETTE

    data = scene.diff_scene()
    ts = caelix_amd.TraceScene(data, load_factor=0.7)
    cam = data.camera
    bounces, spp, seed = 3, 4, 7
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(11)
    emitters = data.emissive_bricks
    r = DifferentiableRenderer(ts, cam, bounces, spp, seed, emitters)
    a0 = r.initial_albedo()
    s0 = torch.ones(len(emitters), dtype=torch.float32, device="cuda")
    a_t = (a0 + torch.from_numpy(rng.normal(0.0, 0.05, size=(256, 3)).astype(np.float32)).cuda()).clamp(0.01, 0.99)
    s_t = s0 + torch.from_numpy(rng.normal(0.0, 0.05, size=len(emitters)).astype(np.float32)).cuda()
    with torch.no_grad():
        target = r.render(a_t, s_t).clone()
    r.set_parameters(a0, s0)
    steps = max(3, min(args.steps, 10))
    fwd_ms = _time_steps(lambda _: r.forward_image(), steps, 1, stream)
    image = r.forward_image().clone()
    grad_image = 2.0 * (image - target)
    adj_ms = _time_steps(lambda _: r.adjoint(grad_image), steps, 1, stream)

    # finite differences at full size, roulette off
    rf = DifferentiableRenderer(ts, cam, bounces, spp, seed, emitters, flags=TRACE_NO_RUSSIAN_ROULETTE)
    with torch.no_grad():
        target_f = rf.render(a_t, s_t).clone()
    rf.set_parameters(a0, s0)
    base = rf.forward_image().clone()
    ga, gs = rf.adjoint(2.0 * (base - target_f))
    ga = ga.double().cpu().numpy()
    gs = gs.double().cpu().numpy()

    def loss(a, sc):
        rf.set_parameters(a, sc)
        return float(((rf.forward_image().double() - target_f.double()) ** 2).sum())

    h = 1e-3
    checks = []
    for flat in np.argsort(-np.abs(ga).ravel())[:2]:
        m, c = divmod(int(flat), 3)
        up, dn = a0.clone(), a0.clone()
        up[m, c] += h
        dn[m, c] -= h
        fd = (loss(up, s0) - loss(dn, s0)) / (2 * h)
        checks.append({"param": f"albedo[{m}][{c}]", "fd": fd, "adjoint": float(ga[m, c]),
                       "rel_err": abs(fd - ga[m, c]) / max(abs(fd), 1e-30)})
    for e in np.argsort(-np.abs(gs))[:2]:
        up, dn = s0.clone(), s0.clone()
        up[int(e)] += h
        dn[int(e)] -= h
        fd = (loss(a0, up) - loss(a0, dn)) / (2 * h)
        checks.append({"param": f"scale[{int(e)}]", "fd": fd, "adjoint": float(gs[e]),
                       "rel_err": abs(fd - gs[e]) / max(abs(fd), 1e-30)})
    r.set_parameters(a0, s0)
    return {"forward_ms": round(fwd_ms, 4), "adjoint_ms": round(adj_ms, 4), "adjoint_over_forward": round(adj_ms / fwd_ms, 3),
            "config": {"workload": "differentiable: 512x512, 4 spp, 3 bounces, 262,144 bricks (2048^3 terrain), "
                                   "16,384 emitters; 768 albedo + 16,384 emission-scale gradients",
                       "bricks": int(len(data.keys)), "emitters": int(len(emitters))},
            "fd_check": checks}


def batch_bench(args) -> dict:
    """The batched form: 256 chunk scenes (brick counts log-uniform in [64, 65536] as far as the
    64^3-brick chunk's terrain holds them, 5 % emissive), each step rebuilding every scene's light
    tree and tracing a 256x256, 1 spp, 2-bounce frame of every scene. Timed four ways on the same
    scenes: launch-per-scene (one build and one cx_trace_frame per scene, launched from Python), the
    same launch-per-scene step captured in a hipGraph and replayed, batched (one batched build and one
    cx_trace_batch), and the batched step captured in a hipGraph and replayed."""
    import numpy as np
    import torch

    from caelix_amd import scene
    from caelix_amd.batch import PlainScenes, TraceBatch

    t0 = time.perf_counter()
    datas = scene.chunk_batch(256)
    gen_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream()
    batch = TraceBatch(datas)
    plain = PlainScenes(datas)
    steps = max(2, min(args.steps, 5))
    plain_ms = _time_steps(lambda i: plain.step(2, 50 + i, 1), steps, 1, stream)

    def batched(i):
        batch.build_trees()
        batch.trace(2, 50 + i, 1)

    batch_ms = _time_steps(batched, steps, 1, stream)
    from caelix_amd import autotune

    tuned_build = autotune.tune_build("batch256-build", batch.build_trees, reps=3, use_cache=False)
    tuned = autotune.tune("batch256-256x256-b2", lambda: batched(0), reps=2, use_cache=False)
    tuned_ms = _time_steps(batched, steps, 1, stream)
    side = torch.cuda.Stream()
    side.wait_stream(stream)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        batched(0)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=side):
            batched(0)
    stream.wait_stream(side)
    graph_ms = _time_steps(lambda _: graph.replay(), steps, 1, stream)
    # The launch-per-scene step captured in a hipGraph too (the same launches without Python and ctypes
    # in between): the baseline the batched form is held against without host launch overhead.
    plain_graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        plain.step(2, 50, 1)
        torch.cuda.synchronize()
        with torch.cuda.graph(plain_graph, stream=side):
            plain.step(2, 50, 1)
    stream.wait_stream(side)
    plain_graph_ms = _time_steps(lambda _: plain_graph.replay(), steps, 1, stream)
    # the rays of one step
    batch.counters.zero_()
    batched(0)
    torch.cuda.synchronize()
    c = batch.counters.cpu().numpy()
    rays = float(c[0] + c[1])
    bricks = np.array([len(d.keys) for d in datas])
    return {
        "plain_ms": round(plain_ms, 3), "batched_default_ms": round(batch_ms, 3), "batched_ms": round(tuned_ms, 3),
        "graph_ms": round(graph_ms, 3), "plain_graph_ms": round(plain_graph_ms, 3),
        "autotuned": {"queue_chunk": tuned["queue_chunk"], "workgroups_per_cu": tuned["workgroups_per_cu"],
                      "trials": len(tuned.get("trials", [])), "digit_bits": tuned_build["digit_bits"],
                      "tile": tuned_build["tile"],
                      "build_trials": [{k: t[k] for k in ("digit_bits", "tile", "ms")}
                                       for t in tuned_build.get("trials", [])]},
        "speedup_batched": round(plain_ms / tuned_ms, 2), "speedup_graph": round(plain_ms / graph_ms, 2),
        "speedup_graph_vs_plain_graph": round(plain_graph_ms / graph_ms, 2),
        "rays_per_step": int(rays), "graph_Mrays_per_s": round(rays / (graph_ms * 1e-3) / 1e6, 1),
        "config": {"workload": "256 chunk scenes (64^3-brick terrain chunks), 5 % emissive, light-tree build + "
                               "256x256 1 spp 2-bounce frame each", "scenes": len(datas),
                   "bricks_total": int(bricks.sum()), "bricks_min": int(bricks.min()), "bricks_max": int(bricks.max()),
                   "light_bricks_total": int(batch.light_bricks.sum()), "scene_generation_s": round(gen_s, 1)},
    }


def copy_bench(args) -> dict:
    """A plain device copy of 1 GiB (torch's copy kernel): the calibration run for the HBM byte
    counters (FETCH_SIZE / WRITE_SIZE per dispatch against 1 GiB read and 1 GiB written)."""
    import torch

    n = 1 << 30
    src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    ms = _time_steps(lambda _: dst.copy_(src), args.steps, args.warmup, stream)
    gbs = 2 * n / (ms * 1e-3) / 1e9
    return {"metric": "device copy GB/s (1 GiB read + 1 GiB written)", "value": round(gbs, 1), "unit": "GB/s",
            "ms_per_step": round(ms, 4), "dtype": "u8", "config": {"workload": "copy 1 GiB", "bytes": n},
            "roofline": {"bound": "hbm", "kernel": "copy", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": 2 * n}}


def cxw_bench(args) -> dict:
    """The 1080p terrain world (1,048,576 bricks) written to a .cxw file (raw deflate, level 1), then
    timed: open (index parse) + decode of every sector on the host cores, and the GPU scene build from
    the decoded bricks (upload, finalize, bulk hash, light records, light tree). One step = one load.
    value = decoded bricks per second of open + decode. Host-side work: no HBM roofline applies to the
    decode; the scene build's kernels are the trace workload's set-up kernels."""
    import tempfile

    import torch

    import caelix_amd
    from caelix_amd import cxw, scene

    data = scene.hd_scene()
    threads = min(os.cpu_count() or 1, 16)
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as tmp:
        path = Path(tmp) / "terrain_1080p.cxw"
        t0 = time.perf_counter()
        cxw.write(path, [cxw.WEntity(cxw.sectors_from_bricks(data.keys, data.payload))], level=1)
        write_s = time.perf_counter() - t0
        file_bytes = path.stat().st_size
        times = []
        decoded = None
        for i in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            with cxw.CxwFile(path) as f:
                decoded = f.decode_all(threads)
                sectors = len(f.sectors)
            if i >= args.warmup:
                times.append(time.perf_counter() - t0)
        bricks = len(decoded["keys"])
        decode_s = sorted(times)[len(times) // 2]
        loaded = cxw.load_scene(path, data.materials, data.light_codes, data.camera, threads=threads)
    torch.cuda.synchronize()
    build_times = []
    for i in range(2):
        t0 = time.perf_counter()
        ts = caelix_amd.TraceScene(loaded, load_factor=0.7)
        torch.cuda.synchronize()
        build_times.append(time.perf_counter() - t0)
        del ts
    inflated = bricks * 1024
    return {"metric": "cxw load: decoded bricks/s (open + decode on host threads)",
            "value": round(bricks / decode_s, 1), "unit": "bricks/s", "ms_per_step": round(decode_s * 1e3, 3),
            "dtype": "u16", "config": {"workload": "1080p terrain world as .cxw: 1,048,576 bricks, 16^3-brick "
                                                   "sectors, raw deflate level 1", "bricks": bricks,
                                       "sectors": sectors, "file_bytes": file_bytes, "threads": threads},
            "roofline": None,
            "cxw": {"decode_s": round(decode_s, 4), "inflated_GB_per_s": round(inflated / decode_s / 1e9, 3),
                    "file_MB_per_s": round(file_bytes / decode_s / 1e6, 1), "write_s": round(write_s, 2),
                    "scene_build_s": round(min(build_times), 4), "voxels": decoded["voxels"]}}


def main() -> int:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--workload", default="trace", choices=["trace", "lighttree", "copy", "cxw"])
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-lighttree", action="store_true", help="trace only, without the light-tree figure")
    parser.add_argument("--no-adjoint", action="store_true", help="without the differentiable-form figure")
    parser.add_argument("--no-batch", action="store_true", help="without the batched-form figure")
    args = parser.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the hot path runs on one GPU and this build has no multi-GPU "
              "path (the design for sharding it is documented); run it with --gpus 1", file=sys.stderr)
        return 2
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload == "trace":
        result = trace_bench(args)
        if not args.no_adjoint:
            result["adjoint"] = diff_bench(args)
        if not args.no_batch:
            result["batch"] = batch_bench(args)
        if not args.no_lighttree:
            lt = lighttree_bench(args, with_cpu=not args.no_cpu_baseline)
            result["lighttree"] = {k: lt[k] for k in ("value", "unit", "ms_per_step", "roofline",
                                                      "pipeline_roofline", "config", "keys64", "cpu_baseline",
                                                      "stages", "sort_tuning")
                                   if k in lt}
    elif args.workload == "copy":
        result = copy_bench(args)
    elif args.workload == "cxw":
        result = cxw_bench(args)
    else:
        result = lighttree_bench(args, with_cpu=not args.no_cpu_baseline)
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
    for key in ("cpu_baseline", "parity", "lighttree", "adjoint", "batch", "pipeline_roofline", "keys64", "stages",
                "sort_tuning", "cxw"):
        if key in result:
            line[key] = result[key]
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
