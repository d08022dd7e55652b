"""Autotuner of the trace's launch tunables (cx_trace_set_tuning) and of the light-tree build's sort
(cx_light_bvh_set_tuning).

For a workload step (a callable that records one batch step or one frame on the current stream), it
times every candidate setting with HIP events over a few repetitions, keeps the fastest, and caches the
choice in a JSON file keyed by the device and a workload key, so later runs reuse it. The tunables
change speed only, never results: ray-queue depth (`queue_chunk`, the items a wave claims per refill
of the persistent traversal) and persistent workgroups per CU (`workgroups_per_cu`, 0 = as many as fit);
for the build, the sort's digit width (`digit_bits`, 6 or 8) and tile (`tile`, 4096 or 8192 elements).
"""
from __future__ import annotations

import ctypes
import itertools
import json
import os
from pathlib import Path

from . import _torch, check, load

QUEUE_CHUNKS = (32, 64, 128, 256)
WORKGROUPS_PER_CU = (0, 2, 4)
DIGIT_BITS = (6, 8)
SORT_TILES = (4096, 8192)


class CxTraceTuning(ctypes.Structure):
    _fields_ = [("queue_chunk", ctypes.c_uint32), ("workgroups_per_cu", ctypes.c_uint32)]


class CxLightBvhTuning(ctypes.Structure):
    _fields_ = [("digit_bits", ctypes.c_uint32), ("tile", ctypes.c_uint32)]


def set_build_tuning(digit_bits: int, tile: int) -> None:
    t = CxLightBvhTuning(digit_bits, tile)
    check(load().cx_light_bvh_set_tuning(ctypes.byref(t)), "cx_light_bvh_set_tuning")


def get_build_tuning() -> dict:
    t = CxLightBvhTuning()
    load().cx_light_bvh_get_tuning(ctypes.byref(t))
    return {"digit_bits": int(t.digit_bits), "tile": int(t.tile)}


def _bind():
    lib = load()
    if not hasattr(lib, "_tuning_bound"):
        lib.cx_trace_set_tuning.restype = ctypes.c_int
        lib.cx_trace_set_tuning.argtypes = [ctypes.POINTER(CxTraceTuning)]
        lib.cx_trace_get_tuning.restype = None
        lib.cx_trace_get_tuning.argtypes = [ctypes.POINTER(CxTraceTuning)]
        lib._tuning_bound = True
    return lib


def set_tuning(queue_chunk: int, workgroups_per_cu: int) -> None:
    t = CxTraceTuning(queue_chunk, workgroups_per_cu)
    check(_bind().cx_trace_set_tuning(ctypes.byref(t)), "cx_trace_set_tuning")


def get_tuning() -> dict:
    t = CxTraceTuning()
    _bind().cx_trace_get_tuning(ctypes.byref(t))
    return {"queue_chunk": int(t.queue_chunk), "workgroups_per_cu": int(t.workgroups_per_cu)}


def default_cache() -> Path:
    return Path(os.environ.get("CAELIX_TUNING_CACHE", Path.home() / ".cache" / "caelix_amd" / "tuning.json"))


def _time(step, reps: int) -> float:
    torch = _torch()
    stream = torch.cuda.current_stream()
    step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / reps


def _cached(entry_key: str, cache: Path):
    if not cache.exists():
        return None
    try:
        return json.loads(cache.read_text()).get(entry_key)
    except (OSError, ValueError):
        return None


def _store(entry_key: str, cache: Path, value: dict) -> None:
    try:
        cache.parent.mkdir(parents=True, exist_ok=True)
        data = json.loads(cache.read_text()) if cache.exists() else {}
        data[entry_key] = value
        cache.write_text(json.dumps(data, indent=1))
    except (OSError, ValueError):
        pass


def tune_build(key: str, step, reps: int = 3, cache: Path | None = None, use_cache: bool = True) -> dict:
    """Picks and sets the fastest sort digit width and tile for a build step (a callable that records
    one or more light-tree builds on the current stream); returns {"digit_bits", "tile", "ms", "trials"}."""
    torch = _torch()
    cache = cache if cache is not None else default_cache()
    entry_key = f"{torch.cuda.get_device_name(0)}|build|{key}"
    known = _cached(entry_key, cache) if use_cache else None
    if known:
        set_build_tuning(known["digit_bits"], known["tile"])
        return dict(known, cached=True)
    trials = []
    for bits, tile in itertools.product(DIGIT_BITS, SORT_TILES):
        set_build_tuning(bits, tile)
        trials.append({"digit_bits": bits, "tile": tile, "ms": round(_time(step, reps), 4)})
    best = min(trials, key=lambda t: t["ms"])
    set_build_tuning(best["digit_bits"], best["tile"])
    _store(entry_key, cache, {k: best[k] for k in ("digit_bits", "tile", "ms")})
    return {"digit_bits": best["digit_bits"], "tile": best["tile"], "ms": best["ms"], "trials": trials, "cached": False}


def tune(key: str, step, reps: int = 3, cache: Path | None = None, use_cache: bool = True) -> dict:
    """Picks and sets the fastest tunables for `step`; returns {"queue_chunk", "workgroups_per_cu",
    "ms", "trials"}."""
    torch = _torch()
    device = torch.cuda.get_device_name(0)
    cache = cache if cache is not None else default_cache()
    entry_key = f"{device}|{key}"
    if use_cache and cache.exists():
        try:
            known = json.loads(cache.read_text()).get(entry_key)
        except (OSError, ValueError):
            known = None
        if known:
            set_tuning(known["queue_chunk"], known["workgroups_per_cu"])
            return dict(known, cached=True)
    trials = []
    for chunk, per_cu in itertools.product(QUEUE_CHUNKS, WORKGROUPS_PER_CU):
        set_tuning(chunk, per_cu)
        trials.append({"queue_chunk": chunk, "workgroups_per_cu": per_cu, "ms": round(_time(step, reps), 4)})
    best = min(trials, key=lambda t: t["ms"])
    set_tuning(best["queue_chunk"], best["workgroups_per_cu"])
    result = {"queue_chunk": best["queue_chunk"], "workgroups_per_cu": best["workgroups_per_cu"], "ms": best["ms"],
              "trials": trials}
    try:
        cache.parent.mkdir(parents=True, exist_ok=True)
        data = json.loads(cache.read_text()) if cache.exists() else {}
        data[entry_key] = {k: result[k] for k in ("queue_chunk", "workgroups_per_cu", "ms")}
        cache.write_text(json.dumps(data, indent=1))
    except (OSError, ValueError):
        pass
    return dict(result, cached=False)
