"""Launch-shape autotuner for batched and fused response launches.

The kernels' results do not depend on the launch shape, only their speed does: the tile
length (output samples per workgroup tile, i.e. samples per lane x 256) and the persistent grid
density (workgroups per CU). `tune(plan)` times every candidate on the plan itself with HIP
events, keeps the fastest on the plan and remembers it per workload key (precision, fused,
TDI generation, equal arms, dt, total samples rounded to a power of two), so later plans of
the same kind start from the cached choice. The cache lives in memory and, if FLR_TUNE_CACHE
names a file, in that JSON file.
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, Iterable, Optional, Tuple

import torch

from ._capi import check, lib

_cache: Dict[str, Tuple[int, int]] = {}
_loaded = False


def candidates(plan) -> Iterable[Tuple[int, int]]:
    """(tile_samples, grid_per_cu) pairs valid for the plan."""
    if not plan.br.fused:
        tiles = (256,)
    elif plan.br.precision != "fp64":
        tiles = (256, 512, 1024)
    else:
        tiles = (512,)
    for t in tiles:
        for g in (4, 8, 16, 32):
            yield (t, g)


def workload_key(plan) -> str:
    br = plan.br
    orbits = type(br.model.response_orbits).__name__
    total = max(1, int(plan.lengths.sum()))
    return (f"{br.precision}|fused={int(br.fused)}|{br.model.tdi}|{orbits}|dt={br.dt:g}|"
            f"n=2^{round(math.log2(total))}")


def _cache_file() -> Optional[str]:
    return os.environ.get("FLR_TUNE_CACHE") or None


def _load():
    global _loaded
    if _loaded:
        return
    _loaded = True
    path = _cache_file()
    if path and os.path.exists(path):
        with open(path) as f:
            for k, v in json.load(f).items():
                _cache[k] = (int(v[0]), int(v[1]))


def _save():
    path = _cache_file()
    if path:
        with open(path, "w") as f:
            json.dump({k: list(v) for k, v in _cache.items()}, f, indent=1, sort_keys=True)


def apply(plan, tile_samples: int, grid_per_cu: int) -> None:
    """Set a plan's launch shape (no launch of the plan may be in flight)."""
    torch.cuda.synchronize()
    check(lib().flr_batch_set_launch(plan._h, int(tile_samples), int(grid_per_cu)))
    plan.launch_shape = (int(tile_samples), int(grid_per_cu))


def tune(plan, reps: int = 3, use_cache: bool = True) -> dict:
    """Time every candidate launch shape on `plan`, keep the fastest; returns
    {"key", "best": (tile, grid), "ms": {shape: ms}, "cached": bool}."""
    _load()
    key = workload_key(plan)
    if use_cache and key in _cache:
        apply(plan, *_cache[key])
        return {"key": key, "best": _cache[key], "ms": {}, "cached": True}
    times = {}
    for shape in candidates(plan):
        apply(plan, *shape)
        plan.run()  # warm-up (first launch of this variant)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            plan.run()
        b.record()
        torch.cuda.synchronize()
        times[shape] = a.elapsed_time(b) / reps
    best = min(times, key=times.get)
    apply(plan, *best)
    plan.check_errors()
    _cache[key] = best
    _save()
    return {"key": key, "best": best, "ms": {f"{t}x{g}": ms for (t, g), ms in times.items()},
            "cached": False}
