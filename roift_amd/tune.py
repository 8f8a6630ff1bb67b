"""Autotuner for the pipeline's launch parameters.

The OIFT kernel's cost is dominated by the latency of its barrier-separated phases, so
the best grid size (workgroups taking part in every grid barrier) and flood budget (LDS
batches a workgroup relaxes locally before handing work to the next round) depend on
the workload's shape: how many voxels, how wide the fronts are.  Whether narrow flood
queues run on a single wavefront (flood mode 0) or workgroup-wide (mode 2), and the plain
relaxation's tile (rows, z planes, XCD grouping), are tuned the same way.

`autotune` walks stages of candidates (coordinate descent: each stage keeps the winners of
the stages before it), times every candidate by the wall clock of `run()` (after one warm-up
run, median of `reps` runs, the device synchronised around each) and leaves the fastest
setting on the context.  The wall clock covers graph replays too (the batched pipeline's
HIP graph records no device events inside it).  Results are cached per workload key for the
life of the process (and in a JSON file when ROIFT_TUNE_CACHE names one).
"""
from __future__ import annotations

import json
import os
import time

# (oift_blocks, flood_budget); blocks 0 = the library's default grid, sized from the arc count
OIFT_SHAPES = [(0, 1024), (64, 1024), (96, 1024), (128, 128), (128, 1024), (192, 1024), (256, 128)]
STAGE_OIFT = ("oift", [{"oift_blocks": b, "flood_budget": g} for b, g in OIFT_SHAPES])
STAGE_FLOOD = ("flood", [{"flood_mode": 0}, {"flood_mode": 2}])
# two-class 6-adjacency relaxation tile: rows | planes << 8 | xcd << 16 (0 = 4 rows, 32 planes, XCD-grouped)
STAGE_RELAX_TILE = ("relax_tile", [{"relax_tile": 0}, {"relax_tile": 8 | (32 << 8) | (1 << 16)},
                                   {"relax_tile": 4 | (16 << 8) | (1 << 16)}, {"relax_tile": 4 | (32 << 8)}])
STAGES_BATCH = [STAGE_OIFT, STAGE_FLOOD]          # the batched relaxation's tiles are fixed
STAGES_PLAIN = [STAGE_OIFT, STAGE_FLOOD, STAGE_RELAX_TILE]
DEFAULTS = {"oift_blocks": 0, "flood_budget": 0, "flood_mode": 0, "relax_tile": 0}

_cache: dict[str, dict] = {}


def _load_file():
    path = os.environ.get("ROIFT_TUNE_CACHE")
    if path and os.path.exists(path):
        with open(path) as f:
            for k, v in json.load(f).items():
                if isinstance(v, dict):
                    _cache.setdefault(k, v)


def _save_file():
    path = os.environ.get("ROIFT_TUNE_CACHE")
    if path:
        with open(path, "w") as f:
            json.dump(_cache, f)


def _apply(ctx, setting):
    for name, value in setting.items():
        ctx.set_param(name, value)


def _wall(run, sync):
    sync()
    t0 = time.perf_counter()
    run()
    sync()
    return (time.perf_counter() - t0) * 1e3


def autotune(ctx, key: str, run, stages=None, reps: int = 3, sync=None) -> dict:
    """Pick launch parameters for the workload `run()` executes on `ctx`.

    stages: list of (name, [setting dict, ...]); default STAGES_BATCH.  Each candidate gets one
    warm-up run and `reps` timed runs (median).  Returns {"setting", "ms", "trials"}; the winner
    is left set on ctx.
    """
    if sync is None:
        import torch
        sync = torch.cuda.synchronize
    _load_file()
    if key in _cache:
        _apply(ctx, _cache[key])
        return {"setting": dict(_cache[key]), "ms": None, "trials": {}, "cached": True}
    best = dict(DEFAULTS)
    _apply(ctx, best)
    trials, best_ms = {}, None
    for name, cands in stages or STAGES_BATCH:
        stage_best, stage_ms = None, None
        for cand in cands:
            setting = {**best, **cand}
            _apply(ctx, setting)
            run()   # warm-up: first-touch allocations, caches, a graph's capture
            times = sorted(_wall(run, sync) for _ in range(reps))
            ms = times[len(times) // 2]
            trials[f"{name}:" + ",".join(f"{k}={v}" for k, v in cand.items())] = round(ms, 3)
            if stage_ms is None or ms < stage_ms:
                stage_best, stage_ms = setting, ms
        best, best_ms = stage_best, stage_ms
    _apply(ctx, best)
    _cache[key] = dict(best)
    _save_file()
    return {"setting": dict(best), "ms": round(best_ms, 3), "trials": trials, "cached": False}


def autotune_oift(ctx, key: str, run, reps: int = 3) -> dict:
    """The grid-size / flood-budget / flood-mode stages only (the batched pipeline)."""
    r = autotune(ctx, key, run, STAGES_BATCH, reps)
    r["blocks"] = r["setting"]["oift_blocks"]
    r["budget"] = r["setting"]["flood_budget"]
    return r
