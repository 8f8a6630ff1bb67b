"""Autotuner for the persistent OIFT kernel's launch shape.

The OIFT kernel's cost is dominated by the latency of its barrier-separated phases, so
the best grid size (workgroups taking part in every grid barrier) and flood budget (LDS
batches a workgroup relaxes locally before handing work to the next round) depend on
the workload's shape: how many voxels, how wide the fronts are.  `autotune_oift` times a
small set of candidates with the library's own device events and leaves the fastest on
the context; results are cached per workload key for the life of the process (and in a
JSON file when ROIFT_TUNE_CACHE names one).
"""
from __future__ import annotations

import json
import os

# blocks 0 = the library's default grid, sized from the workload's arc count
DEFAULT_CANDIDATES = [(0, 1024), (64, 1024), (96, 1024), (128, 128), (128, 1024), (192, 1024), (256, 128)]
_cache: dict[str, tuple[int, int]] = {}


def _load_file():
    path = os.environ.get("ROIFT_TUNE_CACHE")
    if path and os.path.exists(path):
        with open(path) as f:
            for k, v in json.load(f).items():
                _cache.setdefault(k, tuple(v))


def _save_file():
    path = os.environ.get("ROIFT_TUNE_CACHE")
    if path:
        with open(path, "w") as f:
            json.dump({k: list(v) for k, v in _cache.items()}, f)


def autotune_oift(ctx, key: str, run, candidates=None, reps: int = 1) -> dict:
    """Pick (oift_blocks, flood_budget) for the workload `run()` executes on `ctx`.

    run() must launch the OIFT on ctx (alone or inside a pipeline).  Each candidate is
    timed by the OIFT kernel's device-event time (ctx.kernel_timings()), best of `reps`
    runs after one warm-up.  Returns {"blocks", "budget", "ms", "trials"}; the winner is
    left set on ctx (blocks 0: the library's own size-based default).
    """
    _load_file()
    if key in _cache:
        b, g = _cache[key]
        ctx.set_param("oift_blocks", b)
        ctx.set_param("flood_budget", g)
        return {"blocks": b, "budget": g, "ms": None, "trials": {}, "cached": True}
    trials = {}
    for b, g in candidates or DEFAULT_CANDIDATES:
        ctx.set_param("oift_blocks", b)
        ctx.set_param("flood_budget", g)
        run()   # warm-up (first-touch allocations, caches)
        best = None
        for _ in range(reps):
            run()
            t = ctx.kernel_timings()["oift_kernel_ms"]
            best = t if best is None else min(best, t)
        trials[f"{b}x{g}"] = round(best, 3)
    bk = min(trials, key=trials.get)
    b, g = (int(v) for v in bk.split("x"))
    ctx.set_param("oift_blocks", b)
    ctx.set_param("flood_budget", g)
    _cache[key] = (b, g)
    _save_file()
    return {"blocks": b, "budget": g, "ms": trials[bk], "trials": trials, "cached": False}
