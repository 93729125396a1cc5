"""Autotuner for the shell analysis launch shape.

It times the real call over a small search space -- lanes per peak (1 or 16, "gang"; the
library also has 64, one wavefront per peak, which is left out of the default space: at the
1024^3 tile it reads 1.5x fewer bytes but runs 3x longer), the peak-order brick size
(2^brick_log2 cells a side) and the wavefront schedule (exact crossing pass, plain order,
learned bracket) -- and keeps the fastest.  None of these knobs changes a
result; only the time.  Results are cached per workload signature in a JSON file so a
production run tunes once per shape.
"""
from __future__ import annotations

import json
import os
import time

DEFAULT_SPACE = dict(gang=(1, 16), brick_log2=(4, 5, 6, 7), order=("crossing", "plain",
                                                                  "learned"))
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "huntian_hip", "tuning.json")


def signature(**shape) -> str:
    """Workload key: the sizes that decide the best shape (grid side, peak count rounded to
    a power of two, hunt radius, variant, ...)."""
    items = []
    for k in sorted(shape):
        v = shape[k]
        if k == "npeaks":
            v = 1 << max(0, int(v) - 1).bit_length()
        items.append(f"{k}={v}")
    return ";".join(items)


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def tune(run, ctx, *, space=None, has_bracket=False, reps=1, cache=None, key=None, log=None):
    """Time run(order=...) under every knob combination of `space` and set the fastest on
    ctx.  run(order) must perform one complete analysis call.  Returns
    {"best": {...}, "table": [...], "cached": bool}."""
    import torch

    space = dict(DEFAULT_SPACE if space is None else space)
    orders = [o for o in space.get("order", ("crossing",)) if o != "learned" or has_bracket]
    if cache and key:
        hit = _load(cache).get(key)
        if hit:
            b = hit["best"]
            ctx.set_tuning(gang=b["gang"], brick_log2=b["brick_log2"])
            return dict(best=b, table=hit.get("table", []), cached=True)
    table = []
    for gang in space.get("gang", (1,)):
        for L in space.get("brick_log2", (6,)):
            for order in orders:
                ctx.set_tuning(gang=gang, brick_log2=L)
                run(order)  # warm: buffers grow, schedule compiles
                torch.cuda.synchronize()
                best_ms = float("inf")
                for _ in range(reps):
                    t0 = time.perf_counter()
                    run(order)
                    torch.cuda.synchronize()
                    best_ms = min(best_ms, (time.perf_counter() - t0) * 1e3)
                row = dict(gang=gang, brick_log2=L, order=order, ms=round(best_ms, 3))
                table.append(row)
                if log:
                    log(f"autotune {row}")
    best = min(table, key=lambda r: r["ms"])
    ctx.set_tuning(gang=best["gang"], brick_log2=best["brick_log2"])
    if cache and key:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        data = _load(cache)
        data[key] = dict(best=best, table=table)
        tmp = cache + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1)
        os.replace(tmp, cache)
    return dict(best=best, table=table, cached=False)
