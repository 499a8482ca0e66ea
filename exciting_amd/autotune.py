"""Tile-shape autotuner for the setup kernels.

For a stage ("ystage", "mt", and for the grouped launch "bmt"), every compiled variant is
run on the caller's real inputs, timed with the library's per-stage HIP-event timers, and
the fastest is pinned in the context with ``exc_ctx_set_variant``.  Results are keyed by
the basis-size bucket (N rounded up to a multiple of 256) and the APW count naa, and can
be stored as JSON so a later run starts from the tuned table.
"""
from __future__ import annotations

import json
import os
from typing import Callable, Dict


def bucket(n: int) -> int:
    return ((n + 255) // 256) * 256


def sweep(ctx, stage: str, run: Callable[[], None], reps: int = 3, log=None) -> Dict[int, float]:
    """ms per launch of every variant of `stage` for the work `run()` launches."""
    import torch
    out = {}
    for v in range(ctx.num_variants(stage)):
        ctx.set_variant(stage, v)
        run()                               # warm-up (first launch of a variant loads its code)
        torch.cuda.synchronize()
        ctx.read_timing()
        ctx.set_timing(True)
        for _ in range(reps):
            run()
        ctx.set_timing(False)
        ms, cnt = ctx.read_timing()[stage]
        out[v] = ms / max(cnt, 1)
        if log:
            log(f"    {stage} v{v}: {out[v]:.3f} ms")
    return out


def tune(ctx, stages, run: Callable[[], None], reps: int = 3, log=None) -> Dict[str, int]:
    """Sweep each stage in turn (others at their current setting) and pin the fastest variant."""
    chosen = {}
    for st in stages:
        res = sweep(ctx, st, run, reps, log)
        best = min(res, key=res.get)
        ctx.set_variant(st, best)
        chosen[st] = best
        if log:
            log(f"  autotune {st}: " + ", ".join(f"v{v}={t:.3f}ms" for v, t in sorted(res.items())) + f" -> v{best}")
    return chosen


def save(path: os.PathLike, key, chosen: Dict[str, int]):
    table = {}
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
    table[str(key)] = chosen
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(path: os.PathLike, key):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(str(key))
