"""Tile-shape autotuner for the setup kernels.

For a stage ("ystage", "mt", and for the grouped launch "bmt", "crtgemm", "bsplit"), every
compiled variant is run on the caller's real inputs, timed with the library's per-stage
HIP-event timers, and the fastest is pinned in the context with ``exc_ctx_set_variant``.
Results are keyed by the basis-size bucket (the largest N of the batch rounded up to a
multiple of 256), the APW count naa and the atom count, and are kept in a JSON table;
``apply`` pins the table's choice for a batch before it is launched.  The library's built-in
defaults are used for shapes the table does not hold.  ``TABLE`` is the table measured on
MI355X for the shapes of the four GPU workloads.
"""
from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Callable, Dict, Optional

TABLE = Path(__file__).with_name("tuned.json")


def bucket(n: int) -> int:
    return ((n + 255) // 256) * 256


def key(cr, ngps) -> str:
    """Table key of a batch: N bucket of its largest basis, naa, natm."""
    return f"N{bucket(max(ngps))}-naa{cr.maxaa}-natm{cr.natm}"


def lookup(cr, ngps, path: os.PathLike = TABLE) -> Optional[Dict[str, int]]:
    return load(path, key(cr, ngps))


def apply(ctx, cr, ngps, path: os.PathLike = TABLE) -> Dict[str, int]:
    """Pin the tuned variants of this batch shape in ctx (nothing if the table has no entry);
    returns what was pinned."""
    chosen = lookup(cr, ngps, path) or {}
    for st, v in chosen.items():
        ctx.set_variant(st, int(v))
    return dict(chosen)


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
    """Merge `chosen` into the table's entry for `key` (stages it does not name are kept)."""
    table = {}
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
    table[str(key)] = {**table.get(str(key), {}), **chosen}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load(path: os.PathLike, key):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(str(key))
