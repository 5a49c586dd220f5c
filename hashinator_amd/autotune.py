"""Launch-shape autotuner for the batched multi-map kernel and the probe kernels.

Candidates are timed on the caller's own data (HIP events on the stream the
kernels run on), the fastest wins, and the choice is cached per workload
signature for the life of the process (and in a JSON file when a path is
given).  A candidate that errors out is skipped.
"""
from __future__ import annotations

import json
import os

import torch

from ._lib import HN_BATCH_ALL, hn_tune

_CACHE: dict = {}

BATCH_CANDIDATES = [(w, b) for w in (4, 8, 16, 32) for b in (256, 512, 1024)]


def _time(fn, reps):
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def batch_signature(bm) -> str:
    return f"batch/k{bm.kb}/n{bm.n}/slots{bm.total_slots}/ins{int(bm.ins_len.sum())}"


def tune_batch(bm, ops=HN_BATCH_ALL, candidates=None, reps=2, cache_path: str | None = None):
    """Best (probe_width, block_size) for `bm` running `ops`; returns (hn_tune, table of timings)."""
    key = batch_signature(bm) + f"/ops{ops}"
    if cache_path and os.path.exists(cache_path) and key not in _CACHE:
        try:
            _CACHE.update(json.load(open(cache_path)))
        except (OSError, ValueError):
            pass
    if key in _CACHE:
        w, b = _CACHE[key]["best"]
        return hn_tune(w, b, 0, 0), _CACHE[key]["table"]
    table = {}
    for w, b in candidates or BATCH_CANDIDATES:
        t = hn_tune(w, b, 0, 0)

        def step():
            bm.init()
            bm.run(ops, t)

        try:
            step()
            torch.cuda.synchronize()
            table[f"{w}x{b}"] = _time(step, reps)
        except Exception:  # an invalid shape for this device: skip it
            continue
    best = min(table, key=table.get)
    w, b = (int(x) for x in best.split("x"))
    _CACHE[key] = {"best": [w, b], "table": table}
    if cache_path:
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        json.dump(_CACHE, open(cache_path, "w"), indent=1)
    return hn_tune(w, b, 0, 0), table
