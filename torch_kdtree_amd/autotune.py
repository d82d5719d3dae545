"""Leaf-size (tree depth) selection for GPU trees.

The reference picks ``levels = max(1, min(13, int(log2 N) - 3))`` (``nn_distance.py:146-147``),
i.e. leaves of 8-16 points up to N = 2^16 and ever larger leaves beyond (488 points per leaf at
N = 4M).  The traversal kernel here is fastest with small leaves, so ``levels="auto"`` asks this
module for a depth.  The choice comes from ``tuning.json`` next to this file when it holds an
entry for the dims/dtype (written by :func:`tune` from measured query times), else from the
built-in leaf-size table below.
"""
from __future__ import annotations

import json
import math
import os
import time

import torch

_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

# Target points per leaf when no measurement is cached.
DEFAULT_LEAF = {1: 16, 2: 16, 3: 32, 8: 32}


def _levels_for_leaf(n: int, leaf: float) -> int:
    lv = int(round(math.log2(max(n / float(leaf), 1.0))))
    lv = max(1, min(24, lv))
    while lv > 1 and n < (1 << (lv - 1)):
        lv -= 1
    return lv


def _key(dims: int, dtype) -> str:
    return "d%d_%s" % (dims, str(dtype).replace("torch.", ""))


def _load_cache() -> dict:
    try:
        with open(_CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def target_leaf_size(dims: int, dtype=torch.float32) -> float:
    entry = _load_cache().get(_key(dims, dtype))
    if entry and "leaf_size" in entry:
        return float(entry["leaf_size"])
    return float(DEFAULT_LEAF.get(dims, 16))


def pick_levels(n: int, dims: int, device=None, dtype=torch.float32) -> int:
    """Tree depth for ``n`` points: the reference default on CPU trees, tuned on GPU trees."""
    if device is not None and torch.device(device).type == "cpu":
        return int(max(1, min(13, int(math.log(n) / math.log(2)) - 3)))
    return _levels_for_leaf(n, target_leaf_size(dims, dtype))


def tune(points: torch.Tensor, queries: torch.Tensor, k: int, leaf_sizes=(4, 6, 8, 12, 16, 24, 32),
         repeats: int = 3, write: bool = True) -> dict:
    """Times GPU queries for several leaf sizes and caches the fastest one for this dims/dtype."""
    from .nn_distance import build_kd_tree

    n, dims = points.shape
    results = {}
    for leaf in leaf_sizes:
        lv = _levels_for_leaf(n, leaf)
        tree = build_kd_tree(points, device=queries.device, levels=lv)
        tree.query(queries, nr_nns_searches=k)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            tree.query(queries, nr_nns_searches=k)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        results[leaf] = {"levels": lv, "seconds": best}
    leaf_best = min(results, key=lambda s: results[s]["seconds"])
    out = {"leaf_size": leaf_best, "k": k, "n": n, "measured": {str(s): v for s, v in results.items()}}
    if write:
        cache = _load_cache()
        cache[_key(dims, points.dtype)] = out
        with open(_CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return out
