"""Autotuning of the GPU search: leaf size (tree depth) and traversal launch variant.

The reference fixes its launch shape at compile time (``kdtree_g.cu:6,281-283``) and picks
``levels = max(1, min(13, int(log2 N) - 3))`` (``nn_distance.py:146-147``), i.e. ever larger leaves
beyond N = 2^16 (488 points per leaf at N = 4M).  Here two choices are measured on the GPU:

* the leaf size, which sets the depth of trees built with ``levels="auto"``;
* the launch variant of the traversal kernel (``torch_knn.tree_variants()``): the leaf-scan chunk
  staged in LDS per pass, the occupancy target in waves per SIMD and the threads per block.  The
  variants are prebuilt instantiations; the non-default ones exist for float searches at d = 3
  with lists of 8 or 16 entries (k = 5..16).

:func:`tune` times queries over the candidates, checks that every candidate returns the same
results, and stores the fastest in ``tuning.json`` next to this file (or the path in
``TORCH_KDTREE_AMD_TUNING``).  ``build_kd_tree(levels="auto")`` reads the leaf size from there, and
every GPU query reads the variant for its dims, dtype and list width.  Without an entry the
built-in leaf-size table below and variant 0 (the built-in launch shape) apply.
"""
from __future__ import annotations

import json
import math
import os
import time

import torch

_DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

# Target points per leaf when no measurement is cached.
DEFAULT_LEAF = {1: 16, 2: 16, 3: 32, 8: 32}

_memo = {"path": None, "mtime": None, "data": {}}


def cache_path() -> str:
    return os.environ.get("TORCH_KDTREE_AMD_TUNING", _DEFAULT_CACHE)


def _levels_for_leaf(n: int, leaf: float) -> int:
    lv = int(round(math.log2(max(n / float(leaf), 1.0))))
    lv = max(1, min(24, lv))
    while lv > 1 and n < (1 << (lv - 1)):
        lv -= 1
    return lv


def _dtype_name(dtype) -> str:
    return str(dtype).replace("torch.", "")


def _key(dims: int, dtype) -> str:
    return "d%d_%s" % (dims, _dtype_name(dtype))


def _list_width(k: int) -> int:
    kb = 1
    while kb < k:
        kb <<= 1
    return kb


def _load_cache() -> dict:
    """The cache file's contents, re-read only when the file changes."""
    path = cache_path()
    try:
        mtime = os.path.getmtime(path)
    except OSError:
        return {}
    if _memo["path"] != path or _memo["mtime"] != mtime:
        try:
            with open(path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
        _memo.update(path=path, mtime=mtime, data=data)
    return _memo["data"]


def _store(key: str, entry: dict) -> None:
    path = cache_path()
    cache = dict(_load_cache())
    cache[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


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


def _variants_key(tie_break: str) -> str:
    # launch variants are measured per tie policy: their kernels differ (the lower-index list holds
    # 64-bit keys), and so does the fastest launch shape
    return "variants" if tie_break == "reference" else "variants_" + tie_break


def variant_for(dims: int, dtype, k: int, tie_break: str = "reference") -> int:
    """The cached launch variant for GPU queries of this dims, dtype, k and tie policy (0 without an
    entry)."""
    entry = _load_cache().get(_key(dims, dtype))
    if not entry:
        return 0
    return int(entry.get(_variants_key(tie_break), {}).get(str(_list_width(k)), 0))


def _time_query(tree, queries, k, repeats):
    tree.query(queries, nr_nns_searches=k)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        tree.query(queries, nr_nns_searches=k)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(points: torch.Tensor, queries: torch.Tensor, k: int, leaf_sizes=(8, 16, 24, 32, 48),
         variants=None, repeats: int = 3, write: bool = True, tie_break: str = "reference") -> dict:
    """Measures GPU queries of ``queries`` against ``points`` (k neighbours) for each leaf size with the
    built-in launch shape, then each launch variant on the fastest depth, and keeps the fastest of
    each.  Every variant's distances and indices must equal variant 0's on the same tree, and every
    depth's distances must equal the first depth's (the neighbour sets are exact; only tie order and
    structured indices depend on the tree); a mismatch raises.  With ``write`` the choice is stored
    for this dims/dtype (leaf size, measured under the reference tie policy only) and, for the
    tie policy ``tie_break``, list width (variant)."""
    from .nn_distance import build_kd_tree
    from . import torch_knn

    n, dims = points.shape
    dev = queries.device
    if dev.type != "cuda":
        raise RuntimeError("tune() measures GPU queries: pass queries on a GPU device")
    if variants is None:
        variants = range(len(torch_knn.tree_variants()))
    leaf_res, ref_dist = {}, None
    for leaf in leaf_sizes:
        lv = _levels_for_leaf(n, leaf)
        tree = build_kd_tree(points, device=dev, levels=lv, variant=0, tie_break=tie_break)
        d, _ = tree.query(queries, nr_nns_searches=k)
        if ref_dist is None:
            ref_dist = d.clone()
        elif not torch.equal(d, ref_dist):
            raise RuntimeError("tune: leaf size %s changed the distances" % leaf)
        leaf_res[leaf] = {"levels": lv, "seconds": _time_query(tree, queries, k, repeats)}
    leaf_best = min(leaf_res, key=lambda s: leaf_res[s]["seconds"])
    tree = build_kd_tree(points, device=dev, levels=leaf_res[leaf_best]["levels"], variant=0, tie_break=tie_break)
    table = torch_knn.tree_variants()
    var_res, ref_out = {}, None
    for v in variants:
        tree.variant = int(v)
        res_i = torch.empty((queries.shape[0], k), dtype=torch.int32, device=dev)
        d, o = tree.query(queries, nr_nns_searches=k, result_idx=res_i)
        out = (d.clone(), res_i.clone(), o.clone())
        if ref_out is None:
            ref_out = out
        elif not all(torch.equal(a, b) for a, b in zip(out, ref_out)):
            raise RuntimeError("tune: launch variant %d changed the results" % v)
        var_res[int(v)] = dict(table[int(v)], seconds=_time_query(tree, queries, k, repeats))
    var_best = min(var_res, key=lambda v: var_res[v]["seconds"])
    kb = _list_width(k)
    entry = dict(_load_cache().get(_key(dims, points.dtype), {}))
    if tie_break == "reference":
        entry.update(leaf_size=leaf_best, n=int(n), queries=int(queries.shape[0]),
                     measured_leaf={str(s): r for s, r in leaf_res.items()})
    vk = _variants_key(tie_break)
    entry.setdefault(vk, {})[str(kb)] = var_best
    entry.setdefault("measured_" + vk, {})[str(kb)] = {str(v): r for v, r in var_res.items()}
    if write:
        _store(_key(dims, points.dtype), entry)
    return {"leaf_size": leaf_best, "variant": var_best, "list_width": kb, "k": k, "tie_break": tie_break,
            "measured_leaf": leaf_res, "measured_variants": var_res}
