"""The autotuner (torch_kdtree_amd/autotune.py) on the GPU: tune() on a small cloud measures leaf
sizes and every prebuilt launch variant, requires identical results across them, and stores its
choice; the choice is read back by build_kd_tree(levels="auto") and by the queries, whose results
stay equal to the oracle's."""
import json

import numpy as np
import pytest
import torch

import torch_kdtree_amd as tk
from torch_kdtree_amd import autotune, torch_knn
from tests._util import assert_same, gpu_query, oracle_for

pytestmark = pytest.mark.gpu


def test_tune_writes_and_reads_back(cuda, tmp_path, monkeypatch):
    cache = tmp_path / "tuning.json"
    monkeypatch.setenv("TORCH_KDTREE_AMD_TUNING", str(cache))
    rng = np.random.default_rng(21)
    ref = torch.from_numpy(rng.random((200_000, 3)).astype(np.float32))
    qry = torch.from_numpy((rng.random((100_000, 3)) * 1.2 - 0.1).astype(np.float32)).to(cuda)
    n_var = len(torch_knn.tree_variants())
    assert n_var > 1
    out = autotune.tune(ref, qry, 16, leaf_sizes=(16, 32), repeats=1)
    assert set(out["measured_variants"]) == set(range(n_var))
    assert out["leaf_size"] in (16, 32) and 0 <= out["variant"] < n_var
    data = json.loads(cache.read_text())
    entry = data["d3_float32"]
    assert entry["leaf_size"] == out["leaf_size"] and entry["variants"]["16"] == out["variant"]
    # read back: depth for levels="auto", variant for k in 9..16, nothing for other k
    assert autotune.target_leaf_size(3, torch.float32) == out["leaf_size"]
    assert autotune.variant_for(3, torch.float32, 16) == out["variant"]
    assert autotune.variant_for(3, torch.float32, 12) == out["variant"]
    assert autotune.variant_for(3, torch.float32, 4) == 0
    tree = tk.build_kd_tree(ref, device=cuda, levels="auto")
    assert tree.levels == autotune._levels_for_leaf(200_000, out["leaf_size"])
    q_np = qry.cpu().numpy()
    dist, sidx, oidx = gpu_query(tree, q_np, 16, cuda)
    assert tree.kdtree.get_variant() == out["variant"]
    od, oi = oracle_for(tree).search(q_np, 16)
    assert_same(dist, od, "distances with the tuned variant")
    assert_same(sidx, oi, "structured indices with the tuned variant")


def test_tune_lower_index_policy_kept_apart(cuda, tmp_path, monkeypatch):
    """Tuning under the lower-index policy stores its launch variant under its own key and leaves the
    reference policy's variants and the leaf size alone; queries of each policy read their own."""
    cache = tmp_path / "tuning.json"
    monkeypatch.setenv("TORCH_KDTREE_AMD_TUNING", str(cache))
    cache.write_text(json.dumps({"d3_float32": {"leaf_size": 32, "variants": {"8": 2}}}))
    rng = np.random.default_rng(22)
    ref = torch.from_numpy(rng.random((100_000, 3)).astype(np.float32))
    qry = torch.from_numpy(rng.random((50_000, 3)).astype(np.float32)).to(cuda)
    out = autotune.tune(ref, qry, 8, leaf_sizes=(32,), repeats=1, tie_break="lower_index")
    entry = json.loads(cache.read_text())["d3_float32"]
    assert entry["leaf_size"] == 32 and entry["variants"] == {"8": 2}
    assert entry["variants_lower_index"]["8"] == out["variant"]
    assert autotune.variant_for(3, torch.float32, 8, "lower_index") == out["variant"]
    assert autotune.variant_for(3, torch.float32, 8) == 2
    low = tk.build_kd_tree(ref, device=cuda, levels="auto", tie_break="lower_index")
    gpu_query(low, qry.cpu().numpy(), 8, cuda)
    assert low.kdtree.get_variant() == out["variant"]
    refp = tk.build_kd_tree(ref, device=cuda, levels="auto")
    gpu_query(refp, qry.cpu().numpy(), 8, cuda)
    assert refp.kdtree.get_variant() == 2


@pytest.mark.parametrize("tie", ["reference", "lower_index", "lower_index_mode2"])
@pytest.mark.parametrize("k", [6, 16])
def test_every_variant_gives_the_same_results(cuda, tie, k):
    """Each prebuilt launch variant (chunk, occupancy target, block) returns variant 0's results, bit
    for bit, which are the oracle's (reference policy) or brute force's (lower-index policy, with
    the list chosen from the cloud, or the (distance, index) list forced)."""
    import oracle
    rng = np.random.default_rng(31 + k)
    ref = rng.random((120_000, 3)).astype(np.float32)
    qry = (rng.random((40_000, 3)) * 1.2 - 0.1).astype(np.float32)
    mode = 2 if tie.endswith("mode2") else 0
    tie = "lower_index" if tie.startswith("lower_index") else tie
    base = None
    for v in range(len(torch_knn.tree_variants())):
        tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=12, tie_break=tie, variant=v)
        tree.kdtree.set_lower_list(mode)
        out = gpu_query(tree, qry, k, cuda)
        assert tree.kdtree.get_variant() == v
        if base is None:
            base = out
            if tie == "reference":
                od, oi = oracle_for(tree).search(qry, k)
                assert_same(out[0], od, "distances")
                assert_same(out[1], oi, "structured indices")
            else:
                bd, bi = oracle.bruteforce(ref, qry, k)
                assert_same(out[0], bd, "distances")
                assert_same(out[2], bi.astype(np.int64), "original indices")
        for a, b, what in zip(out, base, ("distances", "structured indices", "original indices")):
            assert_same(a, b, "variant %d %s" % (v, what))
