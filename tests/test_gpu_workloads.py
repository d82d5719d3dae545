"""GPU parity on the benchmark workloads at their full sizes (trees and query sets of the real
shape), against the oracle on the same trees: bit-exact distances and indices.  The plain GPU
workload's 16,000,000 queries and the batching workload are in test_gpu_fullsize.py."""
import numpy as np
import pytest
import torch

import workloads
import torch_kdtree_amd as tk
from tests._util import assert_same, gpu_query, oracle_for

pytestmark = pytest.mark.gpu


def _check(tree, queries, k, cuda, what):
    gd, gi, go = gpu_query(tree, np.ascontiguousarray(queries), k, cuda)
    od, oi = oracle_for(tree).search(np.ascontiguousarray(queries), k)
    assert_same(gd, od, what + " distances")
    assert_same(gi, oi, what + " structured indices")
    sh = tree.kdtree.get_shuffled_inds()
    assert_same(go, sh[oi].astype(np.int64), what + " original indices")


def test_differentiation_workload_full(cuda):
    """1,000,000 perturbed-sphere points, all 1,000,000 queries, k=8."""
    w = workloads.diff_sphere()
    tree = tk.build_kd_tree(torch.from_numpy(w["ref"]), device=cuda, levels="auto")
    _check(tree, w["queries"], 8, cuda, "diff_sphere")


def test_d8_workload_full_tree_sample(cuda):
    """d=8: 2,000,000-point anisotropic mixture tree, 16,384 of its queries, k=32."""
    w = workloads.d8_mixture()
    tree = tk.build_kd_tree(torch.from_numpy(w["ref"]), device=cuda, levels="auto")
    _check(tree, w["queries"][:16384], 32, cuda, "d8")


def test_fp64_workload_full(cuda):
    """d=3 fp64: 500,000 points N(0,1)*1e3, all 500,000 queries, k=16."""
    w = workloads.fp64_normal()
    tree = tk.build_kd_tree(torch.from_numpy(w["ref"]), device=cuda, levels="auto")
    _check(tree, w["queries"], 16, cuda, "fp64")
