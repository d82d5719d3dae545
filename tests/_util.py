"""Shared helpers for the parity tests."""
import numpy as np
import torch

import oracle


def gpu_query(tree, queries_np, k, device):
    q = torch.from_numpy(queries_np).to(device)
    res_d = torch.empty((q.shape[0], k), dtype=q.dtype, device=device)
    res_i = torch.empty((q.shape[0], k), dtype=torch.int32, device=device)
    d, inds = tree.query(q, nr_nns_searches=k, result_dists=res_d, result_idx=res_i)
    torch.cuda.synchronize()
    return d.cpu().numpy(), res_i.cpu().numpy(), inds.cpu().numpy()


def oracle_for(tree):
    """Oracle tree with the same arrays as a built TorchKDTree (layout parity is tested separately)."""
    return oracle.tree_from_product(tree)


def assert_same(a, b, what):
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        raise AssertionError("%s differs at %d positions, first %s: %r vs %r" % (
            what, len(bad), bad[0].tolist(), a[tuple(bad[0])], b[tuple(bad[0])]))
