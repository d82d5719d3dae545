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


def with_nan_rows(qry, rng, frac=0.1):
    """A copy of qry with NaN in one or more coordinates of about frac of its rows; returns (copy, rows)."""
    q = qry.copy()
    rows = np.sort(rng.choice(len(q), size=max(1, int(len(q) * frac)), replace=False))
    for r in rows:
        m = rng.random(q.shape[1]) < 0.5
        if not m.any():
            m[rng.integers(q.shape[1])] = True
        q[r, m] = np.nan
    return q, rows


def assert_same(a, b, what):
    """Exact equality; NaN equals NaN (at the same positions) in floating-point arrays."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        raise AssertionError("%s: shapes differ, %s vs %s" % (what, a.shape, b.shape))
    nan_ok = a.dtype.kind == "f" and b.dtype.kind == "f"
    if not np.array_equal(a, b, equal_nan=nan_ok):
        same = (a == b) | (np.isnan(a) & np.isnan(b)) if nan_ok else (a == b)
        bad = np.argwhere(~same)
        raise AssertionError("%s differs at %d positions, first %s: %r vs %r" % (
            what, len(bad), bad[0].tolist(), a[tuple(bad[0])], b[tuple(bad[0])]))
