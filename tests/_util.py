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


def nonfinite_cloud(rng, n, d, dtype, kind):
    """n points N(0,1)*1e3 of dtype with non-finite values planted: "one_nan" one NaN coordinate;
    "nan_rows" NaN in one or more coordinates of about 1% of the rows; "inf_rows" +-inf likewise;
    "mixed" both; "overflow" a cluster near the largest finite value, whose even-count medians
    overflow to inf (0.5 * (a + b) with a + b rounded in the point type)."""
    pts = (rng.standard_normal((n, d)) * 1e3).astype(dtype)
    big = np.finfo(dtype).max
    rows = rng.choice(n, size=max(1, n // 100), replace=False)
    if kind == "one_nan":
        pts[rng.integers(n), rng.integers(d)] = np.nan
    elif kind in ("nan_rows", "inf_rows", "mixed"):
        for j, r in enumerate(rows):
            m = rng.random(d) < 0.5
            if not m.any():
                m[rng.integers(d)] = True
            if kind == "nan_rows" or (kind == "mixed" and j % 2 == 0):
                pts[r, m] = np.nan
            else:
                pts[r, m] = np.where(rng.random(int(m.sum())) < 0.5, np.inf, -np.inf)
    elif kind == "overflow":
        pts[rows] = (big * (0.9 + 0.1 * rng.random((len(rows), d)))).astype(dtype)
    else:
        raise ValueError(kind)
    return pts


def nonfinite_queries(rng, q, d, dtype):
    """q queries N(0,1)*1e3 with NaN rows, +-inf rows and a few at the largest finite value."""
    qry = (rng.standard_normal((q, d)) * 1e3).astype(dtype)
    qry, _ = with_nan_rows(qry, rng, frac=0.05)
    r = rng.choice(q, size=max(2, q // 40), replace=False)
    qry[r[0::2], 0] = np.inf
    qry[r[1::2], d - 1] = -np.inf
    qry[rng.integers(q)] = np.finfo(dtype).max
    return qry


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
