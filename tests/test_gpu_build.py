"""GPU tree build (builder="gpu", tk_tree_build_gpu / tk_tree_build_many_gpu).

On inputs without equal coordinates on any axis the GPU build must equal the CPU build (the
reference's layout) bit for bit: structured points, indices, split values, leaf offsets, and
therefore every query result.  With tied coordinates the reference orders tied points by an
unstable sort, so only the tree's invariants and the neighbour distances are checked, plus the
tie-independent lower-index results, which must equal those of the CPU-built tree exactly.
"""
import numpy as np
import pytest
import torch

import oracle
import workloads
import torch_kdtree_amd as tk
from torch_kdtree_amd import torch_knn
from tests._util import assert_same

pytestmark = pytest.mark.gpu


def _distinct_points(rng, n, d, dtype):
    """n points whose coordinates on every axis are pairwise distinct (no ties in any node)."""
    cols = []
    for _ in range(d):
        x = rng.standard_normal(n) * 100.0
        x = x.astype(dtype)
        # make the column distinct: replace repeats by fresh draws until unique
        while len(np.unique(x)) != n:
            _, first = np.unique(x, return_index=True)
            dup = np.setdiff1d(np.arange(n), first)
            x[dup] = (rng.standard_normal(len(dup)) * 100.0).astype(dtype)
        cols.append(x)
    return np.ascontiguousarray(np.stack(cols, 1))


def _same_tree(a, b, what):
    assert_same(a.kdtree.get_structured_points(), b.kdtree.get_structured_points(), what + " structured points")
    assert_same(a.kdtree.get_shuffled_inds(), b.kdtree.get_shuffled_inds(), what + " indices")
    for x, y, name in zip(a.kdtree.get_nodes(), b.kdtree.get_nodes(), ("axes", "split values", "leaf offsets")):
        assert_same(x, y, what + " " + name)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
@pytest.mark.parametrize("n,levels", [(1, 1), (2, 2), (37, 6), (1000, 5), (4097, 12), (50_001, 11)])
def test_gpu_build_equals_cpu_build_without_ties(cuda, dtype, d, n, levels):
    rng = np.random.default_rng(n * 10 + d + (dtype == np.float64))
    pts = _distinct_points(rng, n, d, dtype)
    cpu = tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels=levels)
    gpu = tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels=levels, builder="gpu")
    _same_tree(gpu, cpu, "n=%d d=%d L=%d" % (n, d, levels))
    q = torch.from_numpy(_distinct_points(rng, 500, d, dtype)).to(cuda)
    k = min(8, n)
    dg, ig = gpu.query(q, nr_nns_searches=k)
    dc, ic = cpu.query(q, nr_nns_searches=k)
    torch.cuda.synchronize()
    assert torch.equal(dg, dc) and torch.equal(ig, ic)


def test_gpu_build_with_ties_is_a_valid_tree(cuda):
    """Integer grid: many equal coordinates.  Node sizes equal the reference's; every node splits
    its points at its median; the lower-index results equal the CPU-built tree's exactly."""
    rng = np.random.default_rng(5)
    n, d, L = 20_000, 3, 9
    pts = rng.integers(-20, 21, (n, d)).astype(np.float32)
    cpu = tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels=L)
    gpu = tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels=L, builder="gpu")
    ax, sp, lo = gpu.kdtree.get_nodes()
    _, _, lo_c = cpu.kdtree.get_nodes()
    assert_same(lo, lo_c, "leaf offsets")
    st = gpu.kdtree.get_structured_points()
    sh = gpu.kdtree.get_shuffled_inds()
    assert_same(st, pts[sh], "structured points are the indexed points")
    assert np.array_equal(np.sort(sh), np.arange(n))
    # node ranges from the leaf offsets: node i at depth l covers leaves [first, last)
    nn = (1 << L) - 1
    for i in range(nn):
        l = int(np.floor(np.log2(i + 1)))
        first = (i + 1 - (1 << l)) << (L - l)
        last = first + (1 << (L - l))
        mid = first + (1 << (L - l - 1))
        a = l % d
        left = st[lo[first]:lo[mid], a]
        right = st[lo[mid]:lo[last], a]
        allv = np.sort(st[lo[first]:lo[last], a])
        m = allv[len(allv) // 2] if len(allv) % 2 else np.float32((allv[len(allv) // 2 - 1] + allv[len(allv) // 2]) * 0.5)
        assert sp[i] == m, i
        assert (left.size == 0 or left.max() <= sp[i]) and (right.size == 0 or right.min() >= sp[i]), i
    q = torch.from_numpy(rng.integers(-22, 23, (3000, d)).astype(np.float32)).to(cuda)
    for tie in ("lower_index",):
        cpu.tie_break = gpu.tie_break = tie
        dg, ig = gpu.query(q, nr_nns_searches=10)
        dc, ic = cpu.query(q, nr_nns_searches=10)
        torch.cuda.synchronize()
        assert torch.equal(dg, dc) and torch.equal(ig, ic), tie
    cpu.tie_break = gpu.tie_break = "reference"
    dg, ig = gpu.query(q, nr_nns_searches=10)
    od, _ = oracle.bruteforce(pts, q.cpu().numpy(), 10)
    assert_same(dg.cpu().numpy(), od, "reference-policy distances vs brute force")
    # every returned index has the returned distance
    diff = q.cpu().numpy()[:, None, :] - pts[ig.cpu().numpy()]
    assert_same((diff * diff).sum(-1).astype(np.float32), dg.cpu().numpy(), "indices match distances")


def test_gpu_build_of_the_plain_workload(cuda):
    """The 4,000,000-point tree of the plain GPU workload (float32 coordinates tie often): node
    sizes equal, lower-index results on 65,536 of its queries equal to the CPU-built tree's."""
    w = workloads.plain_gpu()
    ref = torch.from_numpy(w["ref"])
    cpu = tk.build_kd_tree(ref, device=cuda, levels=17, tie_break="lower_index")
    gpu = tk.build_kd_tree(ref, device=cuda, levels=17, tie_break="lower_index", builder="gpu")
    assert_same(gpu.kdtree.get_nodes()[2], cpu.kdtree.get_nodes()[2], "leaf offsets")
    half = w["halves"][0]
    q = torch.from_numpy(np.concatenate([w["queries"][:32768], w["queries"][half:half + 32768]])).to(cuda)
    dg, ig = gpu.query(q, nr_nns_searches=16)
    dc, ic = cpu.query(q, nr_nns_searches=16)
    torch.cuda.synchronize()
    assert torch.equal(dg, dc) and torch.equal(ig, ic)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gpu_forest_build_equals_per_tree_cpu_builds(cuda, dtype):
    rng = np.random.default_rng(21 + (dtype == np.float64))
    sizes = [1, 5, 300, 4000, 1023, 60_000, 2]
    clouds = [_distinct_points(rng, n, 3, dtype) for n in sizes]
    batch = tk.build_kd_tree_batch([torch.from_numpy(c) for c in clouds], device=cuda, builder="gpu")
    ref = tk.build_kd_tree_batch([torch.from_numpy(c) for c in clouds], device=cuda)
    assert batch.levels == ref.levels
    for c, (pts, L) in enumerate(zip(clouds, batch.levels)):
        single = tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels=L)
        built = tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels=L, builder="gpu")
        _same_tree(built, single, "cloud %d" % c)
    queries = [torch.from_numpy(_distinct_points(rng, 300, 3, dtype)) for _ in sizes]
    ks = [min(4, n) for n in sizes]
    d1, i1 = batch.query(queries, ks)
    d2, i2 = ref.query(queries, ks)
    torch.cuda.synchronize()
    for c in range(len(sizes)):
        assert torch.equal(d1[c], d2[c]) and torch.equal(i1[c], i2[c]), c


def test_gpu_build_errors(cuda):
    pts = np.zeros((10, 3), np.float32)
    pts[:, 0] = np.arange(10)
    with pytest.raises(RuntimeError, match="Ran out of points"):
        tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels=6, builder="gpu")
    bad = pts.copy()
    bad[3, 1] = np.nan
    with pytest.raises(RuntimeError, match="NaN"):
        tk.build_kd_tree(torch.from_numpy(bad), device=cuda, levels=2, builder="gpu")
    with pytest.raises(RuntimeError):
        torch_knn.build_tree_gpu(0, 10, 5, "float32", 2)
