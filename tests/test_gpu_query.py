"""GPU parity: the HIP traversal against the oracle (C++ restatement of the reference CPU path).

Reference policy: distances and indices must equal ``KDTreeKNNSearch`` (src/kdtree.cpp:164-194)
on the same tree, bit for bit, including the order of exact-distance ties.
Lower-index policy: distances and indices must equal the brute-force (distance, index) order.
"""
import numpy as np
import pytest
import torch

import oracle
import workloads
import torch_kdtree_amd as tk
from tests._util import assert_same, gpu_query, nonfinite_cloud, nonfinite_queries, oracle_for, with_nan_rows

pytestmark = pytest.mark.gpu


def _data(rng, n, q, d, dtype, scale=1e3):
    return ((rng.standard_normal((n, d)) * scale).astype(dtype), (rng.standard_normal((q, d)) * scale).astype(dtype))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 16, 33])
def test_reference_policy_matches_oracle(cuda, dtype, d, k):
    rng = np.random.default_rng(100 + d * 7 + k)
    ref, qry = _data(rng, 20000, 3000, d, dtype)
    for levels in (None, 12):
        tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=levels)
        dist, sidx, oidx = gpu_query(tree, qry, k, cuda)
        od, oi = oracle_for(tree).search(qry, k)
        assert_same(dist, od, "distances")
        assert_same(sidx, oi, "structured indices")
        assert_same(oidx, tree.kdtree.get_shuffled_inds()[oi].astype(np.int64), "original indices")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 3, 8])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 8, 32, 100])
def test_lower_index_policy_matches_bruteforce(cuda, dtype, d, k):
    rng = np.random.default_rng(200 + d * 11 + k)
    ref, qry = _data(rng, 6000, 1500, d, dtype)
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=9, tie_break="lower_index")
    dist, sidx, oidx = gpu_query(tree, qry, k, cuda)
    bd, bi = oracle.bruteforce(ref, qry, k)
    assert_same(dist, bd, "distances")
    assert_same(oidx, bi.astype(np.int64), "original indices")
    assert_same(tree.kdtree.get_shuffled_inds()[sidx], bi, "structured -> original")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 3, 8])
@pytest.mark.parametrize("k", [1, 5, 16, 40])
def test_nan_query_rows(cuda, dtype, d, k):
    """Queries with a NaN coordinate.  Reference policy: the oracle's result (NaN distances on the
    last records of the one leaf the reference scans, then (+inf, 0) when that leaf holds fewer
    than k points; leaves here hold about 31 points).  Lower-index policy: brute force's, which
    never inserts a NaN distance (d < worst, src/nndistance_g.cu:57): (+inf, 0) in every slot.
    The other rows are as without them.  k = 40 uses the LDS list."""
    rng = np.random.default_rng(700 + d * 3 + k + (dtype == np.float64))
    ref, qry = _data(rng, 4000, 800, d, dtype)
    qn, rows = with_nan_rows(qry, rng)
    keep = np.setdiff1d(np.arange(len(qry)), rows)
    for policy in ("reference", "lower_index"):
        tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=7, tie_break=policy)
        sh = tree.kdtree.get_shuffled_inds()
        dist, sidx, oidx = gpu_query(tree, qn, k, cuda)
        d0, s0, o0 = gpu_query(tree, qry, k, cuda)
        assert_same(dist[keep], d0[keep], policy + " distances of rows without NaN")
        assert_same(sidx[keep], s0[keep], policy + " indices of rows without NaN")
        if policy == "reference":
            assert np.isnan(dist[rows, 0]).all()
            od, oi = oracle_for(tree).search(qn, k)
            assert_same(dist, od, "distances")
            assert_same(sidx, oi, "structured indices")
            assert_same(oidx, sh[oi].astype(np.int64), "original indices")
        else:
            bd, bi = oracle.bruteforce(ref, qn, k)
            assert_same(dist, bd, "distances")
            assert_same(oidx, bi.astype(np.int64), "original indices")
            assert_same(sh[sidx], bi, "structured -> original")
            assert np.isposinf(dist[rows]).all() and (oidx[rows] == 0).all()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 3, 8])
@pytest.mark.parametrize("kind", ["one_nan", "nan_rows", "mixed", "overflow"])
def test_nonfinite_cloud(cuda, dtype, d, kind):
    """Reference clouds with NaN or infinite coordinates, or medians that overflow.  Reference
    policy: the oracle's result bit for bit, NaN distances included (the reference's own walk and
    list, knn_tree_exact_kernel); lower-index policy: brute force's (NaN distances never
    inserted), with pruning that a NaN split never stops.  Queries hold NaN, +-inf and huge rows."""
    rng = np.random.default_rng(900 + d + 5 * (dtype == np.float64) + 17 * len(kind))
    ref = nonfinite_cloud(rng, 3000, d, dtype, kind)
    qry = nonfinite_queries(rng, 600, d, dtype)
    for levels in (1, 5, 7):
        tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=levels)
        ot = oracle_for(tree)
        assert tree.kdtree.is_finite() == (np.isfinite(ref).all() and np.isfinite(ot.export()["median"]).all())
        sh = tree.kdtree.get_shuffled_inds()
        for k in (1, 5, 16, 40):
            dist, sidx, oidx = gpu_query(tree, qry, k, cuda)
            od, oi = ot.search(qry, k)
            assert_same(dist, od, "levels %d k %d distances" % (levels, k))
            assert_same(sidx, oi, "levels %d k %d structured indices" % (levels, k))
            assert_same(oidx, sh[oi].astype(np.int64), "levels %d k %d original indices" % (levels, k))
        low = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=levels, tie_break="lower_index")
        for k in (1, 5, 16, 40):
            dist, sidx, oidx = gpu_query(low, qry, k, cuda)
            bd, bi = oracle.bruteforce(ref, qry, k)
            assert_same(dist, bd, "lower-index levels %d k %d distances" % (levels, k))
            assert_same(oidx, bi.astype(np.int64), "lower-index levels %d k %d original indices" % (levels, k))


def test_unaligned_output_rows(cuda):
    """Output arrays that start 4 bytes past a 16-byte boundary (written with 4-byte stores) hold
    what aligned ones (16-byte stores when k is the list width) hold."""
    rng = np.random.default_rng(17)
    ref, qry = _data(rng, 20000, 3000, 3, np.float32)
    q = torch.from_numpy(qry).to(cuda)
    for policy in ("reference", "lower_index"):
        tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=9, tie_break=policy)
        for k in (4, 5, 16, 32):
            d0, s0, o0 = gpu_query(tree, qry, k, cuda)
            bd = torch.empty(q.shape[0] * k + 1, dtype=torch.float32, device=cuda)[1:].view(-1, k)
            bi = torch.empty(q.shape[0] * k + 1, dtype=torch.int32, device=cuda)[1:].view(-1, k)
            assert bd.data_ptr() % 16 == 4 and bi.data_ptr() % 16 == 4
            d, o = tree.query(q, nr_nns_searches=k, result_dists=bd, result_idx=bi)
            torch.cuda.synchronize()
            assert_same(bd.cpu().numpy(), d0, "%s k=%d distances" % (policy, k))
            assert_same(bi.cpu().numpy(), s0, "%s k=%d structured indices" % (policy, k))
            assert_same(o.cpu().numpy(), o0, "%s k=%d original indices" % (policy, k))


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
@pytest.mark.parametrize("k", [5, 8, 12, 16, 32])
def test_lower_index_ties_among_distinct_distances(cuda, d, k, mode):
    """Lower-index policy, float lists of 8 to 32, under each list mode (0: chosen from the cloud,
    1: tie-flagged list, each unsure lane searched again, 2: (distance, index) list): a cloud where
    one point in ten has an exact duplicate (every distance to the pair ties) and queries on and off
    the points, so that under mode 1 lanes whose lists meet a tie share wavefronts with lanes that
    meet none."""
    rng = np.random.default_rng(900 + d * 13 + k)
    base = rng.standard_normal((18000, d)).astype(np.float32)
    dup = base[rng.choice(len(base), 1800, replace=False)]
    ref = np.concatenate([base, dup])[rng.permutation(len(base) + len(dup))]
    qry = np.concatenate([rng.standard_normal((3000, d)).astype(np.float32),
                          ref[rng.choice(len(ref), 1000, replace=False)]])
    for levels in (None, 10):
        tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=levels, tie_break="lower_index")
        tree.kdtree.set_lower_list(mode)
        assert tree.kdtree.get_lower_list() == (mode, True)  # the sample sees the duplicates
        dist, sidx, oidx = gpu_query(tree, qry, k, cuda)
        bd, bi = oracle.bruteforce(ref, qry, k)
        assert_same(dist, bd, "distances")
        assert_same(oidx, bi.astype(np.int64), "original indices")
        assert_same(tree.kdtree.get_shuffled_inds()[sidx], bi, "structured -> original")


@pytest.mark.parametrize("policy", ["reference", "lower_index", "lower_index_mode1", "lower_index_mode2"])
def test_exact_ties_on_integer_grid(cuda, policy):
    """Points and queries on a small integer lattice: nearly every distance is tied.  Under the
    lower-index policy also with each list mode forced (the automatic choice takes mode 2 here)."""
    rng = np.random.default_rng(7)
    ref = rng.integers(0, 6, (5000, 3)).astype(np.float32)
    qry = (rng.integers(0, 6, (2000, 3)) + 0.5 * rng.integers(0, 2, (2000, 3))).astype(np.float32)
    mode = int(policy[-1]) if policy.endswith(("mode1", "mode2")) else 0
    policy = "lower_index" if policy.startswith("lower_index") else policy
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=8, tie_break=policy)
    tree.kdtree.set_lower_list(mode)
    for k in (1, 7, 20, 64):
        dist, sidx, oidx = gpu_query(tree, qry, k, cuda)
        if policy == "reference":
            od, oi = oracle_for(tree).search(qry, k)
            assert_same(dist, od, "distances")
            assert_same(sidx, oi, "structured indices")
        else:
            bd, bi = oracle.bruteforce(ref, qry, k)
            assert_same(dist, bd, "distances")
            assert_same(oidx, bi.astype(np.int64), "original indices")


def test_lower_list_choice_from_the_cloud(cuda):
    """The upload's sample test: continuous clouds take the tie-flagged list (mode 0 resolves to 1),
    lattices, rounded coordinates, duplicated points and non-finite trees the (distance, index) list;
    an unknown mode is refused."""
    rng = np.random.default_rng(17)
    cont = rng.random((50000, 3)).astype(np.float32)
    cases = {
        "uniform": (cont, False),
        "normal": ((rng.standard_normal((50000, 3)) * 1e3).astype(np.float32), False),
        "lattice": (rng.integers(0, 64, (50000, 3)).astype(np.float32), True),
        "rounded": (np.round(cont, 3), True),
        "duplicated": (np.concatenate([cont, cont[:5000]]), True),
        "nan": (np.where(np.arange(50000)[:, None] == 7, np.nan, cont).astype(np.float32), True),
    }
    for name, (ref, tie_prone) in cases.items():
        tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, tie_break="lower_index")
        assert tree.kdtree.get_lower_list() == (0, tie_prone), name
    with pytest.raises(RuntimeError):
        tree.kdtree.set_lower_list(3)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
def test_record_group_boxes_big_leaves(cuda, dtype, d):
    """Trees whose leaves hold 64 or more points on average carry a box per run of 16 records of a
    leaf, and a leaf scan skips the runs whose box lies at least the worst distance away.  Random
    points with queries on and off them, a tie-heavy integer lattice with duplicate points, and
    random points with a few duplicates, at 3,000 and 375 points a leaf, under every lower-index
    list mode: every result equals the oracle's (reference policy) and brute force's (lower-index
    policy)."""
    rng = np.random.default_rng(40 + d)
    rnd_ref, rnd_q = _data(rng, 12000, 1500, d, dtype)
    lat_ref = rng.integers(0, 5, (12000, d)).astype(dtype)
    lat_q = (rng.integers(0, 5, (1500, d)) + 0.5 * rng.integers(0, 2, (1500, d))).astype(dtype)
    # a few exact duplicates, too few for the upload's sample test: lower-index float searches keep
    # the tie-flagged list, and the lanes at the duplicates are searched again
    dup = rnd_ref[rng.choice(len(rnd_ref), 24, replace=False)]
    sd_ref = np.concatenate([rnd_ref, dup])[rng.permutation(len(rnd_ref) + len(dup))]
    sd_q = np.ascontiguousarray(np.concatenate([rnd_q, dup, dup + dtype(0.5)]))
    cases = (("random", rnd_ref, np.ascontiguousarray(np.concatenate([rnd_q, rnd_ref[:300]]))),
             ("lattice", lat_ref, lat_q), ("sparse duplicates", sd_ref, sd_q))
    for name, ref, qry in cases:
        for levels in (2, 5):
            for policy in ("reference", "lower_index", "lower_index_mode1", "lower_index_mode2"):
                mode = int(policy[-1]) if policy.endswith(("mode1", "mode2")) else 0
                policy = "lower_index" if policy.startswith("lower_index") else policy
                tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=levels, tie_break=policy)
                tree.kdtree.set_lower_list(mode)
                if name != "lattice":
                    assert not tree.kdtree.get_lower_list()[1], name
                ot = oracle_for(tree) if policy == "reference" else None
                for k in (1, 5, 16, 32, 40):
                    what = "%s levels %d %s mode %d k %d" % (name, levels, policy, mode, k)
                    dist, sidx, oidx = gpu_query(tree, qry, k, cuda)
                    if policy == "reference":
                        od, oi = ot.search(qry, k)
                        assert_same(dist, od, what + " distances")
                        assert_same(sidx, oi, what + " structured indices")
                    else:
                        bd, bi = oracle.bruteforce(ref, qry, k)
                        assert_same(dist, bd, what + " distances")
                        assert_same(oidx, bi.astype(np.int64), what + " original indices")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
def test_presort_does_not_change_results(cuda, dtype, d):
    rng = np.random.default_rng(8 + d)
    ref, qry = _data(rng, 50000, 20000, d, dtype)
    out = {}
    for mode in ("morton", "leaf", "off"):
        tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=13, presort=mode)
        out[mode] = gpu_query(tree, qry, 16, cuda)
    for mode in ("morton", "leaf"):
        for a, b in zip(out[mode], out["off"]):
            assert_same(a, b, "presort %s vs off" % mode)
    od, oi = oracle_for(tree).search(qry, 16)
    assert_same(out["off"][0], od, "distances")
    assert_same(out["off"][1], oi, "structured indices")


def test_small_and_degenerate_shapes(cuda):
    rng = np.random.default_rng(9)
    ref = rng.standard_normal((5, 3)).astype(np.float32)
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda)
    assert tree.levels == 1
    qry = rng.standard_normal((33, 3)).astype(np.float32)
    for k in (1, 5):
        dist, sidx, _ = gpu_query(tree, qry, k, cuda)
        od, oi = oracle_for(tree).search(qry, k)
        assert_same(dist, od, "distances")
        assert_same(sidx, oi, "indices")
    with pytest.raises(RuntimeError):
        tree.query(torch.from_numpy(qry).to(cuda), nr_nns_searches=6)
    empty = torch.empty((0, 3), dtype=torch.float32, device=cuda)
    d, i = tree.query(empty, nr_nns_searches=2)
    assert d.shape == (0, 2) and i.shape == (0, 2)
    one = torch.from_numpy(qry[:1]).to(cuda)
    d, i = tree.query(one, nr_nns_searches=3)
    od, oi = oracle_for(tree).search(qry[:1], 3)
    assert_same(d.cpu().numpy(), od, "single query")


def test_k_equals_n_and_minimum_points_per_level(cuda):
    rng = np.random.default_rng(10)
    ref = rng.standard_normal((64, 2)).astype(np.float64)
    qry = rng.standard_normal((100, 2)).astype(np.float64)
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels=7)  # 2^(7-1) = 64 points: leaves of 0/1
    for k in (1, 64):
        dist, sidx, _ = gpu_query(tree, qry, k, cuda)
        od, oi = oracle_for(tree).search(qry, k)
        assert_same(dist, od, "distances")
        assert_same(sidx, oi, "indices")


def test_gpu_k_limit(cuda):
    ref = np.random.default_rng(11).standard_normal((1000, 3)).astype(np.float32)
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda)
    with pytest.raises(RuntimeError):
        tree.query(torch.from_numpy(ref[:10]).to(cuda), nr_nns_searches=129)


def test_more_queries_than_one_presort_slice(cuda):
    """Over 2^28 queries (the presort sorts at most 2^28 per slice): results on both sides of the
    slice boundary and at the end equal the oracle's."""
    rng = np.random.default_rng(28)
    ref = rng.standard_normal((5000, 3)).astype(np.float32)
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels="auto")
    n = (1 << 28) + 3000
    gen = torch.Generator(device=cuda)
    gen.manual_seed(5)
    q = torch.randn((n, 3), device=cuda, generator=gen)
    d, i = tree.query(q, nr_nns_searches=1)
    torch.cuda.synchronize()
    picks = np.concatenate([np.arange(0, 2000), np.arange((1 << 28) - 2000, (1 << 28) + 2000),
                            np.arange(n - 2000, n)])
    sel = torch.from_numpy(picks).to(cuda)
    qs = q[sel].cpu().numpy()
    gd = d[sel].cpu().numpy()
    gi = i[sel].cpu().numpy()
    del q, d, i
    od, oi = oracle.OracleTree(ref, tree.levels).search(qs, 1)
    assert_same(gd, od, "distances")
    sh = tree.kdtree.get_shuffled_inds()
    assert_same(gi, sh[oi].astype(np.int64), "original indices")


def test_plain_cpu_workload_exact(cuda):
    """Workload 1 (100k refs, 10k queries, k=5, seed 0) at the reference's default depth."""
    w = workloads.plain_cpu()
    tree = tk.build_kd_tree(torch.from_numpy(w["ref"]), device=cuda)
    assert tree.levels == 13
    dist, sidx, oidx = gpu_query(tree, w["queries"], w["k"], cuda)
    od, oi = oracle.OracleTree(w["ref"], 13).search(w["queries"], w["k"])
    assert_same(dist, od, "distances")
    assert_same(sidx, oi, "structured indices")


def test_plain_gpu_workload_sample_exact(cuda):
    """Workload 2 at 1/16 scale (250k refs, 1M queries, k=16), checked in full against the oracle."""
    w = workloads.plain_gpu(n=250_000, q=1_000_000)
    for levels in (None, "auto"):
        tree = tk.build_kd_tree(torch.from_numpy(w["ref"]), device=cuda, levels=levels)
        dist, sidx, _ = gpu_query(tree, w["queries"], 16, cuda)
        od, oi = oracle_for(tree).search(w["queries"], 16)
        assert_same(dist, od, "distances")
        assert_same(sidx, oi, "structured indices")


def test_repeatable(cuda):
    rng = np.random.default_rng(12)
    ref, qry = _data(rng, 100000, 100000, 3, np.float32)
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels="auto")
    a = gpu_query(tree, qry, 16, cuda)
    b = gpu_query(tree, qry, 16, cuda)
    for x, y in zip(a, b):
        assert_same(x, y, "second run")


def test_saved_tree_loads_onto_the_gpu(cuda, tmp_path):
    rng = np.random.default_rng(4242)
    pts = rng.random((20000, 3)).astype(np.float32)
    tree = tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels="auto")
    path = tmp_path / "gpu_tree.bin"
    tree.save(path)
    loaded = tk.load_kd_tree(path, device=cuda)
    q = torch.from_numpy(rng.random((5000, 3)).astype(np.float32)).to(cuda)
    d0, i0 = tree.query(q, nr_nns_searches=8)
    d1, i1 = loaded.query(q, nr_nns_searches=8)
    assert torch.equal(d0, d1) and torch.equal(i0, i1)
