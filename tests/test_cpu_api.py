"""Host-side behaviour of the product (no GPU needed): tree build parity, CPU search parity,
the reference's API contract (tests/test_knn_unit.py restated for device='cpu') and the C ABI."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import oracle
import torch_kdtree_amd as tk
from torch_kdtree_amd import torch_knn
from tests._util import assert_same, nonfinite_cloud, nonfinite_queries, with_nan_rows

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------------------------------------
# Tree build: identical to the oracle's faithful two-sort restatement of createKDTree
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
@pytest.mark.parametrize("kind", ["normal", "duplicates"])
def test_build_matches_oracle(dtype, d, kind):
    rng = np.random.default_rng(d * 3 + (kind == "duplicates"))
    n = 70000
    if kind == "normal":
        pts = (rng.standard_normal((n, d)) * 1e3).astype(dtype)
    else:
        pts = rng.integers(0, 9, (n, d)).astype(dtype)  # sorts full of equal keys
    for levels in (1, oracle.default_levels(n), 14):
        t = tk.build_kd_tree(torch.from_numpy(pts), device="cpu", levels=levels)
        e = oracle.OracleTree(pts, levels).export()
        axis, split, leaf_off = t.kdtree.get_nodes()
        assert_same(t.kdtree.get_structured_points(), e["structured"], "structured points")
        assert_same(t.kdtree.get_shuffled_inds(), e["shuffled"], "shuffled indices")
        assert_same(axis, e["axis"], "axis")
        assert_same(split, e["median"], "medians")
        assert_same(leaf_off[:-1], e["leaf_off"], "leaf offsets")
        assert leaf_off[-1] == n


def test_build_independent_of_thread_count():
    pts = np.random.default_rng(1).integers(0, 50, (200000, 3)).astype(np.float32)
    from torch_kdtree_amd import torch_knn as m
    a = m.KDTreeCPU3DF(pts, 12)
    old = os.environ.get("OMP_NUM_THREADS")
    b = m.KDTreeCPU3DF(pts, 12)
    assert_same(a.get_shuffled_inds(), b.get_shuffled_inds(), "shuffled")
    assert old == os.environ.get("OMP_NUM_THREADS")


# ---------------------------------------------------------------------------------------------
# CPU search: identical to the oracle's restatement of KDTreeKNNSearch
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
def test_cpu_query_matches_oracle(dtype, d):
    rng = np.random.default_rng(50 + d)
    ref = (rng.standard_normal((30000, d)) * 1e3).astype(dtype)
    qry = (rng.standard_normal((4000, d)) * 1e3).astype(dtype)
    for levels in (None, 13):
        t = tk.build_kd_tree(torch.from_numpy(ref), device="cpu", levels=levels)
        assert t.kdtree.is_finite()
        ot = oracle.tree_from_product(t)
        for k in (1, 7, 16, 150):
            res_i = torch.empty((qry.shape[0], k), dtype=torch.int32)
            dist, inds = t.query(torch.from_numpy(qry), nr_nns_searches=k, result_idx=res_i)
            od, oi = ot.search(qry, k)
            assert_same(dist.numpy(), od, "distances")
            assert_same(res_i.numpy(), oi, "structured indices")
            assert_same(inds.numpy(), t.kdtree.get_shuffled_inds()[oi].astype(np.int64), "original indices")


@pytest.mark.parametrize("levels", [13, 15, 17])
def test_cpu_query_plain_gpu_distribution(levels):
    """Workload-2 distribution (uniform refs, queries partly outside the cube, clustered half):
    long climbs and many skipped far sides."""
    import workloads
    w = workloads.plain_gpu(n=400_000, q=200_000)
    t = tk.build_kd_tree(torch.from_numpy(w["ref"]), device="cpu", levels=levels)
    res_i = torch.empty((200_000, 16), dtype=torch.int32)
    dist, _ = t.query(torch.from_numpy(w["queries"]), nr_nns_searches=16, result_idx=res_i)
    od, oi = oracle.tree_from_product(t).search(w["queries"], 16)
    assert_same(dist.numpy(), od, "distances")
    assert_same(res_i.numpy(), oi, "structured indices")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 3, 8])
def test_cpu_query_nan_rows_match_oracle(dtype, d):
    """Queries with a NaN coordinate: NaN distances on the last records of the one leaf the
    reference scans, then (+inf, 0); the other rows as without them."""
    rng = np.random.default_rng(60 + d)
    ref = (rng.standard_normal((5000, d)) * 1e3).astype(dtype)
    qry = (rng.standard_normal((900, d)) * 1e3).astype(dtype)
    qn, rows = with_nan_rows(qry, rng)
    t = tk.build_kd_tree(torch.from_numpy(ref), device="cpu", levels=7)
    ot = oracle.tree_from_product(t)
    keep = np.setdiff1d(np.arange(len(qry)), rows)
    for k in (1, 5, 16, 60):
        res_i = torch.empty((qn.shape[0], k), dtype=torch.int32)
        dist, inds = t.query(torch.from_numpy(qn), nr_nns_searches=k, result_idx=res_i)
        od, oi = ot.search(qn, k)
        assert_same(dist.numpy(), od, "distances")
        assert_same(res_i.numpy(), oi, "structured indices")
        assert_same(inds.numpy(), t.kdtree.get_shuffled_inds()[oi].astype(np.int64), "original indices")
        assert np.isnan(dist.numpy()[rows, 0]).all()
        d0, _ = t.query(torch.from_numpy(qry), nr_nns_searches=k)
        assert_same(dist.numpy()[keep], d0.numpy()[keep], "rows without NaN")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 3, 8])
@pytest.mark.parametrize("kind", ["one_nan", "nan_rows", "inf_rows", "mixed", "overflow"])
def test_cpu_query_nonfinite_cloud_matches_oracle(dtype, d, kind):
    """Reference clouds with NaN or infinite coordinates (or medians that overflow): the build
    equals the oracle's two-sort build, and the search equals the reference's step for step, NaN
    distances included (taken by `d >= worst` failing, placed by upperBound's probes,
    src/nndistance.hpp:41-85), on queries with NaN and infinite coordinates as well."""
    rng = np.random.default_rng(70 + d + 7 * (dtype == np.float64) + 31 * len(kind))
    ref = nonfinite_cloud(rng, 2000, d, dtype, kind)
    qry = nonfinite_queries(rng, 500, d, dtype)
    for levels in (1, 5, 8):
        t = tk.build_kd_tree(torch.from_numpy(ref), device="cpu", levels=levels)
        ob = oracle.OracleTree(ref, levels).export()
        assert t.kdtree.is_finite() == (np.isfinite(ref).all() and np.isfinite(ob["median"]).all())
        assert_same(t.kdtree.get_structured_points(), ob["structured"], "structured points")
        assert_same(t.kdtree.get_shuffled_inds(), ob["shuffled"], "shuffled indices")
        ot = oracle.tree_from_product(t)
        for k in (1, 4, 5, 16, 40):
            res_i = torch.empty((qry.shape[0], k), dtype=torch.int32)
            dist, inds = t.query(torch.from_numpy(qry), nr_nns_searches=k, result_idx=res_i)
            od, oi = ot.search(qry, k)
            assert_same(dist.numpy(), od, "distances")
            assert_same(res_i.numpy(), oi, "structured indices")
            assert_same(inds.numpy(), t.kdtree.get_shuffled_inds()[oi].astype(np.int64), "original indices")


def test_cpu_query_one_nan_coordinate_2000_points():
    """One NaN coordinate in a 2,000-point cloud, depth 5, k = 4: every row equals the oracle."""
    rng = np.random.default_rng(11)
    ref = rng.random((2000, 3)).astype(np.float32)
    ref[1234, 1] = np.nan
    qry = rng.random((500, 3)).astype(np.float32)
    t = tk.build_kd_tree(torch.from_numpy(ref), device="cpu", levels=5)
    res_i = torch.empty((500, 4), dtype=torch.int32)
    dist, _ = t.query(torch.from_numpy(qry), nr_nns_searches=4, result_idx=res_i)
    od, oi = oracle.tree_from_product(t).search(qry, 4)
    assert_same(dist.numpy(), od, "distances")
    assert_same(res_i.numpy(), oi, "structured indices")
    assert np.isnan(od).any(), "the NaN point must reach some lists"


def test_cpu_query_ties_match_oracle():
    rng = np.random.default_rng(3)
    ref = rng.integers(0, 6, (20000, 3)).astype(np.float32)
    qry = (rng.integers(0, 6, (3000, 3)) + 0.5 * rng.integers(0, 2, (3000, 3))).astype(np.float32)
    t = tk.build_kd_tree(torch.from_numpy(ref), device="cpu", levels=9)
    od, oi = oracle.OracleTree(ref, 9).search(qry, 24)
    res_i = torch.empty((qry.shape[0], 24), dtype=torch.int32)
    dist, _ = t.query(torch.from_numpy(qry), nr_nns_searches=24, result_idx=res_i)
    assert_same(dist.numpy(), od, "distances")
    assert_same(res_i.numpy(), oi, "structured indices")


def test_known_answers_cpu():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "known_answers.json")) as f:
        cases = json.load(f)
    for case in cases:
        pts = torch.tensor(case["points"], dtype=torch.float32)
        t = tk.build_kd_tree(pts, device="cpu", levels=case["levels"])
        for q in case["queries"]:
            res_i = torch.empty((1, q["k"]), dtype=torch.int32)
            d, i = t.query(torch.tensor([q["q"]], dtype=torch.float32), nr_nns_searches=q["k"], result_idx=res_i)
            assert d[0].tolist() == q["dist"] and res_i[0].tolist() == q["idx"] and i[0].tolist() == q["orig"]


# ---------------------------------------------------------------------------------------------
# The reference's own test grid (tests/test_knn_unit.py:30-70), device='cpu'
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("nr_refs", [5, 10, 20, 30, 1000, 10000])
@pytest.mark.parametrize("nr_query", [5, 30, 1000])
@pytest.mark.parametrize("k", [1, 5, 10, 20])
@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("requires_grad", [True, False])
def test_reference_grid_cpu(nr_refs, nr_query, k, d, dtype, requires_grad):
    from scipy.spatial import cKDTree
    g = torch.Generator().manual_seed(nr_refs * 31 + nr_query * 7 + k + d)
    points_ref = torch.randn(size=(nr_refs, d), dtype=dtype, generator=g) * 1e3
    points_query = (torch.randn(size=(nr_query, d), dtype=dtype, generator=g) * 1e3).requires_grad_(requires_grad)
    t = tk.build_kd_tree(points_ref, device=torch.device("cpu"))
    if k > nr_refs:
        with pytest.raises(Exception):
            t.query(points_query, nr_nns_searches=k)
        return
    dists, inds = t.query(points_query, nr_nns_searches=k)
    assert inds.shape == (nr_query, k) and dists.shape == (nr_query, k) and inds.dtype == torch.int64
    dref, iref = cKDTree(points_ref.numpy()).query(points_query.detach().numpy(), k)
    dref = np.asarray(dref).reshape(nr_query, k)
    iref = np.asarray(iref).reshape(nr_query, k)
    assert np.allclose(dref ** 2, dists.detach().numpy(), atol=1e-5, rtol=1e-5 if dtype == torch.float64 else 1e-4)
    assert np.sum(iref == inds.numpy()) / iref.size > 0.95
    od, oi = oracle.tree_from_product(t).search(points_query.detach().numpy(), k)
    assert np.array_equal(dists.detach().numpy(), od)


def test_uncompiled_dimension():
    points_ref = np.random.default_rng(0).uniform(size=(10, 436))
    with pytest.raises(Exception):
        tk.build_kd_tree(points_ref, device="cpu")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_wrong_dtype(dtype):
    other = torch.float64 if dtype == torch.float32 else torch.float32
    t = tk.build_kd_tree(torch.randn(size=(10, 3), dtype=dtype), device="cpu")
    with pytest.raises(Exception):
        t.query(torch.randn(size=(10, 3), dtype=other))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_grad_cpu(dtype):
    """tests/test_knn_unit.py:90-107 on a CPU tree (the reference's torch recomputation path)."""
    torch.manual_seed(0)
    points_ref = torch.randn(size=(10, 3), dtype=dtype, requires_grad=True)
    t = tk.build_kd_tree(points_ref)
    points_query = torch.randn(size=(5, 3), dtype=dtype, requires_grad=True)
    dists, inds = t.query(points_query, nr_nns_searches=2)
    (0.5 * torch.sum(dists)).backward()
    assert torch.allclose(points_query.grad, torch.sum((points_query[:, None] - points_ref[inds]), axis=-2))
    ref_grad = torch.zeros_like(points_ref)
    for i, ind in enumerate(inds):
        ref_grad[ind] += points_ref[ind] - points_query[i]
    assert torch.allclose(points_ref.grad, ref_grad)


def test_levels_and_point_count_errors():
    pts = torch.randn(100, 3)
    with pytest.raises(AssertionError):
        tk.build_kd_tree(pts, device="cpu", levels=0)
    with pytest.raises(AssertionError):
        tk.build_kd_tree(pts, device="cpu", levels=25)
    with pytest.raises(RuntimeError, match="Ran out of points"):
        tk.build_kd_tree(pts, device="cpu", levels=8)  # needs 128 points
    tk.build_kd_tree(torch.randn(128, 3), device="cpu", levels=8)
    with pytest.raises(ValueError):
        tk.build_kd_tree(pts, device="cpu", tie_break="random")
    with pytest.raises(RuntimeError, match="needs a GPU"):
        tk.build_kd_tree(pts, device="cpu", builder="gpu")
    with pytest.raises(ValueError, match="builder"):
        tk.build_kd_tree(pts, device="cpu", builder="fast")


def test_unsupported_dtype_raises():
    with pytest.raises(RuntimeError):
        tk.build_kd_tree(torch.randn(100, 3).half(), device="cpu")


def test_noncontiguous_result_buffers_are_written():
    ref = torch.randn(500, 3)
    qry = torch.randn(40, 3)
    t = tk.build_kd_tree(ref, device="cpu")
    buf_d = torch.zeros(4, 40).t()
    buf_i = torch.zeros(4, 40, dtype=torch.int32).t()
    assert not buf_d.is_contiguous()
    d, _ = t.query(qry, nr_nns_searches=4, result_dists=buf_d, result_idx=buf_i)
    assert torch.equal(buf_d, d)
    ref_i = torch.empty(40, 4, dtype=torch.int32)
    t.query(qry, nr_nns_searches=4, result_idx=ref_i)
    assert torch.equal(buf_i, ref_i)


def test_squared_distances_false():
    ref = torch.randn(500, 3, dtype=torch.float64)
    qry = torch.randn(40, 3, dtype=torch.float64)
    d2, i2 = tk.build_kd_tree(ref, device="cpu").query(qry, nr_nns_searches=3)
    d1, i1 = tk.build_kd_tree(ref, device="cpu", squared_distances=False).query(qry, nr_nns_searches=3)
    assert torch.equal(i1, i2) and torch.equal(d1, torch.sqrt(d2))


def test_numpy_input_and_module_surface():
    assert tk.__version__ == "1.0" and torch_knn.__version__ == "1.0"
    assert tk.gpu_available is True
    for dev in ("CPU", "GPU"):
        for d in (1, 2, 3, 8):
            for p in ("", "F"):
                assert hasattr(torch_knn, "KDTree%s%dD%s" % (dev, d, p))
    t = tk.build_kd_tree(np.random.default_rng(0).standard_normal((300, 2)), device="cpu")
    assert t.dtype == torch.float64 and t.dims == 2 and t.nr_ref_points == 300
    assert t.shuffled_ind.dtype == torch.int64 and t.structured_points.shape == (300, 2)


# ---------------------------------------------------------------------------------------------
# C ABI
# ---------------------------------------------------------------------------------------------
def _header_symbols():
    with open(os.path.join(ROOT, "include", "torch_kdtree_amd.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(tk_\w+)\s*\(", text, re.M)))


def test_c_abi_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "torch_kdtree_amd", "libtorch_kdtree_amd.so"))
    syms = _header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s


def test_c_abi_build_query_cpu_and_errors():
    lib = ctypes.CDLL(os.path.join(ROOT, "torch_kdtree_amd", "libtorch_kdtree_amd.so"))
    lib.tk_last_error.restype = ctypes.c_char_p
    lib.tk_version.restype = ctypes.c_char_p
    assert lib.tk_version() == b"1.0"
    pts = np.random.default_rng(0).standard_normal((1000, 3)).astype(np.float32)
    h = ctypes.c_void_p()
    rc = lib.tk_tree_build(pts.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(1000), 3, 0, 7, 0, ctypes.byref(h))
    assert rc == 0
    q = pts[:10].copy()
    dist = np.empty((10, 4), np.float32)
    idx = np.empty((10, 4), np.int32)
    orig = np.empty((10, 4), np.int64)
    rc = lib.tk_knn_query_cpu(h, q.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(10), 4,
                              dist.ctypes.data_as(ctypes.c_void_p), idx.ctypes.data_as(ctypes.c_void_p),
                              orig.ctypes.data_as(ctypes.c_void_p), 0)
    assert rc == 0 and np.array_equal(orig[:, 0], np.arange(10)) and np.all(dist[:, 0] == 0)
    rc = lib.tk_knn_query_cpu(h, q.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(10), 1001,
                              dist.ctypes.data_as(ctypes.c_void_p), idx.ctypes.data_as(ctypes.c_void_p), None, 0)
    assert rc != 0 and b"more nearest neighbors" in lib.tk_last_error()
    # a CPU-only tree refuses GPU queries with an error, not a crash
    rc = lib.tk_knn_query_gpu(h, None, ctypes.c_int64(10), 4, None, None, None, 0, -1, None, ctypes.c_size_t(0), None)
    assert rc != 0 and b"no GPU copy" in lib.tk_last_error()
    lib.tk_tree_free(h)
    bad = ctypes.c_void_p()
    assert lib.tk_tree_build(pts.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(1000), 5, 0, 7, 0,
                             ctypes.byref(bad)) != 0
    assert b"dimensionality" in lib.tk_last_error()


def test_c_abi_gpu_build_validates_before_touching_a_device():
    """tk_tree_build_gpu / tk_tree_build_many_gpu reject bad arguments with the host build's
    messages before any device call (so these run without a GPU)."""
    lib = ctypes.CDLL(os.path.join(ROOT, "torch_kdtree_amd", "libtorch_kdtree_amd.so"))
    lib.tk_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    out = vp()
    cases = [((10, 5, 0, 2), b"dimensionality"), ((10, 3, 7, 2), b"dtype"), ((10, 3, 0, 0), b"levels"),
             ((10, 3, 0, 25), b"levels"), ((10, 3, 0, 6), b"Ran out of points")]
    for (n, d, dt, lv), msg in cases:
        rc = lib.tk_tree_build_gpu(vp(1), ctypes.c_int64(n), d, dt, lv, -1, None, ctypes.byref(out))
        assert rc != 0 and msg in lib.tk_last_error(), (n, d, dt, lv, lib.tk_last_error())
        assert not out.value
    outs = (vp * 2)()
    ns = (ctypes.c_int64 * 2)(100, 3)
    lvs = (ctypes.c_int32 * 2)(3, 3)
    rc = lib.tk_tree_build_many_gpu(vp(1), ns, 2, 3, 0, lvs, -1, None, outs)
    assert rc != 0 and b"Ran out of points" in lib.tk_last_error()
    assert not outs[0] and not outs[1]
    assert lib.tk_tree_build_many_gpu(vp(1), ns, 0, 3, 0, lvs, -1, None, outs) != 0


# --- tree serialization --------------------------------------------------------------------------

@pytest.mark.parametrize("dtype,d", [(np.float32, 3), (np.float64, 2), (np.float32, 8), (np.float64, 1)])
def test_tree_save_load_roundtrip(tmp_path, dtype, d):
    rng = np.random.default_rng(900 + d)
    pts = np.round(rng.standard_normal((3000, d)) * 4).astype(dtype)  # duplicates included
    tree = tk.build_kd_tree(torch.from_numpy(pts), device="cpu", levels=7)
    path = tmp_path / "tree.bin"
    tree.save(path)
    loaded = tk.load_kd_tree(path, device="cpu")
    assert loaded.levels == tree.levels and loaded.dims == d and loaded.dtype == tree.dtype
    np.testing.assert_array_equal(loaded.kdtree.get_structured_points(), tree.kdtree.get_structured_points())
    np.testing.assert_array_equal(loaded.kdtree.get_shuffled_inds(), tree.kdtree.get_shuffled_inds())
    for a, b in zip(loaded.kdtree.get_nodes(), tree.kdtree.get_nodes()):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(loaded.points_ref_bak.numpy(), pts)
    q = torch.from_numpy(rng.standard_normal((500, d)).astype(dtype) * 4)
    d0, i0 = tree.query(q, nr_nns_searches=5)
    d1, i1 = loaded.query(q, nr_nns_searches=5)
    assert torch.equal(d0, d1) and torch.equal(i0, i1)


def test_tree_load_rejects_damaged_files(tmp_path):
    pts = np.random.default_rng(3).random((1000, 3)).astype(np.float32)
    path = tmp_path / "tree.bin"
    tk.build_kd_tree(torch.from_numpy(pts), device="cpu", levels=5).save(path)
    good = path.read_bytes()
    for name, data in [("flipped", good[:100] + bytes([good[100] ^ 1]) + good[101:]),
                       ("truncated", good[:-9]),
                       ("empty", b""),
                       ("magic", b"XKDTREE\0" + good[8:])]:
        bad = tmp_path / ("bad_" + name)
        bad.write_bytes(data)
        with pytest.raises(RuntimeError):
            tk.load_kd_tree(bad, device="cpu")
    with pytest.raises(RuntimeError):
        tk.load_kd_tree(tmp_path / "missing.bin", device="cpu")


def test_c_abi_serialize_roundtrip():
    lib = ctypes.CDLL(os.path.join(ROOT, "torch_kdtree_amd", "libtorch_kdtree_amd.so"))
    vp = ctypes.c_void_p
    lib.tk_last_error.restype = ctypes.c_char_p
    lib.tk_tree_build.argtypes = [vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(vp)]
    lib.tk_tree_serialize.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_size_t)]
    lib.tk_tree_deserialize.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(vp)]
    lib.tk_tree_get_shuffled_inds.argtypes = [vp, vp]
    lib.tk_tree_free.argtypes = [vp]
    pts = np.random.default_rng(5).random((777, 2)).astype(np.float64)
    h = vp()
    assert lib.tk_tree_build(pts.ctypes.data, 777, 2, 1, 6, 0, ctypes.byref(h)) == 0, lib.tk_last_error()
    n = ctypes.c_size_t(0)
    assert lib.tk_tree_serialize(h, None, ctypes.byref(n)) == 0
    buf = ctypes.create_string_buffer(n.value)
    small = ctypes.c_size_t(n.value - 1)
    assert lib.tk_tree_serialize(h, buf, ctypes.byref(small)) != 0  # too small a buffer is an error
    assert lib.tk_tree_serialize(h, buf, ctypes.byref(n)) == 0
    h2 = vp()
    assert lib.tk_tree_deserialize(buf, n.value, ctypes.byref(h2)) == 0, lib.tk_last_error()
    a = np.empty(777, np.int32)
    b = np.empty(777, np.int32)
    lib.tk_tree_get_shuffled_inds(h, a.ctypes.data)
    lib.tk_tree_get_shuffled_inds(h2, b.ctypes.data)
    np.testing.assert_array_equal(a, b)
    lib.tk_tree_free(h)
    lib.tk_tree_free(h2)


def test_autotune_cache_lookup(tmp_path, monkeypatch):
    """The tuning cache: leaf size per dims/dtype, launch variant per list width; missing entries
    fall back to the built-in table and variant 0; CPU trees keep the reference's default depth."""
    import json
    from torch_kdtree_amd import autotune
    cache = tmp_path / "t.json"
    monkeypatch.setenv("TORCH_KDTREE_AMD_TUNING", str(cache))
    assert autotune.variant_for(3, torch.float32, 16) == 0
    assert autotune.target_leaf_size(3, torch.float32) == autotune.DEFAULT_LEAF[3]
    cache.write_text(json.dumps({"d3_float32": {"leaf_size": 24, "variants": {"16": 5, "8": 2}}}))
    assert autotune.target_leaf_size(3, torch.float32) == 24
    assert autotune.variant_for(3, torch.float32, 9) == 5 and autotune.variant_for(3, torch.float32, 16) == 5
    assert autotune.variant_for(3, torch.float32, 5) == 2 and autotune.variant_for(3, torch.float32, 4) == 0
    assert autotune.variant_for(3, torch.float64, 16) == 0
    # launch variants are kept per tie policy: the reference policy's do not apply to the other
    assert autotune.variant_for(3, torch.float32, 16, "lower_index") == 0
    cache.write_text(json.dumps({"d3_float32": {"leaf_size": 24, "variants": {"16": 5},
                                                "variants_lower_index": {"16": 1}}}))
    assert autotune.variant_for(3, torch.float32, 16, "lower_index") == 1
    assert autotune.variant_for(3, torch.float32, 16) == 5
    assert autotune.pick_levels(4_000_000, 3, "cpu") == 13
    assert autotune.pick_levels(4_000_000, 3, "cuda") == autotune._levels_for_leaf(4_000_000, 24)
    from torch_kdtree_amd import torch_knn
    table = torch_knn.tree_variants()
    assert table[0] == {"chunk": 0, "waves": 0, "block": 0} and len(table) == 9
    assert all(v["chunk"] in (4, 8, 16) and v["block"] in (64, 128, 256) for v in table[1:])


def test_debug_flavour_library_runs_the_reference_cross_checks():
    """libtorch_kdtree_amd_debug.so (make FLAVOUR=debug, built by __graft_entry__.build()): the same
    C ABI with the reference's debug cross-checks compiled in; a CPU build and search on a tie-heavy
    cloud pass them and equal the release library's results."""
    lib_path = os.path.join(ROOT, "torch_kdtree_amd", "libtorch_kdtree_amd_debug.so")
    if not os.path.exists(lib_path):
        pytest.skip("debug flavour not built (make -C torch_kdtree_amd/csrc FLAVOUR=debug)")
    dbg = ctypes.CDLL(lib_path)
    rel = ctypes.CDLL(os.path.join(ROOT, "torch_kdtree_amd", "libtorch_kdtree_amd.so"))
    rng = np.random.default_rng(4)
    pts = rng.integers(0, 7, (20000, 3)).astype(np.float32)
    qry = (rng.integers(0, 7, (3000, 3)) + 0.5 * rng.integers(0, 2, (3000, 3))).astype(np.float32)
    outs = []
    for lib in (dbg, rel):
        lib.tk_last_error.restype = ctypes.c_char_p
        tree = ctypes.c_void_p()
        assert lib.tk_tree_build(pts.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(pts)), 3, 0, 9, 0,
                                 ctypes.byref(tree)) == 0, lib.tk_last_error()
        d = np.empty((3000, 12), np.float32)
        i = np.empty((3000, 12), np.int32)
        rc = lib.tk_knn_query_cpu(tree, qry.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(3000), 12,
                                  d.ctypes.data_as(ctypes.c_void_p), i.ctypes.data_as(ctypes.c_void_p), None, 0)
        assert rc == 0, lib.tk_last_error()
        lib.tk_tree_free(tree)
        outs.append((d, i))
    assert_same(outs[0][0], outs[1][0], "debug vs release distances")
    assert_same(outs[0][1], outs[1][1], "debug vs release indices")


def test_search_stats_counts_scanned_points():
    """tk_knn_search_stats: with k equal to the cloud size nothing can be pruned, so every query scans
    both leaves of a one-level tree; with small k on a deep tree it scans few points, at least k."""
    rng = np.random.default_rng(2)
    ref = rng.random((500, 3)).astype(np.float32)
    qry = rng.random((40, 3)).astype(np.float32)
    t = tk.build_kd_tree(torch.from_numpy(ref), device="cpu", levels=1)
    assert t.kdtree.search_stats(qry, 500) == (80, 40 * 500)
    big = rng.random((100000, 3)).astype(np.float32)
    t2 = tk.build_kd_tree(torch.from_numpy(big), device="cpu", levels=12)
    leaves, points = t2.kdtree.search_stats(qry, 8)
    assert 40 <= leaves <= 40 * 64 and 40 * 8 <= points <= 40 * 3000
