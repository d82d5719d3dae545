"""GPU parity of the batched many-cloud mode.

Every cloud's batched results must equal the oracle's on that cloud's tree (reference tie policy)
or brute force (lower-index policy), and equal the single-tree GPU path bit for bit: distances,
structured and original indices, dL/dq and dL/dr.  Graph replays must equal eager runs.
"""
import numpy as np
import pytest
import torch

import oracle
import torch_kdtree_amd as tk
from torch_kdtree_amd import torch_knn
from tests._util import assert_same, with_nan_rows

pytestmark = pytest.mark.gpu


def _clouds(rng, n_clouds, d, dtype, sizes=(1, 5000)):
    clouds, queries, ks = [], [], []
    for c in range(n_clouds):
        n = int(rng.integers(sizes[0], sizes[1] + 1))
        pts = rng.standard_normal((n, d)) * (1.0 + c)
        if c % 3 == 1:  # integer grid: many exact ties
            pts = np.round(pts * 2.0)
        nq = int(rng.integers(0, 2 * n + 2)) if c % 4 != 2 else 0
        qry = rng.standard_normal((nq, d)) * (1.0 + c)
        if c % 3 == 1:
            qry = np.round(qry * 2.0)
        clouds.append(pts.astype(dtype))
        queries.append(qry.astype(dtype))
        ks.append(int(rng.integers(1, min(32, n) + 1)))
    return clouds, queries, ks


def _plan_forward(plan, q_cat, total_out, dtype, dev):
    d = torch.empty(total_out, dtype=dtype, device=dev)
    i = torch.empty(total_out, dtype=torch.int32, device=dev)
    o = torch.empty(total_out, dtype=torch.int64, device=dev)
    plan.forward(q_cat.data_ptr(), d.data_ptr(), i.data_ptr(), o.data_ptr(), torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    return d, i, o


@pytest.mark.parametrize("tie", ["reference", "lower_index"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
def test_batch_forward_matches_oracle_and_single_tree(cuda, tie, dtype, d):
    rng = np.random.default_rng(500 + 10 * d + (dtype == np.float64) + 2 * (tie == "lower_index"))
    clouds, queries, ks = _clouds(rng, 9, d, dtype)
    batch = tk.build_kd_tree_batch([torch.from_numpy(c) for c in clouds], device=cuda, levels="auto", tie_break=tie)
    plan = batch.plan([q.shape[0] for q in queries], ks)
    total_q, total_out, total_ref, n_clouds = plan.sizes()
    assert n_clouds == len(clouds) and total_ref == sum(c.shape[0] for c in clouds)
    q_cat = torch.from_numpy(np.concatenate(queries)).to(cuda)
    dd, ii, oo = _plan_forward(plan, q_cat, total_out, q_cat.dtype, cuda)
    q_off, out_off, _ = plan.offsets()
    for c, (pts, qry, k) in enumerate(zip(clouds, queries, ks)):
        bd = dd[out_off[c]:out_off[c + 1]].view(-1, k).cpu().numpy()
        bi = ii[out_off[c]:out_off[c + 1]].view(-1, k).cpu().numpy()
        bo = oo[out_off[c]:out_off[c + 1]].view(-1, k).cpu().numpy()
        assert bd.shape[0] == qry.shape[0]
        if qry.shape[0] == 0:
            continue
        single = tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels=batch.levels[c], tie_break=tie)
        q = torch.from_numpy(qry).to(cuda)
        sd = torch.empty((q.shape[0], k), dtype=q.dtype, device=cuda)
        si = torch.empty((q.shape[0], k), dtype=torch.int32, device=cuda)
        _, so = single.query(q, nr_nns_searches=k, result_dists=sd, result_idx=si)
        torch.cuda.synchronize()
        assert_same(bd, sd.cpu().numpy(), "cloud %d distances vs single tree" % c)
        assert_same(bi, si.cpu().numpy(), "cloud %d structured indices vs single tree" % c)
        assert_same(bo, so.cpu().numpy(), "cloud %d original indices vs single tree" % c)
        if tie == "reference":
            ot = oracle.OracleTree(pts, batch.levels[c])
            od, oi = ot.search(qry, k)
            assert_same(bd, od, "cloud %d distances vs oracle" % c)
            assert_same(bi, oi, "cloud %d structured indices vs oracle" % c)
        else:
            od, oi = oracle.bruteforce(pts, qry, k)
            assert_same(bd, od, "cloud %d distances vs brute force" % c)
            assert_same(bo, oi.astype(np.int64), "cloud %d indices vs brute force" % c)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n_clouds,sizes", [(7, (20, 3000)), (40, (1800, 3000))])
def test_batch_backward_matches_single_tree(cuda, dtype, n_clouds, sizes):
    """With 40 clouds of at most 3,000 points (over 65,536 points in all) the pair sort keys carry
    the cloud in their high bits and sort by the record within the cloud (two passes instead of
    three); with 7 small clouds they are arena records."""
    rng = np.random.default_rng(77 + (dtype == np.float64) + n_clouds)
    clouds, queries, ks = _clouds(rng, n_clouds, 3, dtype, sizes=sizes)
    batch = tk.build_kd_tree_batch(clouds, device=cuda)
    plan = batch.plan([q.shape[0] for q in queries], ks)
    total_q, total_out, total_ref, _ = plan.sizes()
    q_cat = torch.from_numpy(np.concatenate(queries)).to(cuda)
    dd, ii, oo = _plan_forward(plan, q_cat, total_out, q_cat.dtype, cuda)
    g = torch.from_numpy(rng.random(total_out).astype(dtype)).to(cuda)
    dq = torch.empty_like(q_cat)
    dr = torch.empty((total_ref, 3), dtype=q_cat.dtype, device=cuda)
    stream = torch.cuda.current_stream(cuda).cuda_stream
    plan.backward(q_cat.data_ptr(), ii.data_ptr(), g.data_ptr(), dq.data_ptr(), dr.data_ptr(), stream)
    dq2 = torch.empty_like(q_cat)
    dr2 = torch.empty_like(dr)
    plan.backward(q_cat.data_ptr(), ii.data_ptr(), g.data_ptr(), dq2.data_ptr(), dr2.data_ptr(), stream)
    torch.cuda.synchronize()
    assert torch.equal(dq, dq2) and torch.equal(dr, dr2), "batched backward is not repeatable"
    q_off, out_off, r_off = plan.offsets()
    dtype_name = "float32" if dtype == np.float32 else "float64"
    for c, (pts, qry, k) in enumerate(zip(clouds, queries, ks)):
        r = torch.from_numpy(pts).to(cuda)
        q = torch.from_numpy(qry).to(cuda)
        idx = oo[out_off[c]:out_off[c + 1]].contiguous()
        gc = g[out_off[c]:out_off[c + 1]].contiguous()
        sq = torch.empty_like(q)
        sr = torch.empty_like(r)
        if q.shape[0] > 0:
            torch_knn.knn_backward(q.data_ptr(), r.data_ptr(), q.shape[0], r.shape[0], 3, dtype_name, k, idx.data_ptr(),
                                   gc.data_ptr(), sq.data_ptr(), sr.data_ptr(), stream)
        else:
            sr.zero_()
        torch.cuda.synchronize()
        assert_same(dq[q_off[c]:q_off[c + 1]].cpu().numpy(), sq.cpu().numpy(), "cloud %d dL/dq" % c)
        assert_same(dr[r_off[c]:r_off[c + 1]].cpu().numpy(), sr.cpu().numpy(), "cloud %d dL/dr" % c)


@pytest.mark.parametrize("tie", ["reference", "lower_index"])
def test_batch_nan_query_rows(cuda, tie):
    """Clouds whose query sets hold rows with NaN coordinates: forward equal to the single-tree
    path (whose NaN rows follow the oracle) and to the oracle; backward (reference policy) equal to
    the single-tree backward, NaN in the same places."""
    rng = np.random.default_rng(611 + (tie == "lower_index"))
    clouds, queries, ks = _clouds(rng, 8, 3, np.float32, sizes=(40, 4000))
    queries = [with_nan_rows(q, rng)[0] if q.shape[0] else q for q in queries]
    batch = tk.build_kd_tree_batch(clouds, device=cuda, tie_break=tie)
    plan = batch.plan([q.shape[0] for q in queries], ks)
    total_q, total_out, total_ref, _ = plan.sizes()
    q_cat = torch.from_numpy(np.concatenate(queries)).to(cuda)
    dd, ii, oo = _plan_forward(plan, q_cat, total_out, q_cat.dtype, cuda)
    q_off, out_off, r_off = plan.offsets()
    stream = torch.cuda.current_stream(cuda).cuda_stream
    if tie == "reference":
        g = torch.from_numpy(rng.random(total_out).astype(np.float32)).to(cuda)
        dq = torch.empty_like(q_cat)
        dr = torch.empty((total_ref, 3), dtype=q_cat.dtype, device=cuda)
        plan.backward(q_cat.data_ptr(), ii.data_ptr(), g.data_ptr(), dq.data_ptr(), dr.data_ptr(), stream)
        torch.cuda.synchronize()
    for c, (pts, qry, k) in enumerate(zip(clouds, queries, ks)):
        if qry.shape[0] == 0:
            continue
        bd = dd[out_off[c]:out_off[c + 1]].view(-1, k).cpu().numpy()
        bi = ii[out_off[c]:out_off[c + 1]].view(-1, k).cpu().numpy()
        bo = oo[out_off[c]:out_off[c + 1]].view(-1, k).cpu().numpy()
        single = tk.build_kd_tree(torch.from_numpy(pts), device=cuda, levels=batch.levels[c], tie_break=tie)
        q = torch.from_numpy(qry).to(cuda)
        sd = torch.empty((q.shape[0], k), dtype=q.dtype, device=cuda)
        si = torch.empty((q.shape[0], k), dtype=torch.int32, device=cuda)
        _, so = single.query(q, nr_nns_searches=k, result_dists=sd, result_idx=si)
        torch.cuda.synchronize()
        assert_same(bd, sd.cpu().numpy(), "cloud %d distances vs single tree" % c)
        assert_same(bi, si.cpu().numpy(), "cloud %d structured indices vs single tree" % c)
        assert_same(bo, so.cpu().numpy(), "cloud %d original indices vs single tree" % c)
        if tie == "reference":
            od, oi = oracle.OracleTree(pts, batch.levels[c]).search(qry, k)
            assert_same(bd, od, "cloud %d distances vs oracle" % c)
            assert_same(bi, oi, "cloud %d structured indices vs oracle" % c)
            r = torch.from_numpy(pts).to(cuda)
            idx = oo[out_off[c]:out_off[c + 1]].contiguous()
            gc = g[out_off[c]:out_off[c + 1]].contiguous()
            sq = torch.empty_like(q)
            sr = torch.empty_like(r)
            torch_knn.knn_backward(q.data_ptr(), r.data_ptr(), q.shape[0], r.shape[0], 3, "float32", k,
                                   idx.data_ptr(), gc.data_ptr(), sq.data_ptr(), sr.data_ptr(), stream)
            torch.cuda.synchronize()
            assert_same(dq[q_off[c]:q_off[c + 1]].cpu().numpy(), sq.cpu().numpy(), "cloud %d dL/dq" % c)
            assert_same(dr[r_off[c]:r_off[c + 1]].cpu().numpy(), sr.cpu().numpy(), "cloud %d dL/dr" % c)
        else:
            od, oi = oracle.bruteforce(pts, qry, k)
            assert_same(bd, od, "cloud %d distances vs brute force" % c)
            assert_same(bo, oi.astype(np.int64), "cloud %d indices vs brute force" % c)


def test_batch_graph_replays_equal_eager(cuda):
    rng = np.random.default_rng(91)
    clouds, queries, ks = _clouds(rng, 12, 3, np.float32, sizes=(50, 20000))
    batch = tk.build_kd_tree_batch(clouds, device=cuda)
    qt = [torch.from_numpy(q) for q in queries]
    weights = [torch.from_numpy(rng.random((q.shape[0], k)).astype(np.float32)) for q, k in zip(queries, ks)]
    step = tk.BatchStep(batch, qt, ks, weights)
    # eager reference with the same plan
    plan = step.plan
    total_q, total_out, total_ref, _ = plan.sizes()
    ed, ei, eo = _plan_forward(plan, step.queries, total_out, torch.float32, cuda)
    edq = torch.empty_like(step.queries)
    edr = torch.empty((total_ref, 3), dtype=torch.float32, device=cuda)
    stream = torch.cuda.current_stream(cuda).cuda_stream
    plan.backward(step.queries.data_ptr(), ei.data_ptr(), step.weights.data_ptr(), edq.data_ptr(), edr.data_ptr(),
                  stream)
    torch.cuda.synchronize()
    other = tk.build_kd_tree(torch.from_numpy(clouds[0]), device=cuda)
    for rep in range(3):
        for buf in (step.dist, step.grad_queries, step.grad_refs):
            buf.fill_(-1.0)
        step.idx.fill_(-1)
        step.replay()
        torch.cuda.synchronize()
        assert torch.equal(step.dist, ed), "replay %d distances" % rep
        assert torch.equal(step.idx, ei), "replay %d indices" % rep
        assert torch.equal(step.orig, eo), "replay %d original indices" % rep
        assert torch.equal(step.grad_queries, edq), "replay %d dL/dq" % rep
        assert torch.equal(step.grad_refs, edr), "replay %d dL/dr" % rep
        # eager work between replays
        other.query(torch.from_numpy(queries[1]).to(cuda) if queries[1].shape[0] else
                    torch.zeros((1, 3), device=cuda), nr_nns_searches=1)
        torch.cuda.synchronize()


def test_batch_autograd_matches_per_cloud_autograd(cuda):
    rng = np.random.default_rng(13)
    clouds, queries, ks = _clouds(rng, 5, 3, np.float32, sizes=(30, 4000))
    refs = [torch.from_numpy(c).to(cuda).requires_grad_(True) for c in clouds]
    qs = [torch.from_numpy(q).to(cuda).requires_grad_(True) for q in queries]
    ws = [torch.from_numpy(rng.random((q.shape[0], k)).astype(np.float32)).to(cuda) for q, k in zip(queries, ks)]
    batch = tk.build_kd_tree_batch(refs, device=cuda)
    dists, inds = batch.query(qs, ks)
    loss = sum((w * d).sum() for w, d in zip(ws, dists))
    loss.backward()
    for c in range(len(clouds)):
        r2 = torch.from_numpy(clouds[c]).to(cuda).requires_grad_(True)
        q2 = torch.from_numpy(queries[c]).to(cuda).requires_grad_(True)
        tree = tk.build_kd_tree(r2, device=cuda, levels=batch.levels[c])
        d2, i2 = tree.query(q2, nr_nns_searches=ks[c])
        assert torch.equal(d2.detach(), dists[c].detach()) and torch.equal(i2, inds[c])
        if q2.shape[0]:
            (ws[c] * d2).sum().backward()
            assert torch.equal(q2.grad, qs[c].grad), "cloud %d dL/dq" % c
            assert torch.equal(r2.grad, refs[c].grad), "cloud %d dL/dr" % c


def test_batch_errors(cuda):
    batch = tk.build_kd_tree_batch([np.zeros((10, 3), np.float32), np.ones((5, 3), np.float32)], device=cuda)
    with pytest.raises(RuntimeError):
        batch.plan([3, 3], [1, 33])
    with pytest.raises(RuntimeError):
        batch.query([torch.zeros((2, 3)), torch.zeros((2, 3))], [1, 6])


def test_two_steps_of_equal_shapes_keep_their_own_graphs(cuda):
    """Two BatchSteps with the same query counts and k but different queries and weights: each
    replay writes its own step's buffers with its own inputs, equal to an eager run of those."""
    rng = np.random.default_rng(321)
    clouds, queries, ks = _clouds(rng, 6, 3, np.float32, sizes=(100, 8000))
    batch = tk.build_kd_tree_batch(clouds, device=cuda)
    qa = [torch.from_numpy(q) for q in queries]
    qb = [torch.from_numpy((q + rng.normal(0, 0.01, q.shape)).astype(np.float32)) for q in queries]
    wa = [torch.from_numpy(rng.random((q.shape[0], k)).astype(np.float32)) for q, k in zip(queries, ks)]
    wb = [torch.from_numpy(rng.random((q.shape[0], k)).astype(np.float32)) for q, k in zip(queries, ks)]
    step_a = tk.BatchStep(batch, qa, ks, wa)
    step_b = tk.BatchStep(batch, qb, ks, wb)
    assert step_a.plan is not step_b.plan
    stream = torch.cuda.current_stream(cuda).cuda_stream
    want = {}
    for name, step in (("a", step_a), ("b", step_b)):
        plan = batch.new_plan(step.nq, step.ks)
        total_q, total_out, total_ref, _ = plan.sizes()
        ed, ei, eo = _plan_forward(plan, step.queries, total_out, torch.float32, cuda)
        edq = torch.empty_like(step.queries)
        edr = torch.empty((total_ref, 3), dtype=torch.float32, device=cuda)
        plan.backward(step.queries.data_ptr(), ei.data_ptr(), step.weights.data_ptr(), edq.data_ptr(),
                      edr.data_ptr(), stream)
        torch.cuda.synchronize()
        want[name] = (ed, ei, eo, edq, edr)
    assert not torch.equal(want["a"][0], want["b"][0])
    for rep in range(2):
        for name, step in (("a", step_a), ("b", step_b), ("a", step_a)):
            for buf in (step.dist, step.grad_queries, step.grad_refs):
                buf.fill_(-1.0)
            step.replay()
            torch.cuda.synchronize()
            ed, ei, eo, edq, edr = want[name]
            got = (step.dist, step.idx, step.orig, step.grad_queries, step.grad_refs)
            for what, g, w in zip(("dist", "idx", "orig", "dL/dq", "dL/dr"), got, (ed, ei, eo, edq, edr)):
                assert torch.equal(g, w), "step %s replay %d %s" % (name, rep, what)


def test_batch_autograd_with_mixed_array_and_tensor_clouds(cuda):
    """Clouds given partly as numpy arrays, partly as tensors requiring grad: the tensor clouds get
    their per-cloud gradients, the array clouds none, and the queries theirs."""
    rng = np.random.default_rng(77)
    clouds, queries, ks = _clouds(rng, 4, 3, np.float32, sizes=(200, 3000))
    given = [torch.from_numpy(c).to(cuda).requires_grad_(True) if i % 2 == 0 else c for i, c in enumerate(clouds)]
    qs = [torch.from_numpy(q).to(cuda).requires_grad_(True) for q in queries]
    ws = [torch.from_numpy(rng.random((q.shape[0], k)).astype(np.float32)).to(cuda) for q, k in zip(queries, ks)]
    batch = tk.build_kd_tree_batch(given, device=cuda)
    dists, _ = batch.query(qs, ks)
    sum((w * d).sum() for w, d in zip(ws, dists)).backward()
    for c in range(len(clouds)):
        r2 = torch.from_numpy(clouds[c]).to(cuda).requires_grad_(True)
        q2 = torch.from_numpy(queries[c]).to(cuda).requires_grad_(True)
        d2, _ = tk.build_kd_tree(r2, device=cuda, levels=batch.levels[c]).query(q2, nr_nns_searches=ks[c])
        if q2.shape[0]:
            (ws[c] * d2).sum().backward()
            assert torch.equal(q2.grad, qs[c].grad), "cloud %d dL/dq" % c
            if c % 2 == 0:
                assert torch.equal(r2.grad, given[c].grad), "cloud %d dL/dr" % c


def test_batch_refuses_nonfinite_clouds_under_reference_policy(cuda):
    """A cloud with a NaN coordinate: the reference policy's batched plan refuses it with a message
    (its single-tree query follows the reference's own steps); the lower-index plan accepts it and
    equals brute force, which never inserts a NaN distance."""
    rng = np.random.default_rng(5)
    a = rng.random((3000, 3)).astype(np.float32)
    b = rng.random((2000, 3)).astype(np.float32)
    b[17, 2] = np.nan
    qa, qb = rng.random((500, 3)).astype(np.float32), rng.random((400, 3)).astype(np.float32)
    batch = tk.build_kd_tree_batch([a, b], device=cuda)
    with pytest.raises(RuntimeError, match="non-finite"):
        batch.plan([500, 400], [4, 4])
    low = tk.build_kd_tree_batch([a, b], device=cuda, tie_break="lower_index")
    dists, inds = low.query([torch.from_numpy(qa), torch.from_numpy(qb)], [4, 4])
    for pts, qry, d, i in ((a, qa, dists[0], inds[0]), (b, qb, dists[1], inds[1])):
        bd, bi = oracle.bruteforce(pts, qry, 4)
        assert_same(d.cpu().numpy(), bd, "distances")
        assert_same(i.cpu().numpy(), bi.astype(np.int64), "indices")


@pytest.mark.parametrize("tie", ["reference", "lower_index"])
def test_persistent_launch_equals_per_bucket_launches(cuda, tie):
    """The persistent traversal (one launch, per-XCD work queues of 64-query units) gives the
    per-bucket launches' outputs bit for bit, eager and in a captured fwd+bwd replayed twice, with k
    buckets 1 to 32 present."""
    rng = np.random.default_rng(404 + (tie == "lower_index"))
    clouds, queries, _ = _clouds(rng, 14, 3, np.float32, sizes=(40, 9000))
    ks = [[1, 2, 3, 4, 7, 8, 12, 16, 20, 32][c % 10] for c in range(len(clouds))]
    ks = [min(k, c.shape[0]) for k, c in zip(ks, clouds)]
    weights = [torch.from_numpy(rng.random((q.shape[0], k)).astype(np.float32)) for q, k in zip(queries, ks)]
    qt = [torch.from_numpy(q) for q in queries]
    outs = {}
    for persistent in (False, True):
        batch = tk.build_kd_tree_batch(clouds, device=cuda, tie_break=tie, persistent=persistent)
        step = tk.BatchStep(batch, qt, ks, weights)
        got = []
        for _ in range(2):
            step.replay()
            torch.cuda.synchronize()
            got.append([t.clone() for t in (step.dist, step.idx, step.orig, step.grad_queries, step.grad_refs)])
        for a, b in zip(got[0], got[1]):
            assert torch.equal(a, b), "replays differ"
        plan = batch.new_plan([q.shape[0] for q in queries], ks)
        total_q, total_out, _, _ = plan.sizes()
        q_cat = torch.cat([t.to(cuda) for t in qt])
        ed, ei, eo = _plan_forward(plan, q_cat, total_out, torch.float32, cuda)
        outs[persistent] = got[0] + [ed, ei, eo]
    for what, a, b in zip(("dist", "idx", "orig", "dL/dq", "dL/dr", "eager dist", "eager idx", "eager orig"),
                          outs[False], outs[True]):
        assert torch.equal(a, b), "persistent vs per-bucket: " + what
