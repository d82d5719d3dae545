"""Full-size parity of the benchmark configurations: all 16,000,000 queries of the plain GPU workload
against the oracle (at the auto depth and at the reference's default depth), all 256 clouds of the batching workload (forward against the oracle, captured
fwd+bwd against the single-tree path and fp64 autograd), the d=8 k=32 backward, and the backward
beyond 2^28 (query, neighbour) pairs."""
import numpy as np
import pytest
import torch

import oracle
import workloads
import torch_kdtree_amd as tk
from torch_kdtree_amd import torch_knn
from tests._util import assert_same, gpu_query, oracle_for

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

EPS32 = 2.0 ** -23


def _fp64_grads(q, r, idx, g, chunk=1 << 18):
    """dL/dq and dL/dr of loss = sum(g * sum_a (q - r[idx])^2) in fp64 (torch autograd's formula of
    nn_distance.py:112, written out), with per-entry sums of |term| and term counts for the
    recursive-summation error bound of the fp32 results."""
    nq, k = idx.shape
    d = q.shape[1]
    dq = torch.zeros((nq, d), dtype=torch.float64, device=q.device)
    aq = torch.zeros_like(dq)
    dr = torch.zeros((r.shape[0], d), dtype=torch.float64, device=q.device)
    ar = torch.zeros_like(dr)
    cr = torch.zeros(r.shape[0], dtype=torch.float64, device=q.device)
    r64 = r.double()
    for s in range(0, nq, chunk):
        e = min(nq, s + chunk)
        ii = idx[s:e]
        t = 2.0 * g[s:e, :, None].double() * (q[s:e, None, :].double() - r64[ii])
        dq[s:e] = t.sum(dim=1)
        aq[s:e] = t.abs().sum(dim=1)
        flat = ii.reshape(-1)
        dr.index_add_(0, flat, -t.reshape(-1, d))
        ar.index_add_(0, flat, t.abs().reshape(-1, d))
        cr.index_add_(0, flat, torch.ones(flat.shape[0], dtype=torch.float64, device=q.device))
    return dq, aq, dr, ar, cr


def _within_sum_bound(got, want, abs_sum, n_terms, what):
    """|fp32 result - exact| <= (n + 4) * 2^-23 * sum|terms|: recursive summation of n rounded
    terms, each a product of rounded factors (a loose form of the gamma_n bound)."""
    err = (got.double() - want).abs()
    bound = (n_terms + 4.0) * EPS32 * abs_sum + 1e-30
    bad = err > bound
    assert not bool(bad.any()), "%s: %d entries beyond the summation bound, worst ratio %.3g" % (
        what, int(bad.sum()), float((err / bound).max()))


def test_plain_gpu_all_queries(cuda):
    """The plain GPU workload in full: 4,000,000-point tree (auto depth), all 16,000,000 queries,
    k = 16, bit-exact against the oracle on the same tree."""
    w = workloads.plain_gpu()
    tree = tk.build_kd_tree(torch.from_numpy(w["ref"]), device=cuda, levels="auto")
    gd, gi, go = gpu_query(tree, w["queries"], 16, cuda)
    od, oi = oracle_for(tree).search(w["queries"], 16)
    assert_same(gd, od, "distances")
    assert_same(gi, oi, "structured indices")
    del gd, od
    assert_same(go, tree.kdtree.get_shuffled_inds()[oi].astype(np.int64), "original indices")


def test_plain_gpu_all_queries_reference_default_depth(cuda):
    """The same 16,000,000 queries on the tree a plain build_kd_tree(points) gives (depth 13, 488
    points a leaf, record-group boxes), k = 16, bit-exact against the oracle on the same tree."""
    w = workloads.plain_gpu()
    tree = tk.build_kd_tree(torch.from_numpy(w["ref"]), device=cuda)
    assert tree.levels == 13
    gd, gi, go = gpu_query(tree, w["queries"], 16, cuda)
    od, oi = oracle_for(tree).search(w["queries"], 16)
    assert_same(gd, od, "distances")
    assert_same(gi, oi, "structured indices")
    del gd, od
    assert_same(go, tree.kdtree.get_shuffled_inds()[oi].astype(np.int64), "original indices")


def test_batch_clouds_all_256(cuda):
    """The batching workload in full (256 clouds, seed 3, 18.7M points, 22.1M queries, k in
    {1, 4, 16}): one captured fwd+bwd replay.  Forward of every cloud bit-exact against the oracle
    on the cloud's own tree; dL/dq and dL/dr of every cloud bit-exact against the single-tree
    backward, and within the fp32 summation bound of fp64 autograd of the reference's
    recomputation (nn_distance.py:112)."""
    clouds = workloads.batch_clouds()
    refs = [c["ref"] for c in clouds]
    qs = [c["queries"] for c in clouds]
    ks = [c["k"] for c in clouds]
    rng = np.random.default_rng(5)
    ws = [rng.random((q.shape[0], k)).astype(np.float32) for q, k in zip(qs, ks)]
    batch = tk.build_kd_tree_batch(refs, device=cuda)
    step = tk.BatchStep(batch, [torch.from_numpy(q) for q in qs], ks, [torch.from_numpy(w) for w in ws])
    step.replay()
    torch.cuda.synchronize()
    dists, idxs, origs = step.split(step.dist), step.split(step.idx), step.split(step.orig)
    gqs, grs = step.split(step.grad_queries, "query"), step.split(step.grad_refs, "ref")
    stream = torch.cuda.current_stream(cuda).cuda_stream
    for c in range(len(clouds)):
        k = ks[c]
        host = tk.build_kd_tree(torch.from_numpy(refs[c]), device="cpu", levels=batch.levels[c])
        od, oi = oracle.tree_from_product(host).search(qs[c], k)
        assert_same(dists[c].cpu().numpy(), od, "cloud %d distances" % c)
        assert_same(idxs[c].cpu().numpy(), oi, "cloud %d structured indices" % c)
        assert_same(origs[c].cpu().numpy(), host.kdtree.get_shuffled_inds()[oi].astype(np.int64),
                    "cloud %d original indices" % c)
        qd = torch.from_numpy(qs[c]).to(cuda)
        rd = torch.from_numpy(refs[c]).to(cuda)
        wd = torch.from_numpy(ws[c]).to(cuda)
        sq, sr = torch.empty_like(qd), torch.empty_like(rd)
        oc = origs[c].contiguous()
        torch_knn.knn_backward(qd.data_ptr(), rd.data_ptr(), qd.shape[0], rd.shape[0], 3, "float32", k,
                               oc.data_ptr(), wd.data_ptr(), sq.data_ptr(), sr.data_ptr(), stream)
        torch.cuda.synchronize()
        assert torch.equal(gqs[c], sq), "cloud %d dL/dq vs single tree" % c
        assert torch.equal(grs[c], sr), "cloud %d dL/dr vs single tree" % c
        if c % 16 == 0 and qd.shape[0]:
            dq64, aq, dr64, ar, cr = _fp64_grads(qd, rd, oc, wd)
            _within_sum_bound(gqs[c], dq64, aq, k, "cloud %d dL/dq vs fp64" % c)
            _within_sum_bound(grs[c], dr64, ar, cr[:, None], "cloud %d dL/dr vs fp64" % c)


def test_d8_fwd_bwd_full(cuda):
    """The d=8 workload in full (2,000,000 points and queries, k = 32): fwd+bwd gradients within the
    fp32 summation bound of fp64 autograd of the reference's recomputation, and bit-identical on a
    second run."""
    w = workloads.d8_mixture()
    rng = np.random.default_rng(9)
    wt = torch.from_numpy(rng.random((w["queries"].shape[0], 32)).astype(np.float32)).to(cuda)
    runs = []
    for _ in range(2):
        r = torch.from_numpy(w["ref"]).to(cuda).requires_grad_(True)
        q = torch.from_numpy(w["queries"]).to(cuda).requires_grad_(True)
        tree = tk.build_kd_tree(r, levels="auto")
        dist, idx = tree.query(q, nr_nns_searches=32)
        (wt * dist).sum().backward()
        torch.cuda.synchronize()
        runs.append((dist.detach(), idx, q.grad, r.grad))
    for a, b, what in zip(runs[0], runs[1], ("distances", "indices", "dL/dq", "dL/dr")):
        assert torch.equal(a, b), what + " repeat"
    dist, idx, gq, gr = runs[0]
    q = torch.from_numpy(w["queries"]).to(cuda)
    r = torch.from_numpy(w["ref"]).to(cuda)
    dq64, aq, dr64, ar, cr = _fp64_grads(q, r, idx, wt)
    _within_sum_bound(gq, dq64, aq, 32, "dL/dq")
    _within_sum_bound(gr, dr64, ar, cr[:, None], "dL/dr")


def test_backward_beyond_2_28_pairs(cuda):
    """2^21 + 3 queries with k = 128 (2^28 + 384 pairs, so dL/dr runs in two slices of the canonical
    order) against 1,000,000 points, with synthetic neighbour indices: dL/dq and dL/dr within the
    fp32 summation bound of fp64, bit-identical on a second call, and dL/dq alone (no dL/dr, no
    workspace) equal to the dL/dq of the full call."""
    n, nq, k = 1_000_000, (1 << 21) + 3, 128
    gen = torch.Generator(device=cuda).manual_seed(0)
    ref = torch.rand((n, 3), device=cuda, generator=gen)
    q = torch.rand((nq, 3), device=cuda, generator=gen)
    idx = torch.randint(0, n, (nq, k), device=cuda, generator=gen, dtype=torch.int64)
    g = torch.rand((nq, k), device=cuda, generator=gen)
    stream = torch.cuda.current_stream(cuda).cuda_stream
    outs = []
    for _ in range(2):
        dq, dr = torch.empty_like(q), torch.empty_like(ref)
        torch_knn.knn_backward(q.data_ptr(), ref.data_ptr(), nq, n, 3, "float32", k, idx.data_ptr(), g.data_ptr(),
                               dq.data_ptr(), dr.data_ptr(), stream)
        torch.cuda.synchronize()
        outs.append((dq, dr))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1]), "repeat"
    dq_only = torch.empty_like(q)
    torch_knn.knn_backward(q.data_ptr(), ref.data_ptr(), nq, n, 3, "float32", k, idx.data_ptr(), g.data_ptr(),
                           dq_only.data_ptr(), 0, stream)
    torch.cuda.synchronize()
    assert torch.equal(dq_only, outs[0][0]), "dL/dq alone"
    dq64, aq, dr64, ar, cr = _fp64_grads(q, ref, idx, g)
    _within_sum_bound(outs[0][0], dq64, aq, k, "dL/dq")
    _within_sum_bound(outs[0][1], dr64, ar, cr[:, None], "dL/dr")
