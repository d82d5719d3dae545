"""The reference's own test grid (tests/test_knn_unit.py:30-70) on device='cuda': nr_refs x nr_query
x k x d x dtype x requires_grad.  Beyond the reference's checks (shapes, devices, the error for
k > nr_refs, scipy's cKDTree within 1e-5 and 95 % index agreement), every output is compared with
the oracle bit for bit (distances, structured and original indices), and with requires_grad the
gradients of 0.5 * sum(dists) with the closed forms of tests/test_knn_unit.py:90-107."""
import numpy as np
import pytest
import torch

import oracle
import torch_kdtree_amd as tk

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nr_refs", [5, 10, 20, 30, 1000, 10000])
@pytest.mark.parametrize("nr_query", [5, 10, 20, 30, 1000])
@pytest.mark.parametrize("k", [1, 5, 10, 20])
@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("requires_grad", [True, False])
def test_reference_grid_cuda(cuda, nr_refs, nr_query, k, d, dtype, requires_grad):
    from scipy.spatial import cKDTree
    g = torch.Generator().manual_seed(nr_refs * 31 + nr_query * 7 + k + d + 1000 * requires_grad)
    points_ref = (torch.randn(size=(nr_refs, d), dtype=dtype, generator=g) * 1e3).to(cuda)
    points_query = (torch.randn(size=(nr_query, d), dtype=dtype, generator=g) * 1e3).to(cuda)
    points_ref.requires_grad_(requires_grad)
    points_query.requires_grad_(requires_grad)
    t = tk.build_kd_tree(points_ref, device=cuda)
    if k > nr_refs:
        with pytest.raises(Exception):
            t.query(points_query, nr_nns_searches=k)
        return
    res_i = torch.empty((nr_query, k), dtype=torch.int32, device=cuda)
    dists, inds = t.query(points_query, nr_nns_searches=k, result_idx=res_i)
    assert dists.device.type == "cuda" and inds.device.type == "cuda"
    assert inds.shape == (nr_query, k) and dists.shape == (nr_query, k) and inds.dtype == torch.int64
    r_np = points_ref.detach().cpu().numpy()
    q_np = points_query.detach().cpu().numpy()
    dref, iref = cKDTree(r_np).query(q_np, k)
    dref = np.asarray(dref).reshape(nr_query, k)
    iref = np.asarray(iref).reshape(nr_query, k)
    d_np = dists.detach().cpu().numpy()
    i_np = inds.cpu().numpy()
    assert np.allclose(dref ** 2, d_np, atol=1e-5, rtol=1e-5 if dtype == torch.float64 else 1e-4)
    assert np.sum(iref == i_np) / iref.size > 0.95
    ot = oracle.tree_from_product(t)
    od, oi = ot.search(q_np, k)
    assert np.array_equal(d_np, od), "distances vs oracle"
    assert np.array_equal(res_i.cpu().numpy(), oi), "structured indices vs oracle"
    assert np.array_equal(i_np, t.kdtree.get_shuffled_inds()[oi].astype(np.int64)), "original indices vs oracle"
    if requires_grad:
        (0.5 * torch.sum(dists)).backward()
        q64, r64 = q_np.astype(np.float64), r_np.astype(np.float64)
        gq = np.sum(q64[:, None] - r64[i_np], axis=-2)
        gr = np.zeros_like(r64)
        np.add.at(gr, i_np.reshape(-1), (r64[i_np] - q64[:, None]).reshape(-1, d))
        rtol = 1e-10 if dtype == torch.float64 else 1e-5
        scale = max(1.0, float(np.abs(gq).max()), float(np.abs(gr).max()))
        assert np.allclose(points_query.grad.cpu().numpy(), gq, rtol=rtol, atol=rtol * scale), "dL/dq closed form"
        assert np.allclose(points_ref.grad.cpu().numpy(), gr, rtol=rtol, atol=rtol * scale), "dL/dr closed form"
