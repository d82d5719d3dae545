"""GPU parity of the brute-force kernel and of the hand-written backward."""
import numpy as np
import pytest
import torch

import oracle
import workloads
import torch_kdtree_amd as tk
from torch_kdtree_amd import torch_knn
from tests._util import assert_same, with_nan_rows

pytestmark = pytest.mark.gpu


def _bruteforce_gpu(ref, qry, k, device):
    r = torch.from_numpy(ref).to(device)
    q = torch.from_numpy(qry).to(device)
    d = torch.empty((q.shape[0], k), dtype=r.dtype, device=device)
    i = torch.empty((q.shape[0], k), dtype=torch.int32, device=device)
    torch_knn.knn_bruteforce(r.data_ptr(), r.shape[0], q.data_ptr(), q.shape[0], r.shape[1],
                             str(r.dtype).replace("torch.", ""), k, d.data_ptr(), i.data_ptr(),
                             torch.cuda.current_stream(device).cuda_stream)
    torch.cuda.synchronize()
    return d.cpu().numpy(), i.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
@pytest.mark.parametrize("k", [1, 4, 16, 32, 33, 64, 100, 128])
def test_bruteforce_matches_oracle(cuda, dtype, d, k):
    rng = np.random.default_rng(300 + d + k)
    ref = (rng.standard_normal((3000, d)) * 1e3).astype(dtype)
    qry = (rng.standard_normal((2500, d)) * 1e3).astype(dtype)
    gd, gi = _bruteforce_gpu(ref, qry, k, cuda)
    od, oi = oracle.bruteforce(ref, qry, k)
    assert_same(gd, od, "distances")
    assert_same(gi, oi, "indices")


@pytest.mark.parametrize("k", [40, 128])
@pytest.mark.parametrize("nq", [100, 70000])
def test_bruteforce_large_k_chunked(cuda, k, nq):
    """Lists beyond 32 entries with the reference set split over chunks (merged in chunk order) and
    with one chunk; ties from an integer lattice go to the lower index, as in the oracle."""
    rng = np.random.default_rng(k + nq)
    ref = rng.integers(0, 12, (20000, 3)).astype(np.float32)
    qry = (rng.integers(0, 12, (nq, 3)) + 0.5 * rng.integers(0, 2, (nq, 3))).astype(np.float32)
    gd, gi = _bruteforce_gpu(ref, qry, k, cuda)
    od, oi = oracle.bruteforce(ref, qry, k)
    assert_same(gd, od, "distances")
    assert_same(gi, oi, "indices")


def test_bruteforce_k_above_128_is_refused(cuda):
    with pytest.raises(RuntimeError, match="128"):
        _bruteforce_gpu(np.zeros((300, 3), np.float32), np.zeros((5, 3), np.float32), 129, cuda)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("k", [1, 7, 32, 64])
def test_bruteforce_nan_query_rows(cuda, dtype, k):
    """A query with a NaN coordinate: the reference's brute force inserts only when d < worst
    (src/nndistance_g.cu:57), which a NaN never is, so the row stays (+inf, 0), as in the oracle;
    reference points with a NaN coordinate are never returned.  The other rows as without them."""
    rng = np.random.default_rng(320 + k)
    ref = (rng.standard_normal((3000, 3)) * 1e3).astype(dtype)
    qry = (rng.standard_normal((2500, 3)) * 1e3).astype(dtype)
    qn, rows = with_nan_rows(qry, rng)
    gd, gi = _bruteforce_gpu(ref, qn, k, cuda)
    od, oi = oracle.bruteforce(ref, qn, k)
    assert_same(gd, od, "distances")
    assert_same(gi, oi, "indices")
    assert np.isposinf(gd[rows]).all() and (gi[rows] == 0).all()
    rn, rrows = with_nan_rows(ref, rng, frac=0.05)
    gd, gi = _bruteforce_gpu(rn, qn, k, cuda)
    od, oi = oracle.bruteforce(rn, qn, k)
    assert_same(gd, od, "distances, NaN reference rows")
    assert_same(gi, oi, "indices, NaN reference rows")
    assert not np.isin(gi, rrows).any()


def test_backward_with_nan_query_rows(cuda):
    """fwd+bwd where some queries hold NaN: dL/dq of those queries and dL/dr of the points they
    returned are NaN on the NaN axes, as in autograd of the reference's recomputation
    (nn_distance.py:112) in fp64; every other gradient entry within the fp32 tolerance of that
    recomputation."""
    w = workloads.diff_sphere(n=20000)
    rng = np.random.default_rng(8)
    qn, rows = with_nan_rows(w["queries"], rng, frac=0.02)
    wt = torch.from_numpy(w["weights"]).to(cuda)
    r = torch.from_numpy(w["ref"]).to(cuda).requires_grad_(True)
    q = torch.from_numpy(qn).to(cuda).requires_grad_(True)
    tree = tk.build_kd_tree(r, levels="auto")
    dist, idx = tree.query(q, nr_nns_searches=8)
    (wt * dist).sum().backward()
    r64 = r.detach().double().requires_grad_(True)
    q64 = q.detach().double().requires_grad_(True)
    (wt.double() * torch.sum((q64[:, None] - r64[idx]) ** 2, dim=-1)).sum().backward()
    for got, want, what in ((q.grad, q64.grad, "dL/dq"), (r.grad, r64.grad, "dL/dr")):
        got = got.double().cpu().numpy()
        want = want.cpu().numpy()
        assert np.array_equal(np.isnan(got), np.isnan(want)), what + " NaN positions"
        fin = ~np.isnan(want)
        scale = np.abs(want[fin]).max()
        assert np.abs(got[fin] - want[fin]).max() <= 1e-5 * scale, what
    assert np.isnan(q.grad.cpu().numpy()[rows]).any(axis=1).all()


def test_bruteforce_ties_go_to_lower_index(cuda):
    rng = np.random.default_rng(13)
    ref = rng.integers(0, 4, (4000, 3)).astype(np.float32)
    qry = rng.integers(0, 4, (1000, 3)).astype(np.float32)
    gd, gi = _bruteforce_gpu(ref, qry, 16, cuda)
    od, oi = oracle.bruteforce(ref, qry, 16)
    assert_same(gd, od, "distances")
    assert_same(gi, oi, "indices")


def test_bruteforce_cross_checks_tree_on_plain_gpu_subset(cuda):
    """Workload 2's brute-force cross-check: 65,536 refs x 65,536 queries (half from each query half), k=16."""
    w = workloads.plain_gpu(n=4_000_000, q=16_000_000)
    ref = np.ascontiguousarray(w["ref"][:65536])
    half = w["halves"][0]
    qry = np.ascontiguousarray(np.concatenate([w["queries"][:32768], w["queries"][half:half + 32768]]))
    gd, gi = _bruteforce_gpu(ref, qry, 16, cuda)
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=cuda, levels="auto", tie_break="lower_index")
    td, ti = tree.query(torch.from_numpy(qry).to(cuda), nr_nns_searches=16)
    assert_same(gd, td.cpu().numpy(), "brute force vs tree distances")
    assert_same(gi.astype(np.int64), ti.cpu().numpy(), "brute force vs tree indices")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_grad_closed_form(cuda, dtype):
    """The reference's test_grad (tests/test_knn_unit.py:90-107) on the GPU path."""
    torch.manual_seed(0)
    points_ref = torch.randn(size=(10, 3), dtype=dtype, device=cuda, requires_grad=True)
    t_kdtree = tk.build_kd_tree(points_ref)
    points_query = torch.randn(size=(5, 3), dtype=dtype, device=cuda, requires_grad=True)
    dists, inds = t_kdtree.query(points_query, nr_nns_searches=2)
    (0.5 * torch.sum(dists)).backward()
    assert torch.allclose(points_query.grad, torch.sum((points_query[:, None] - points_ref[inds]), axis=-2))
    ref_grad = torch.zeros_like(points_ref)
    for i, ind in enumerate(inds):
        ref_grad[ind] += points_ref[ind] - points_query[i]
    assert torch.allclose(points_ref.grad, ref_grad)


def _loss_fd(ref, qry, idx, w):
    """Loss of fixed neighbours in float64, for central differences."""
    diff = qry[:, None, :] - ref[idx]
    return float(np.sum(w * np.sum(diff * diff, axis=-1)))


def test_grad_central_differences_fp64(cuda):
    """Workload 3 subset in fp64: gradients vs central differences (h=1e-6), skipping
    perturbations that change a neighbour index."""
    w = workloads.diff_sphere(n=2000, dtype=np.float64)
    ref = w["ref"]
    qry = w["queries"]
    wt = w["weights"][:, :8].astype(np.float64)
    r = torch.from_numpy(ref).to(cuda).requires_grad_(True)
    q = torch.from_numpy(qry).to(cuda).requires_grad_(True)
    tree = tk.build_kd_tree(r, levels="auto")
    dist, idx = tree.query(q, nr_nns_searches=8)
    (torch.from_numpy(wt).to(cuda) * dist).sum().backward()
    gq = q.grad.cpu().numpy()
    gr = r.grad.cpu().numpy()
    idx = idx.cpu().numpy()
    h = 1e-6
    rng = np.random.default_rng(0)
    checked = 0
    for which, arr, grad in (("q", qry, gq), ("r", ref, gr)):
        rows = rng.choice(arr.shape[0], 60, replace=False)
        for row in rows:
            for a in range(3):
                plus = arr.copy()
                minus = arr.copy()
                plus[row, a] += h
                minus[row, a] -= h
                if which == "q":
                    fp = _loss_fd(ref, plus, idx, wt)
                    fm = _loss_fd(ref, minus, idx, wt)
                    _, ip = oracle.bruteforce(ref, plus[row:row + 1], 8)
                    _, im = oracle.bruteforce(ref, minus[row:row + 1], 8)
                    if not (np.array_equal(np.sort(ip[0]), np.sort(idx[row])) and
                            np.array_equal(np.sort(im[0]), np.sort(idx[row]))):
                        continue
                else:
                    fp = _loss_fd(plus, qry, idx, wt)
                    fm = _loss_fd(minus, qry, idx, wt)
                fd = (fp - fm) / (2 * h)
                assert abs(fd - grad[row, a]) <= 1e-6 * max(1.0, abs(fd)), (which, row, a, fd, grad[row, a])
                checked += 1
    assert checked > 200


def test_grad_matches_reference_formula_and_is_deterministic(cuda):
    """Gradients vs torch autograd of the reference's recomputation (nn_distance.py:112) in fp32,
    and bit-identical across two runs."""
    w = workloads.diff_sphere(n=50000)
    ref = torch.from_numpy(w["ref"]).to(cuda)
    qry = torch.from_numpy(w["queries"]).to(cuda)
    wt = torch.from_numpy(w["weights"]).to(cuda)
    grads = []
    for _ in range(2):
        r = ref.clone().requires_grad_(True)
        q = qry.clone().requires_grad_(True)
        tree = tk.build_kd_tree(r, levels="auto")
        dist, idx = tree.query(q, nr_nns_searches=8)
        (wt * dist).sum().backward()
        grads.append((q.grad.clone(), r.grad.clone(), idx))
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])
    r = ref.double().clone().requires_grad_(True)
    q = qry.double().clone().requires_grad_(True)
    idx = grads[0][2]
    d_ref = torch.sum((q[:, None] - r[idx]) ** 2, dim=-1)
    (wt.double() * d_ref).sum().backward()
    # fp32 kernel vs fp64 autograd of the same formula: relative 1e-5 of the gradient scale
    for got, want in ((grads[0][0], q.grad), (grads[0][1], r.grad)):
        err = (got.double() - want).abs().max().item()
        scale = want.abs().max().item()
        assert err <= 1e-5 * scale, (err, scale)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
@pytest.mark.parametrize("k", [1, 5, 16, 40])
def test_tree_backward_equals_original_index_backward(cuda, dtype, d, k):
    """tk_knn_backward_tree_gpu (neighbours by structured index, points gathered into the tree's
    structured order) gives tk_knn_backward_gpu's dL/dq and dL/dr bit for bit, with and without
    dL/dr, and the autograd path of a GPU query (which takes the tree form) gives the same."""
    gen = torch.Generator(device=cuda).manual_seed(50 + d + k)
    n, nq = 20000, 7000
    ref = torch.randn((n, d), device=cuda, generator=gen, dtype=dtype) * 10
    q = torch.randn((nq, d), device=cuda, generator=gen, dtype=dtype) * 10
    g = torch.rand((nq, k), device=cuda, generator=gen, dtype=dtype)
    tree = tk.build_kd_tree(ref, device=cuda, levels="auto")
    sidx = torch.empty((nq, k), dtype=torch.int32, device=cuda)
    _, idx = tree.query(q, nr_nns_searches=k, result_idx=sidx)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(cuda).cuda_stream
    name = str(dtype).replace("torch.", "")
    full = None
    for with_r in (True, False):
        dq0, dr0 = torch.empty_like(q), torch.empty_like(ref)
        torch_knn.knn_backward(q.data_ptr(), ref.data_ptr(), nq, n, d, name, k, idx.data_ptr(), g.data_ptr(),
                               dq0.data_ptr(), dr0.data_ptr() if with_r else 0, stream)
        dq1, dr1 = torch.empty_like(q), torch.empty_like(ref)
        tree.kdtree.backward(q.data_ptr(), ref.data_ptr(), nq, k, sidx.data_ptr(), g.data_ptr(), dq1.data_ptr(),
                             dr1.data_ptr() if with_r else 0, stream)
        torch.cuda.synchronize()
        assert torch.equal(dq0, dq1), "dL/dq"
        if with_r:
            assert torch.equal(dr0, dr1), "dL/dr"
            full = (dq0, dr0)
    qq = q.clone().requires_grad_(True)
    rr = ref.clone().requires_grad_(True)
    t2 = tk.build_kd_tree(rr, device=cuda, levels="auto")
    dist, _ = t2.query(qq, nr_nns_searches=k)
    (g * dist).sum().backward()
    assert torch.equal(qq.grad, full[0]) and torch.equal(rr.grad, full[1]), "autograd path"

