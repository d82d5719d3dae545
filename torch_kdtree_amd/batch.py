"""Batched many-cloud KNN on one GPU.

Many point clouds, each with its own tree, k and query set, are queried together: one presort
over all queries, one traversal launch per k bucket and one backward over all clouds, instead of a
loop of ``TorchKDTree.query`` calls (``nn_distance.py:49-117`` of the reference, once per cloud).
Each cloud's results equal that loop's results bit for bit, under the same tie policy.

    batch = build_kd_tree_batch([refs_0, refs_1, ...], device="cuda")
    dists, inds = batch.query([q_0, q_1, ...], nr_nns_searches=[16, 4, ...])

``BatchStep`` captures forward + backward with fixed buffers in a HIP graph and replays it.
"""
from typing import List, Sequence, Union

import numpy as np
import torch

from . import autotune
from .nn_distance import _TIE_POLICIES, _dtype_name, _stream_handle, default_levels, MAX_LEVELS, torch_knn

__all__ = ["KDTreeBatch", "BatchStep", "build_kd_tree_batch"]

_DTYPE_CODES = {torch.float32: 0, torch.float64: 1}


def _as_host_array(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(x)


class _BatchKnnSquaredDistances(torch.autograd.Function):
    """Squared distances of a batched query with the batched HIP backward (tk_batch_backward)."""

    @staticmethod
    def forward(ctx, q_cat, r_cat, plan, idx, dists):
        ctx.plan = plan
        ctx.save_for_backward(q_cat, idx)
        ctx.n_ref = r_cat.shape[0]
        return dists.clone()

    @staticmethod
    def backward(ctx, grad):
        q_cat, idx = ctx.saved_tensors
        need_q, need_r = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_q or need_r):
            return None, None, None, None, None
        q = q_cat.detach().contiguous()
        g = grad.contiguous()
        dq = torch.empty_like(q) if need_q else None
        dr = torch.empty((ctx.n_ref, q.shape[1]), dtype=q.dtype, device=q.device) if need_r else None
        ctx.plan.backward(q.data_ptr(), idx.data_ptr(), g.data_ptr(), dq.data_ptr() if need_q else 0,
                          dr.data_ptr() if need_r else 0, _stream_handle(q.device))
        return dq, dr, None, None, None


class KDTreeBatch:
    """Trees of many point clouds (same dims and dtype), built in parallel on the host.

    Parameters
    ----------
    clouds : sequence of (N_c, d) tensors or arrays
    device : torch device of the queries (a GPU)
    levels : "auto" (a depth chosen per cloud for the GPU), None (the reference's default rule per
        cloud), an int for all clouds, or one value per cloud
    tie_break : "reference" or "lower_index", as for ``build_kd_tree``
    builder : "cpu" (default: the reference's build, in parallel over clouds) or "gpu" (one GPU
        build of all trees; see ``build_kd_tree`` for how it can differ on tied coordinates)
    persistent : the traversal as one persistent launch for every k bucket, whose wavefronts take
        units of 64 queries from per-XCD work queues (True), or one launch per k bucket (False,
        the default: 5 % faster on the 256-cloud workload, because each bucket's kernel keeps its
        own register budget and occupancy); the results are identical.
    """

    def __init__(self, clouds: Sequence, device=None, levels: Union[str, int, None, List] = "auto",
                 tie_break: str = "reference", n_threads: int = 0, builder: str = "cpu", persistent: bool = False):
        if tie_break not in _TIE_POLICIES:
            raise ValueError("tie_break must be one of %s" % sorted(_TIE_POLICIES))
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else "cpu"
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("the batched mode runs on a GPU device")
        self.tie_break = tie_break
        self.persistent = bool(persistent)
        if builder not in ("cpu", "gpu"):
            raise ValueError("builder must be 'cpu' or 'gpu'")
        if not len(clouds):
            raise ValueError("no clouds")
        if builder == "gpu":
            tensors = [c.detach() if isinstance(c, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(c))
                       for c in clouds]
            self.dims = tensors[0].shape[1]
            self.dtype = tensors[0].dtype
            n = [t.shape[0] for t in tensors]
        else:
            arrays = [_as_host_array(c) for c in clouds]
            self.dims = arrays[0].shape[1]
            self.dtype = torch.from_numpy(arrays[0][:0]).dtype
            n = [a.shape[0] for a in arrays]
        if isinstance(levels, (list, tuple)):
            lv = [int(x) for x in levels]
        elif levels == "auto":
            lv = [autotune.pick_levels(m, self.dims, self.device, self.dtype) for m in n]
        elif levels is None:
            lv = [default_levels(m) for m in n]
        else:
            lv = [int(levels)] * len(n)
        if len(lv) != len(n) or not all(1 <= x <= MAX_LEVELS for x in lv):
            raise ValueError("levels must be in [1, %d], one per cloud" % MAX_LEVELS)
        if builder == "gpu":
            for t in tensors:
                if t.dim() != 2 or t.shape[1] != self.dims or t.dtype != self.dtype:
                    raise RuntimeError("all clouds must have the same dims and dtype")
            pts = torch.cat([t.to(self.device) for t in tensors]).contiguous()
            self.trees = torch_knn.TreeSet.from_device(pts.data_ptr(), n, lv, self.dims, _DTYPE_CODES[self.dtype],
                                                       self.device.index if self.device.index is not None else
                                                       torch.cuda.current_device(), _stream_handle(self.device))
        else:
            self.trees = torch_knn.TreeSet(arrays, lv, n_threads)
        self.levels = list(self.trees.levels)
        self.nr_ref_points = n
        self._clouds = [c if isinstance(c, torch.Tensor) else None for c in clouds]
        self._plans = {}

    def new_plan(self, n_queries: Sequence[int], ks: Sequence[int]):
        """A plan of its own (arena, scratch and, once captured, graph) for these query counts and k."""
        p = torch_knn.BatchPlan(self.trees, [int(x) for x in n_queries], [int(x) for x in ks],
                                _TIE_POLICIES[self.tie_break],
                                self.device.index if self.device.index is not None else -1)
        p.set_persistent(self.persistent)
        return p

    def plan(self, n_queries: Sequence[int], ks: Sequence[int]):
        """The plan ``query`` uses for these query counts and k, created once and then reused (eager
        forward and backward only; a captured graph needs a plan of its own, ``new_plan``)."""
        key = (tuple(int(x) for x in n_queries), tuple(int(x) for x in ks))
        p = self._plans.get(key)
        if p is None:
            p = self.new_plan(key[0], key[1])
            self._plans[key] = p
        return p

    def _refs_for_grad(self):
        """The clouds concatenated for autograd when any tensor cloud requires grad: tensor clouds as
        given (so gradients flow back to each of them), clouds given as arrays as zeros without grad."""
        refs = self._clouds
        if not any(r is not None and r.requires_grad for r in refs):
            return None
        parts = [r.to(self.device) if r is not None else
                 torch.zeros((n, self.dims), dtype=self.dtype, device=self.device)
                 for r, n in zip(refs, self.nr_ref_points)]
        return torch.cat(parts)

    def query(self, points_query: Sequence[torch.Tensor], nr_nns_searches: Union[int, Sequence[int]] = 1):
        """KNN of every cloud's queries in its own tree.

        Returns (dists, inds): lists with one [Q_c, k_c] tensor per cloud, squared distances and
        int64 original indices, as ``TorchKDTree.query`` returns them.  Gradients flow to the
        queries and to the clouds given as tensors that require grad (clouds given as arrays get
        none).  The gradient of the reference points is computed at the coordinates the trees were
        built from (the single-tree path uses the tensor's current values, ``points_ref_bak``).
        """
        if len(points_query) != len(self.nr_ref_points):
            raise ValueError("one query tensor per cloud is required")
        ks = [int(nr_nns_searches)] * len(points_query) if isinstance(nr_nns_searches, int) else \
            [int(x) for x in nr_nns_searches]
        for k, n in zip(ks, self.nr_ref_points):
            if k > n:
                raise RuntimeError("You requested more nearest neighbors than there are in the KD-Tree")
        qs = [q.to(self.device) for q in points_query]
        for q in qs:
            assert q.dtype == self.dtype and q.dim() == 2 and q.shape[1] == self.dims
        nq = [q.shape[0] for q in qs]
        plan = self.plan(nq, ks)
        q_cat = torch.cat(qs) if len(qs) > 1 else qs[0]
        q_flat = q_cat.detach().contiguous()
        total_q, total_out, _, _ = plan.sizes()
        dist = torch.empty(total_out, dtype=self.dtype, device=self.device)
        idx = torch.empty(total_out, dtype=torch.int32, device=self.device)
        orig = torch.empty(total_out, dtype=torch.int64, device=self.device)
        plan.forward(q_flat.data_ptr(), dist.data_ptr(), idx.data_ptr(), orig.data_ptr(), _stream_handle(self.device))
        r_cat = self._refs_for_grad()
        if (q_cat.requires_grad or r_cat is not None) and torch.is_grad_enabled():
            if r_cat is None:
                r_cat = torch.empty((sum(self.nr_ref_points), self.dims), dtype=self.dtype, device=self.device)
            dist = _BatchKnnSquaredDistances.apply(q_cat, r_cat, plan, idx, dist)
        sizes = [n * k for n, k in zip(nq, ks)]
        dists = [d.view(n, k) for d, n, k in zip(torch.split(dist, sizes), nq, ks)]
        inds = [i.view(n, k) for i, n, k in zip(torch.split(orig, sizes), nq, ks)]
        return dists, inds


class BatchStep:
    """Forward + backward of a batched query with fixed buffers, captured in a HIP graph.

    ``weights`` (one [Q_c, k_c] tensor per cloud) is the upstream gradient of the squared
    distances, i.e. the loss is sum_c sum(weights_c * dists_c).  After ``replay()`` the buffers
    ``dist``, ``idx``, ``orig``, ``grad_queries`` and ``grad_refs`` hold the results (flat, in the
    plan's layout; ``split`` cuts a flat buffer per cloud).  Each step owns its plan and graph, so
    several steps of equal shapes (double-buffered inputs, a train and an eval step) coexist.
    """

    def __init__(self, batch: KDTreeBatch, queries: Sequence[torch.Tensor], nr_nns_searches, weights=None):
        ks = [int(nr_nns_searches)] * len(queries) if isinstance(nr_nns_searches, int) else \
            [int(x) for x in nr_nns_searches]
        dev = batch.device
        self.batch = batch
        self.nq = [q.shape[0] for q in queries]
        self.ks = ks
        self.plan = batch.new_plan(self.nq, ks)
        total_q, total_out, total_ref, _ = self.plan.sizes()
        self.queries = torch.cat([q.to(dev) for q in queries]).detach().contiguous()
        self.dist = torch.empty(total_out, dtype=batch.dtype, device=dev)
        self.idx = torch.empty(total_out, dtype=torch.int32, device=dev)
        self.orig = torch.empty(total_out, dtype=torch.int64, device=dev)
        self.weights = None
        self.grad_queries = self.grad_refs = None
        if weights is not None:
            self.weights = torch.cat([w.to(dev).reshape(-1) for w in weights]).contiguous()
            self.grad_queries = torch.empty_like(self.queries)
            self.grad_refs = torch.empty((total_ref, batch.dims), dtype=batch.dtype, device=dev)
        stream = _stream_handle(dev)
        ptr = lambda t: t.data_ptr() if t is not None else 0  # noqa: E731
        self.plan.capture(ptr(self.queries), ptr(self.dist), ptr(self.idx), ptr(self.orig), ptr(self.weights),
                          ptr(self.grad_queries), ptr(self.grad_refs), stream)

    def replay(self):
        self.plan.launch(_stream_handle(self.batch.device))

    def split(self, flat: torch.Tensor, per: str = "output"):
        """Cuts a flat buffer into per-cloud tensors ("output": [Q_c, k_c]; "query": [Q_c, d];
        "ref": [N_c, d])."""
        if per == "output":
            return [t.view(n, k) for t, n, k in zip(torch.split(flat, [n * k for n, k in zip(self.nq, self.ks)]),
                                                    self.nq, self.ks)]
        if per == "query":
            return list(torch.split(flat, self.nq))
        return list(torch.split(flat, self.batch.nr_ref_points))


def build_kd_tree_batch(clouds: Sequence, device=None, levels="auto", tie_break: str = "reference",
                        builder: str = "cpu", persistent: bool = False) -> KDTreeBatch:
    """Builds the trees of many clouds for batched queries (see ``KDTreeBatch``)."""
    return KDTreeBatch(clouds, device=device, levels=levels, tie_break=tie_break, builder=builder,
                       persistent=persistent)
