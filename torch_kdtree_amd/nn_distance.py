"""Python surface of the MI355X KD-tree: ``build_kd_tree`` and ``TorchKDTree.query``.

Mirrors ``torch_kdtree/nn_distance.py`` of thomgrand/torch_kdtree (same names, argument
meaning, return values and error behaviour) over this package's own ``torch_knn`` module.

Differences, all opt-in or fixes of reference defects:

* GPU queries write the original ``int64`` indices from the kernel (the reference gathers
  ``shuffled_ind[result_idx.long()]`` afterwards in torch, ``nn_distance.py:109``).
* With autograd, the GPU path returns the kernel's own squared distances through a
  ``torch.autograd.Function`` whose backward is a HIP kernel; the reference recomputes the
  distances with torch ops (``nn_distance.py:111-112``).  CPU trees keep that recomputation.
* ``levels`` may go up to 24 (reference: 13, ``nn_distance.py:155``); ``levels=None`` keeps
  the reference's default (``nn_distance.py:146-147``) and ``levels="auto"`` picks a leaf
  size tuned for the GPU.
* ``tie_break`` ("reference" or "lower_index") selects how exact distance ties are ordered on
  the GPU; "reference" (default) reproduces the reference CPU search exactly.
* Non-contiguous ``result_dists``/``result_idx`` buffers receive the results (the reference
  silently writes to a contiguous copy, ``nn_distance.py:93-97``).
"""
from __future__ import annotations

from typing import Union

import numpy as np
import torch

from . import torch_knn
from . import autotune

gpu_available = torch_knn.check_for_gpu()
if not gpu_available:
    print("The library was not successfully compiled using CUDA. Only the CPU version will be available.")

_transl_torch_device = {"cpu": "CPU", "cuda": "GPU"}
_TIE_POLICIES = {"reference": 0, "lower_index": 1}
_PRESORT = {"auto": -1, "off": 0, "on": 1, "morton": 1, "leaf": 2, None: -1, False: 0, True: 1}
MAX_LEVELS = 24


def default_levels(nr_points: int) -> int:
    """The reference's default depth, ``nn_distance.py:146-147``."""
    return int(np.maximum(1, np.minimum(13, int(np.log(int(nr_points)) / np.log(2)) - 3)))


def _dtype_name(dtype: torch.dtype) -> str:
    return {torch.float32: "float32", torch.float64: "float64"}[dtype]


def _stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class _KnnSquaredDistances(torch.autograd.Function):
    """Squared KNN distances with a HIP backward (``tk_knn_backward_tree_gpu`` given the tree and the
    query's structured indices, else ``tk_knn_backward_gpu``; the two give the same bits).

    forward returns the distances the query kernel computed; backward recomputes the
    neighbour differences from the coordinates and returns
    dL/dq_i = sum_j 2 g_ij (q_i - r_idx_ij) and dL/dr_m = -sum_{idx_ij = m} 2 g_ij (q_i - r_m),
    the gradient of ``nn_distance.py:112`` of the reference.
    """

    @staticmethod
    def forward(ctx, points_query, points_ref, inds, dists, sidx=None, kdtree=None):
        ctx.save_for_backward(points_query, points_ref, inds, sidx)
        ctx.kdtree = kdtree
        return dists.clone()

    @staticmethod
    def backward(ctx, grad):
        points_query, points_ref, inds, sidx = ctx.saved_tensors
        need_q, need_r = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_q or need_r):
            return None, None, None, None, None, None
        q = points_query.detach().contiguous()
        r = points_ref.detach().contiguous()
        g = grad.contiguous()
        inds = inds.contiguous()
        dq = torch.empty_like(q) if need_q else None
        dr = torch.empty_like(r) if need_r else None
        if ctx.kdtree is not None and sidx is not None:
            # the points gathered in the tree's structured order: spatially coherent reads
            ctx.kdtree.backward(q.data_ptr(), r.data_ptr(), q.shape[0], inds.shape[1], sidx.contiguous().data_ptr(),
                                g.data_ptr(), dq.data_ptr() if need_q else 0, dr.data_ptr() if need_r else 0,
                                _stream_handle(q.device))
        else:
            torch_knn.knn_backward(
                q.data_ptr(), r.data_ptr(), q.shape[0], r.shape[0], q.shape[1], _dtype_name(q.dtype), inds.shape[1],
                inds.data_ptr(), g.data_ptr(), dq.data_ptr() if need_q else 0, dr.data_ptr() if need_r else 0,
                _stream_handle(q.device))
        return dq, dr, None, None, None, None


class TorchKDTree:
    def __init__(self, points_ref: torch.Tensor, device: torch.device, levels: int, squared_distances: bool,
                 tie_break: str = "reference", presort="auto", _kdtree=None, builder: str = "cpu", variant=None):
        """Builds the KDTree. See :ref:`build_kd_tree` for more details.  (``_kdtree``: an already
        built ``torch_knn`` tree of ``points_ref``, see ``load_kd_tree``.)"""
        assert device.type in ['cpu', 'cuda']
        assert points_ref.shape[0] < 2**31, "Only 32 bit signed indexing implemented"

        self.dtype = points_ref.dtype
        self.dims = points_ref.shape[-1]
        self.nr_ref_points = points_ref.shape[0]
        precision = {torch.float32: "F", torch.float64: ""}.get(self.dtype)
        kdtree_str = ("KDTree" + _transl_torch_device[device.type] + "%dD" % (self.dims) + precision
                      if precision is not None else None)

        try:
            if kdtree_str is None:
                raise AttributeError("unsupported dtype %s" % self.dtype)
            tree_cls = getattr(torch_knn, kdtree_str)
        except AttributeError as err:
            raise RuntimeError("Could not find the KD-Tree for your specified options. This probably means the library "
                               "was not compiled for the specified dimensionality, precision or the targeted device. "
                               "Original error:", err)

        kwargs = {}
        if device.type == "cuda":
            kwargs["device"] = device.index if device.index is not None else torch.cuda.current_device()
            device = torch.device("cuda", kwargs["device"])
        if _kdtree is not None:
            if type(_kdtree).__name__ != kdtree_str:
                raise RuntimeError("loaded tree is a %s, expected %s" % (type(_kdtree).__name__, kdtree_str))
            self.kdtree = _kdtree
        elif builder == "gpu":
            if device.type != "cuda":
                raise RuntimeError("builder='gpu' needs a GPU device")
            pts = points_ref.detach().to(device).contiguous()
            self.kdtree = torch_knn.build_tree_gpu(pts.data_ptr(), pts.shape[0], self.dims, _dtype_name(self.dtype),
                                                   levels, kwargs["device"], _stream_handle(device))
        elif builder == "cpu":
            self.kdtree = tree_cls(points_ref.detach().cpu().numpy(), levels, **kwargs)
        else:
            raise ValueError("builder must be 'cpu' or 'gpu'")
        self.levels = levels

        self.structured_points = torch.from_numpy(self.kdtree.get_structured_points())
        self.shuffled_ind = torch.from_numpy(self.kdtree.get_shuffled_inds()).long()
        self.use_gpu = use_gpu = (device.type == 'cuda')
        self.device = device
        self.dtype_idx = torch.int32  # Restriction in the compiled library
        self.ref_requires_grad = points_ref.requires_grad
        self.points_ref_bak = points_ref
        self.squared_distances = squared_distances
        if tie_break not in _TIE_POLICIES:
            raise ValueError("tie_break must be one of %s" % sorted(_TIE_POLICIES))
        self.tie_break = tie_break
        if presort not in _PRESORT:
            raise ValueError("presort must be 'auto', 'on' ('morton'), 'leaf' or 'off'")
        self.presort = presort
        # launch variant of GPU queries: None = the autotuner's choice for each k (autotune.variant_for)
        self.variant = variant
        self._applied_variant = 0

        if self.use_gpu:
            self.structured_points = self.structured_points.to(self.device)
            self.shuffled_ind = self.shuffled_ind.to(self.device)

    def save(self, path: str) -> None:
        """Writes the tree (structured points, indices, nodes and leaves) to ``path``; ``load_kd_tree``
        reads it back without rebuilding."""
        self.kdtree.save(str(path))

    def query(self, points_query: torch.Tensor, nr_nns_searches: int = 1,
              result_dists: torch.Tensor = None, result_idx: torch.Tensor = None):
        """Searches the specified KD-Tree for KNN of the given points

        Parameters
        ----------
        points_query : torch.Tensor, float32 or float64, shape (Q, d)
            The query points.
        nr_nns_searches : int, optional
            k, the number of neighbours returned per query (default 1).
        result_dists : torch.Tensor, optional
            Caller-owned (Q, k) buffer of the tree's dtype for the squared distances; allocated when
            omitted.
        result_idx : torch.Tensor, optional
            Caller-owned (Q, k) int32 buffer for the neighbour positions in the tree's reordered
            point array (what the reference's C++ layer writes); allocated when omitted.

        Returns
        -------
        tuple
            * dists (tensor of the tree's dtype, (Q, k)) : squared distances (distances when the tree
              was built with ``squared_distances=False``), ascending per row
            * inds (tensor of int64, (Q, k)) : rows of ``points_ref`` of the neighbours

        Raises
        ------
        RuntimeError
            If more neighbours are requested than the tree holds, or the query fails.
        """
        if nr_nns_searches > self.nr_ref_points:
            raise RuntimeError("You requested more nearest neighbors than there are in the KD-Tree")

        points_query = points_query.to(self.device)
        nq = points_query.shape[0]

        if result_dists is None:
            result_dists = torch.empty(size=[nq, nr_nns_searches], dtype=self.dtype, device=self.device)
        if result_idx is None:
            result_idx = torch.empty(size=[nq, nr_nns_searches], dtype=self.dtype_idx, device=self.device)

        assert list(result_dists.shape) == [nq, nr_nns_searches]
        assert result_dists.dtype == self.dtype
        assert list(result_idx.shape) == [nq, nr_nns_searches]
        assert result_idx.dtype == self.dtype_idx
        assert points_query.dtype == self.dtype

        user_dists, user_idx = result_dists, result_idx
        if not result_dists.is_contiguous():
            result_dists = result_dists.contiguous()
        if not result_idx.is_contiguous():
            result_idx = result_idx.contiguous()
        if not points_query.is_contiguous():
            points_query = points_query.contiguous()

        inds = torch.empty(size=[nq, nr_nns_searches], dtype=torch.int64, device=self.device)
        if self.use_gpu:
            v = self.variant if self.variant is not None else autotune.variant_for(self.dims, self.dtype,
                                                                                   nr_nns_searches, self.tie_break)
            if v != self._applied_variant:
                self.kdtree.set_variant(int(v))
                self._applied_variant = v
            self.kdtree.query(points_query.data_ptr(), nq, nr_nns_searches, result_dists.data_ptr(),
                              result_idx.data_ptr(), orig_idx_ptr=inds.data_ptr(),
                              tie=_TIE_POLICIES[self.tie_break], presort=_PRESORT[self.presort],
                              stream=_stream_handle(self.device))
        else:
            self.kdtree.query(points_query.data_ptr(), nq, nr_nns_searches, result_dists.data_ptr(),
                              result_idx.data_ptr(), orig_idx_ptr=inds.data_ptr())

        if user_dists is not result_dists:
            user_dists.copy_(result_dists)
        if user_idx is not result_idx:
            user_idx.copy_(result_idx)

        dists = result_dists
        if (points_query.requires_grad or self.ref_requires_grad) and torch.is_grad_enabled():
            points_ref = self.points_ref_bak.to(self.device)
            if self.use_gpu:
                dists = _KnnSquaredDistances.apply(points_query, points_ref, inds, result_dists, result_idx, self.kdtree)
            else:
                dists = torch.sum((points_query[:, None] - points_ref[inds])**2, dim=-1)

        if not self.squared_distances:
            dists = torch.sqrt(dists)

        return dists, inds


def build_kd_tree(points_ref: Union[torch.Tensor, np.ndarray], device: torch.device = None,
                  squared_distances=True, levels: int = None, tie_break: str = "reference", presort="auto",
                  builder: str = "cpu", variant=None):
    """Builds the KD-Tree for subsequent queries.

    By default the tree is built on the CPU (the reference's build, exactly) and then transferred
    to the GPU if necessary.

    Parameters
    ----------
    points_ref : torch.Tensor or np.ndarray
        Points from which to build the KD-Tree, shape (N, d), d in {1, 2, 3, 8}
    device : torch.device
        Target torch device of the KD-Tree; ``points_ref.device`` if not specified.
    squared_distances : bool
        If true, the squared euclidean distances will be returned, by default True.
    levels : int, "auto" or None
        Levels of the KD-Tree (1 to 24).  None picks the reference's default; "auto" picks a leaf
        size tuned for the GPU path.
    tie_break : str
        "reference" (default): exact distance ties in the reference CPU search's order.
        "lower_index": ties ordered by the lower original index.
    presort : "auto", "on"/"morton", "leaf" or "off"
        How GPU queries are ordered before the traversal (along a space-filling curve over their
        positions, Hilbert for d <= 3 and Morton for d = 8, or by tree leaf); the order changes
        speed only, never results.
    builder : "cpu" (default) or "gpu"
        "gpu" builds a GPU tree on the device.  It gives the same tree as the CPU build whenever no
        node has two points with equal coordinates on its split axis; points with such ties may be
        ordered and split differently (the reference orders them by an unstable sort), so
        structured indices can differ while the neighbour distances do not.
    variant : None (default) or int
        Launch variant of the GPU traversal (``torch_knn.tree_variants()``): None takes the one
        ``autotune.tune`` stored for this dims, dtype and k (0, the built-in shape, without one).
        Results never depend on it.

    Returns
    -------
    TorchKDTree
    """
    if issubclass(type(points_ref), np.ndarray):
        points_ref = torch.from_numpy(points_ref)

    if device is None:
        device = points_ref.device

    if issubclass(type(device), str):
        device = torch.device(device)

    if levels is None:
        levels = default_levels(points_ref.shape[0])
    elif isinstance(levels, str):
        if levels != "auto":
            raise ValueError("levels must be an int, None or 'auto'")
        levels = autotune.pick_levels(int(points_ref.shape[0]), int(points_ref.shape[-1]), device)

    assert levels >= 1 and levels <= MAX_LEVELS
    assert issubclass(type(points_ref), torch.Tensor)
    assert device.type != 'cuda' or gpu_available, \
        "You requested the KD-Tree on the GPU, but the library was compiled with CPU support only"
    assert device.type in ['cuda', 'cpu']

    return TorchKDTree(points_ref, device, levels, squared_distances, tie_break=tie_break, presort=presort,
                       builder=builder, variant=variant)


def load_kd_tree(path: str, device: torch.device = None, squared_distances=True, tie_break: str = "reference",
                 presort="auto") -> TorchKDTree:
    """Loads a tree written by ``TorchKDTree.save`` onto ``device`` (default: the CPU), without
    rebuilding it.  The reference points, in their original order, are recovered from the file."""
    if device is None:
        device = torch.device("cpu")
    if issubclass(type(device), str):
        device = torch.device(device)
    assert device.type in ['cuda', 'cpu']
    gpu = device.type == 'cuda'
    dev_index = (device.index if device.index is not None else torch.cuda.current_device()) if gpu else -1
    kd = torch_knn.load_tree(str(path), gpu, dev_index)
    structured = torch.from_numpy(kd.get_structured_points())
    shuffled = torch.from_numpy(kd.get_shuffled_inds()).long()
    points = torch.empty_like(structured)
    points[shuffled] = structured
    return TorchKDTree(points, device, kd.levels, squared_distances, tie_break=tie_break, presort=presort, _kdtree=kd)
