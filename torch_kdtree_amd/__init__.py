"""torch_kdtree_amd — KD-tree KNN for AMD Instinct MI355X (gfx950).

Drop-in for the query path of thomgrand/torch_kdtree (``torch_kdtree/__init__.py``):
``build_kd_tree(points).query(q, nr_nns_searches=k)``.
"""
from .nn_distance import build_kd_tree, gpu_available, load_kd_tree, TorchKDTree
from .batch import build_kd_tree_batch, KDTreeBatch, BatchStep

__all__ = ["build_kd_tree", "gpu_available", "load_kd_tree", "TorchKDTree", "build_kd_tree_batch", "KDTreeBatch", "BatchStep"]
__version__ = "1.0"
