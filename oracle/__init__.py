"""ORACLE — test infrastructure only.

ctypes loader for ``oracle/_build/liboracle.so``, the C++ restatement of the reference
KD-tree path (thomgrand/torch_kdtree ``src/kdtree.hpp``, ``src/kdtree.cpp``,
``src/nndistance.hpp``; see the header of ``kdtree_oracle.cpp`` for the file:line map).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import this module, and only as the checker.  The product package ``torch_kdtree_amd``
never imports it.

Every function here has a pin in ``tests/test_oracle.py``: brute force, the reference's
iterative-vs-recursive equality (``src/kdtree.cpp:182-189``), hand-traced known-answer
vectors (``tests/golden/``), tree invariants and a numpy restatement of the distance.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build_library() -> None:
    """Compiles the oracle with its Makefile (g++ -O3 -fopenmp -ffp-contract=off -DNDEBUG)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build_library()
    lib = ctypes.CDLL(_LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    ci = ctypes.c_int
    lib.oracle_last_error.restype = ctypes.c_char_p
    lib.oracle_build.restype = vp
    lib.oracle_build.argtypes = [vp, i32, ci, ci, ci]
    lib.oracle_tree_from_arrays.restype = vp
    lib.oracle_tree_from_arrays.argtypes = [vp, vp, vp, vp, vp, vp, i32, ci, ci, ci]
    lib.oracle_free.argtypes = [vp]
    lib.oracle_tree_info.argtypes = [vp, vp]
    lib.oracle_tree_export.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.oracle_search.restype = ci
    lib.oracle_search.argtypes = [vp, vp, i64, ci, vp, vp, ci, ci]
    lib.oracle_bruteforce.restype = ci
    lib.oracle_bruteforce.argtypes = [vp, i32, vp, i64, ci, ci, vp, vp, ci, ci]
    _lib = lib
    return lib


def _dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float32:
        return 0
    if dt == np.float64:
        return 1
    raise TypeError(f"oracle: unsupported dtype {dt}")


def default_levels(n_points: int) -> int:
    """``nn_distance.py:146-147``: max(1, min(13, int(log(N)/log(2)) - 3))."""
    return int(np.maximum(1, np.minimum(13, int(np.log(int(n_points)) / np.log(2)) - 3)))


class OracleTree:
    """Reference tree (``createKDTree``) plus the reference CPU search (``KDTreeKNNSearch``)."""

    def __init__(self, points: np.ndarray, levels: int, _handle=None):
        lib = _load()
        pts = np.ascontiguousarray(points)
        self.dtype = pts.dtype
        self.dims = pts.shape[1]
        if _handle is None:
            h = lib.oracle_build(pts.ctypes.data, pts.shape[0], self.dims, int(levels), _dtype_code(pts.dtype))
            if not h:
                raise RuntimeError(lib.oracle_last_error().decode())
        else:
            h = _handle
        self._h = h
        info = np.zeros(4, np.int32)
        lib.oracle_tree_info(self._h, info.ctypes.data)
        self.levels, self.n_nodes, self.n_leaves, self.n_points = (int(x) for x in info)

    @classmethod
    def from_arrays(cls, structured, shuffled, axis, median, leaf_n, leaf_off, levels):
        lib = _load()
        structured = np.ascontiguousarray(structured)
        shuffled = np.ascontiguousarray(shuffled, np.int32)
        axis = np.ascontiguousarray(axis, np.int32)
        median = np.ascontiguousarray(median, structured.dtype)
        leaf_n = np.ascontiguousarray(leaf_n, np.int32)
        leaf_off = np.ascontiguousarray(leaf_off, np.int32)
        h = lib.oracle_tree_from_arrays(
            structured.ctypes.data, shuffled.ctypes.data, axis.ctypes.data, median.ctypes.data,
            leaf_n.ctypes.data, leaf_off.ctypes.data, structured.shape[0], structured.shape[1],
            int(levels), _dtype_code(structured.dtype))
        if not h:
            raise RuntimeError(lib.oracle_last_error().decode())
        obj = cls.__new__(cls)
        obj.dtype = structured.dtype
        obj.dims = structured.shape[1]
        OracleTree.__init__(obj, structured, levels, _handle=h)
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.oracle_free(h)
            self._h = None

    def export(self) -> dict:
        """structured points, shuffled indices, per-node axis/median, per-leaf size/offset."""
        st = np.empty((self.n_points, self.dims), self.dtype)
        sh = np.empty(self.n_points, np.int32)
        ax = np.empty(self.n_nodes, np.int32)
        md = np.empty(self.n_nodes, self.dtype)
        ln = np.empty(self.n_leaves, np.int32)
        lo = np.empty(self.n_leaves, np.int32)
        _load().oracle_tree_export(self._h, st.ctypes.data, sh.ctypes.data, ax.ctypes.data,
                                   md.ctypes.data, ln.ctypes.data, lo.ctypes.data)
        return dict(structured=st, shuffled=sh, axis=ax, median=md, leaf_n=ln, leaf_off=lo)

    def search(self, queries: np.ndarray, k: int, nthreads: int = 0, recursive: bool = False):
        """Squared distances [Q,k] (T) and structured int32 indices [Q,k]."""
        q = np.ascontiguousarray(queries, self.dtype)
        assert q.ndim == 2 and q.shape[1] == self.dims
        d = np.empty((q.shape[0], k), self.dtype)
        i = np.empty((q.shape[0], k), np.int32)
        rc = _load().oracle_search(self._h, q.ctypes.data, q.shape[0], int(k), d.ctypes.data, i.ctypes.data,
                                   int(nthreads), int(recursive))
        if rc:
            raise RuntimeError(_load().oracle_last_error().decode())
        return d, i


def tree_from_product(torch_kdtree) -> OracleTree:
    """An OracleTree over the arrays of a product tree (``TorchKDTree``), for running the
    reference search on exactly that tree.  Layout parity of the two builds is tested
    separately (tests/test_cpu_api.py)."""
    axis, split, leaf_off = torch_kdtree.kdtree.get_nodes()
    st = torch_kdtree.kdtree.get_structured_points()
    sh = torch_kdtree.kdtree.get_shuffled_inds()
    return OracleTree.from_arrays(st, sh, axis, split, np.diff(leaf_off), leaf_off[:-1], torch_kdtree.levels)


def bruteforce(ref: np.ndarray, queries: np.ndarray, k: int, nthreads: int = 0):
    """Exact KNN ordered by (squared distance, original index); distances as the reference computes them."""
    ref = np.ascontiguousarray(ref)
    q = np.ascontiguousarray(queries, ref.dtype)
    d = np.empty((q.shape[0], k), ref.dtype)
    i = np.empty((q.shape[0], k), np.int32)
    rc = _load().oracle_bruteforce(ref.ctypes.data, ref.shape[0], q.ctypes.data, q.shape[0], ref.shape[1], int(k),
                                   d.ctypes.data, i.ctypes.data, _dtype_code(ref.dtype), int(nthreads))
    if rc:
        raise RuntimeError(_load().oracle_last_error().decode())
    return d, i
