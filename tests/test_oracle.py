"""Pins of the oracle (CPU restatement of the reference path) — no GPU needed.

Each oracle function is checked against something that does not share its code:
* known-answer vectors traced by hand through the reference algorithm (tests/golden/);
* a numpy float32 brute force with explicit per-step rounding (tests/golden/bruteforce_f32.npz);
* the oracle's own brute force on random data (the KNN set is unique up to exact ties);
* the reference's own debug cross-check, iterative == recursive (src/kdtree.cpp:182-189);
* tree-layout invariants of createKDTree (src/kdtree.hpp:418-534);
* the default-levels closed form (nn_distance.py:146-147) at hand-computed points.
A deliberately broken oracle variant is not needed: every check above fails on a dropped
term, a swapped operand or a wrong tie rule (see test_plants_are_caught).
"""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _known():
    with open(os.path.join(GOLD, "known_answers.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", _known(), ids=lambda c: c["name"])
def test_known_answers(case):
    pts = np.array(case["points"], np.float32)
    t = oracle.OracleTree(pts, case["levels"])
    e = t.export()
    tr = case["tree"]
    assert np.array_equal(e["structured"], np.array(tr["structured"], np.float32))
    assert e["shuffled"].tolist() == tr["shuffled"]
    assert e["axis"].tolist() == tr["axis"]
    assert e["median"].tolist() == tr["median"]
    assert np.append(e["leaf_off"], len(pts)).tolist() == tr["leaf_off"]
    for q in case["queries"]:
        d, i = t.search(np.array([q["q"]], np.float32), q["k"])
        assert d[0].tolist() == q["dist"]
        assert i[0].tolist() == q["idx"]
        assert e["shuffled"][i[0]].tolist() == q["orig"]
        bd, bi = oracle.bruteforce(pts, np.array([q["q"]], np.float32), q["k"])
        assert bi[0].tolist() == q["orig_lower"]
        assert bd[0].tolist() == q["dist"]


@pytest.mark.parametrize("d", [2, 3])
def test_bruteforce_matches_numpy_fixture(d):
    z = np.load(os.path.join(GOLD, "bruteforce_f32.npz"))
    bd, bi = oracle.bruteforce(z["ref_d%d" % d], z["qry_d%d" % d], 7)
    assert np.array_equal(bd, z["dist_d%d" % d])
    assert np.array_equal(bi, z["idx_d%d" % d])


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("levels", [1, 3, 6])
def test_tree_search_on_fixture_ties(d, levels):
    """Reference search on the tie-heavy fixture: distances equal brute force everywhere, and
    the index multiset equals brute force outside the last tie group of each row."""
    z = np.load(os.path.join(GOLD, "bruteforce_f32.npz"))
    ref, qry = z["ref_d%d" % d], z["qry_d%d" % d]
    t = oracle.OracleTree(ref, levels)
    dist, idx = t.search(qry, 7)
    assert np.array_equal(dist, z["dist_d%d" % d])
    orig = t.export()["shuffled"][idx]
    for row in range(qry.shape[0]):
        last = dist[row, -1]
        inner = dist[row] < last
        assert sorted(orig[row][inner]) == sorted(z["idx_d%d" % d][row][inner])


def _numpy_sq_dist(q, r):
    s = np.zeros(q.shape[0], q.dtype)
    for a in range(q.shape[1]):
        v = (q[:, a] - r[:, a]).astype(q.dtype)
        s = (s + (v * v).astype(q.dtype)).astype(q.dtype)
    return s


def test_thread_count_argument_leaves_the_process_default():
    """A search or brute force with an explicit thread count runs its own team of that size and
    leaves the process's OpenMP default as it was (the library's multi-tree host build takes that
    default, and ran on one thread after a one-thread oracle run when it did not)."""
    import ctypes
    gomp = ctypes.CDLL("libgomp.so.1")
    before = gomp.omp_get_max_threads()
    rng = np.random.default_rng(5)
    ref = rng.random((2000, 3)).astype(np.float32)
    t = oracle.OracleTree(ref, 6)
    t.search(ref[:100], 4, nthreads=1)
    oracle.bruteforce(ref, ref[:100], 4, nthreads=1)
    assert gomp.omp_get_max_threads() == before


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
def test_search_against_bruteforce_and_recursive(dtype, d):
    rng = np.random.default_rng(d + (10 if dtype == np.float64 else 0))
    ref = (rng.standard_normal((4000, d)) * 1e3).astype(dtype)
    qry = (rng.standard_normal((600, d)) * 1e3).astype(dtype)
    levels = oracle.default_levels(4000)
    t = oracle.OracleTree(ref, levels)
    e = t.export()
    for k in (1, 6, 17):
        dist, idx = t.search(qry, k)
        rd, ri = t.search(qry, k, recursive=True)
        assert np.array_equal(dist, rd) and np.array_equal(idx, ri)  # src/kdtree.cpp:182-189
        bd, bi = oracle.bruteforce(ref, qry, k)
        assert np.array_equal(dist, bd)
        assert np.array_equal(e["shuffled"][idx], bi)  # continuous data: no exact ties
        for j in range(k):  # distances are the reference's arithmetic, restated in numpy
            assert np.array_equal(dist[:, j], _numpy_sq_dist(qry, e["structured"][idx[:, j]]))


def _nan_query_closed_form(e, q, k):
    """The reference's list for a query with a NaN coordinate, from its comparisons alone: the
    insertion (src/nndistance.hpp:77-85) rejects only when d >= worst and places a candidate at the
    upper bound found with `value >= *it`, so a NaN distance is always taken and put in front; the
    bound tests `< worst` all fail, so only the leaf the descent reaches (a NaN coordinate never
    compares below a split) is scanned."""
    axis, split, leaf_off, leaf_n = e["axis"], e["median"], e["leaf_off"], e["leaf_n"]
    n_nodes = len(split)
    node = 0
    while node < n_nodes:
        node = 2 * node + (1 if q[axis[node]] < split[node] else 2)
    leaf = node - n_nodes
    dist = [np.inf] * k
    idx = [0] * k
    for s in range(leaf_off[leaf], leaf_off[leaf] + leaf_n[leaf]):
        dist = [np.nan] + dist[:-1]
        idx = [s] + idx[:-1]
    return np.array(dist), np.array(idx)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 3, 8])
def test_nan_queries_follow_the_insertion_rule(dtype, d):
    rng = np.random.default_rng(40 + d)
    for n, levels in ((40, 3), (3000, 7), (9, 1)):
        ref = (rng.standard_normal((n, d)) * 10).astype(dtype)
        t = oracle.OracleTree(ref, levels)
        e = t.export()
        qry = (rng.standard_normal((150, d)) * 10).astype(dtype)
        for i in range(len(qry)):
            m = rng.random(d) < 0.5
            m[i % d] = True
            qry[i, m] = np.nan
        for k in sorted({1, 3, min(7, n), min(60, n)}):
            dist, idx = t.search(qry, k)
            rd, ri = t.search(qry, k, recursive=True)
            assert np.array_equal(dist, rd, equal_nan=True) and np.array_equal(idx, ri)
            for i in range(len(qry)):
                cd, ci = _nan_query_closed_form(e, qry[i], k)
                assert np.array_equal(dist[i], cd, equal_nan=True) and np.array_equal(idx[i], ci), (n, k, i)
            # brute force inserts only when d < worst (src/nndistance_g.cu:57): nothing for NaN
            bd, bi = oracle.bruteforce(ref, qry, k)
            assert np.isposinf(bd).all() and (bi == 0).all()


def _tree_1level(structured, median, n_left, shuffled=None):
    """A one-level d=1 tree given as arrays: node 0 splits at `median`, leaf 0 holds the first
    n_left structured points, leaf 1 the rest."""
    pts = np.asarray(structured, np.float32).reshape(-1, 1)
    n = pts.shape[0]
    sh = np.arange(n, dtype=np.int32) if shuffled is None else np.asarray(shuffled, np.int32)
    return oracle.OracleTree.from_arrays(pts, sh, np.array([0], np.int32), np.array([median], np.float32),
                                         np.array([n_left, n - n_left], np.int32),
                                         np.array([0, n_left], np.int32), 1)


def test_nan_distance_insertion_known_answer():
    """The reference's insertion with NaN distances, traced by hand (src/nndistance.hpp:41-85:
    reject iff d >= worst, else insert at upperBound, whose probes compare `value >= *it`).
    Query 0, k = 3; the near leaf's points 0.9, 1.0, NaN, 0.7, 0.8, 1.7 give the distances
    a=0.81, b=1.0, NaN, c=0.49, e=0.64, f=2.89 in scan order:
      a -> [a, inf, inf]; b -> [a, b, inf]; NaN fails every probe -> [NaN, a, b];
      c: probe 1 (a) fails, probe 0 (NaN) fails -> [c, NaN, a];
      e: e < a passes the pre-check; probe 1 (NaN) fails, probe 0 (c) passes -> [c, e, NaN];
      f: f >= NaN is false, so it is taken; probe 1 (e) passes, probe 2 (NaN) fails -> [c, e, f].
    The far leaf's bound (100 - 0)^2 is not below worst, so it is never scanned.  A brute force
    (d < worst, src/nndistance_g.cu:57) drops the NaN and returns the three smallest."""
    f32 = np.float32
    x = [0.9, 1.0, np.nan, 0.7, 0.8, 1.7]
    t = _tree_1level(x + [200.0, 300.0], 100.0, 6)
    d, i = t.search(np.zeros((1, 1), f32), 3)
    sq = [f32(v) * f32(v) for v in (0.7, 0.8, 1.7)]
    assert d[0].tolist() == [float(v) for v in sq] and i[0].tolist() == [3, 4, 5]
    # with k = 4 the NaN never reaches the last slot: [c, e, NaN, a] after e, then f >= a is rejected
    d4, i4 = t.search(np.zeros((1, 1), f32), 4)
    assert i4[0].tolist() == [3, 4, 2, 0] and np.isnan(d4[0, 2])
    assert d4[0, [0, 1, 3]].tolist() == [float(sq[0]), float(sq[1]), float(f32(0.9) * f32(0.9))]
    bd, bi = oracle.bruteforce(np.array(x + [200.0, 300.0], f32).reshape(-1, 1), np.zeros((1, 1), f32), 3)
    assert bi[0].tolist() == [3, 4, 0]


def test_nan_split_known_answer():
    """A NaN split value, traced by hand.  Points (d=1) 1, NaN, 3, 2 at one level: libstdc++'s
    std::sort on 4 elements is an insertion sort (`val < prev` moves left): NaN compares false
    both ways and stays, 3 stays behind NaN, 2 moves left past 3 and stops at NaN, so the
    structured order is 1, NaN, 2, 3 (shuffled 0, 1, 3, 2) and the median 0.5*(2 + NaN) is NaN.
    Query 0: 0 < NaN is false, so the near leaf is the right one (2, 3): [4, 9, inf] with
    structured indices [2, 3, 0]; the far leaf's bound ((0-0) + (NaN - 0))^2 is NaN, not below
    worst, so the point at distance 1 is never seen."""
    pts = np.array([[1.0], [np.nan], [3.0], [2.0]], np.float32)
    t = oracle.OracleTree(pts, 1)
    e = t.export()
    assert e["shuffled"].tolist() == [0, 1, 3, 2] and np.isnan(e["median"][0])
    d, i = t.search(np.zeros((1, 1), np.float32), 3)
    assert d[0].tolist() == [4.0, 9.0, float("inf")] and i[0].tolist() == [2, 3, 0]


def test_infinite_query_known_answer():
    """A query at +inf against finite points: every distance is +inf (rejected, inf >= inf) and
    every bound NaN (inf - inf); the list stays (+inf, 0).  The same from brute force."""
    rng = np.random.default_rng(5)
    pts = rng.standard_normal((300, 3)).astype(np.float32)
    q = np.array([[np.inf, 0.0, 0.0], [0.0, -np.inf, 1.0]], np.float32)
    t = oracle.OracleTree(pts, 4)
    d, i = t.search(q, 5)
    assert np.isposinf(d).all() and (i == 0).all()
    bd, bi = oracle.bruteforce(pts, q, 5)
    assert np.isposinf(bd).all() and (bi == 0).all()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 8])
@pytest.mark.parametrize("n,levels", [(1000, 6), (777, 5), (64, 7), (5, 1), (3001, 10)])
def test_tree_layout_invariants(dtype, d, n, levels):
    rng = np.random.default_rng(n + d)
    pts = rng.integers(0, 7, (n, d)).astype(dtype)  # duplicates on every axis
    t = oracle.OracleTree(pts, levels)
    e = t.export()
    assert t.n_nodes == 2 ** levels - 1 and t.n_leaves == 2 ** levels
    assert sorted(e["shuffled"].tolist()) == list(range(n))
    assert np.array_equal(e["structured"], pts[e["shuffled"]])
    assert np.array_equal(np.cumsum(np.concatenate([[0], e["leaf_n"][:-1]])), e["leaf_off"])
    assert e["leaf_n"].sum() == n

    def check(node, lo, hi, depth):
        a = e["axis"][node]
        assert a == depth % d
        cnt = hi - lo
        m = e["median"][node]
        seg = np.sort(e["structured"][lo:hi, a])  # children re-sort their ranges on other axes
        half = cnt // 2
        expect = seg[half] if cnt % 2 else dtype(0.5 * (seg[half] + seg[half - 1]))
        assert m == expect
        assert np.all(e["structured"][lo:lo + half, a] <= m) and np.all(e["structured"][lo + half:hi, a] >= m)
        if depth == levels - 1:
            leaf = 2 * node + 1 - t.n_nodes
            assert e["leaf_off"][leaf] == lo and e["leaf_n"][leaf] == half and e["leaf_n"][leaf + 1] == cnt - half
            return
        check(2 * node + 1, lo, lo + half, depth + 1)
        check(2 * node + 2, lo + half, hi, depth + 1)

    check(0, 0, n, 0)


def test_ran_out_of_points():
    pts = np.zeros((63, 3), np.float32)
    with pytest.raises(RuntimeError, match="Ran out of points"):
        oracle.OracleTree(pts, 7)  # needs 2^(7-1) = 64 points
    oracle.OracleTree(np.zeros((64, 3), np.float32), 7)


def test_default_levels_closed_form():
    # int(log(N)/log(2)) - 3 clamped to [1, 13], evaluated by hand
    table = {5: 1, 10: 1, 20: 1, 30: 1, 64: 3, 1000: 6, 1024: 7, 10000: 10, 100000: 13, 4_000_000: 13}
    for n, lv in table.items():
        assert oracle.default_levels(n) == lv, n


def test_plants_are_caught():
    """Mistakes a restatement could make are visible to the checks above."""
    z = np.load(os.path.join(GOLD, "bruteforce_f32.npz"))
    ref, qry = z["ref_d3"], z["qry_d3"]
    good = z["dist_d3"]
    # r - q + square of the sum instead of the sum of squares
    wrong = np.array([[np.float32(np.sum(q - r)) ** 2 for r in ref[:7]] for q in qry[:3]])
    assert not np.array_equal(np.sort(wrong, axis=1), good[:3])
    # lower-index vs reference tie order differ on the hand-traced case
    case = _known()[1]["queries"][2]
    assert case["orig"] != case["orig_lower"]
