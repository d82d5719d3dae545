"""Writes the golden fixtures of this directory.

known_answers.json
    Small trees and queries whose expected outputs were traced by hand through the reference
    algorithm (src/kdtree.hpp:392-534 build, src/kdtree.cpp:16-110 search,
    src/nndistance.hpp:41-85 insertion).  The traces are stored next to the numbers.  For
    sorts of at most 16 elements libstdc++'s std::sort ends in a plain insertion sort, which
    keeps equal keys in input order; the traces use that.
bruteforce_f32.npz
    Random inputs with many duplicate coordinates and the exact KNN in (distance, index)
    order, computed here with numpy float32 scalars one rounding at a time (q - r per axis,
    squares summed left to right), independently of the C++ oracle.

Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

KNOWN = [
    {
        "name": "1d_tie_order",
        "dtype": "float32",
        "levels": 1,
        "points": [[1.0], [2.0], [0.0], [3.0]],
        "tree": {"structured": [[0.0], [1.0], [2.0], [3.0]], "shuffled": [2, 0, 1, 3], "axis": [0],
                 "median": [1.5], "leaf_off": [0, 2, 4]},
        "queries": [{"q": [1.5], "k": 2, "dist": [0.25, 0.25], "idx": [2, 1], "orig": [1, 0], "orig_lower": [0, 1]}],
        "trace": "sort x -> [0,1,2,3] from rows [2,0,1,3]; median 0.5*(2+1)=1.5. q=1.5: 1.5<1.5 false -> right leaf "
                 "first: s2 d=.25, s3 d=2.25 -> [.25(s2), 2.25(s3)]; bound 0 < 2.25 -> left leaf: s0 d=2.25 >= worst "
                 "rejected, s1 d=.25 inserted after the equal .25 -> [.25(s2), .25(s1)] -> rows [1, 0].",
    },
    {
        "name": "2d_two_levels",
        "dtype": "float32",
        "levels": 2,
        "points": [[1, 1], [3, 1], [1, 3], [3, 3], [0, 2], [4, 2], [2, 0], [2, 4]],
        "tree": {"structured": [[2, 0], [1, 1], [0, 2], [1, 3], [3, 1], [4, 2], [3, 3], [2, 4]],
                 "shuffled": [6, 0, 4, 2, 1, 5, 3, 7], "axis": [0, 1, 1], "median": [2.0, 1.5, 2.5],
                 "leaf_off": [0, 2, 4, 6, 8]},
        "queries": [
            {"q": [2, 2], "k": 4, "dist": [2, 2, 2, 2], "idx": [4, 6, 3, 1], "orig": [1, 3, 2, 0],
             "orig_lower": [0, 1, 2, 3]},
            {"q": [2, 2], "k": 6, "dist": [2, 2, 2, 2, 4, 4], "idx": [4, 6, 3, 1, 5, 7], "orig": [1, 3, 2, 0, 5, 7],
             "orig_lower": [0, 1, 2, 3, 4, 5]},
            {"q": [0.5, 0.5], "k": 2, "dist": [0.5, 2.5], "idx": [1, 0], "orig": [0, 6], "orig_lower": [0, 4]},
        ],
        "trace": "root x-sort (stable, n<=16): rows 4,0,2,6 | 7,1,3,5, median 0.5*(2+2)=2. node1 y-sort: 6,0 | 4,2, "
                 "median 0.5*(2+1)=1.5. node2 y-sort: 1,5 | 3,7, median 0.5*(3+2)=2.5. "
                 "q=(2,2),k=4: 2<2 false -> node2; 2<2.5 -> leaf2: s4 d2, s5 d4; bound (0,.5)->.25<inf -> leaf3: s6 d2 "
                 "(after s4), s7 d4 -> [2(4),2(6),4(5),4(7)]; up: pp[1]=2; root: pp[0]=2, bound 0<4 -> node1: "
                 "2<1.5 false -> leaf1: s2 d4 rejected (>= worst 4), s3 d2 -> [2(4),2(6),2(3),4(5)]; bound .25<4 -> "
                 "leaf0: s0 d4 rejected, s1 d2 -> [2(4),2(6),2(3),2(1)]. "
                 "q=(.5,.5),k=2: root .5<2 -> node1 .5<1.5 -> leaf0: s0 d2.5, s1 d.5 -> [.5(1),2.5(0)]; bound 1<2.5 -> "
                 "leaf1: s2 d2.5 rejected (tie with worst), s3 d6.5 rejected; root: pp[0]=2, bound 2.25<2.5 -> node2: "
                 "leaf2 s4 d6.5, s5 d14.5 rejected; bound 2.25+4 >= 2.5 -> done. Lower-index policy takes row 4 "
                 "(point (0,2)) instead of row 6 for the 2.5 tie.",
    },
]


def f32_sq_dist(q, r):
    s = np.float32(0.0)
    for a in range(len(q)):
        v = np.float32(q[a]) - np.float32(r[a])
        s = np.float32(s + np.float32(v * v))
    return s


def brute_force_lexicographic(ref, qry, k):
    dist = np.empty((qry.shape[0], k), np.float32)
    idx = np.empty((qry.shape[0], k), np.int32)
    for i in range(qry.shape[0]):
        cand = sorted((float(f32_sq_dist(qry[i], ref[j])), j) for j in range(ref.shape[0]))[:k]
        dist[i] = [c[0] for c in cand]
        idx[i] = [c[1] for c in cand]
    return dist, idx


def main():
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(KNOWN, f, indent=1)
    rng = np.random.default_rng(2024)
    out = {}
    for d in (2, 3):
        ref = (rng.integers(-20, 21, (300, d)) * 0.25).astype(np.float32)
        qry = (rng.integers(-20, 21, (60, d)) * 0.25 + rng.integers(0, 2, (60, d)) * 0.125).astype(np.float32)
        dist, idx = brute_force_lexicographic(ref, qry, 7)
        out.update({"ref_d%d" % d: ref, "qry_d%d" % d: qry, "dist_d%d" % d: dist, "idx_d%d" % d: idx})
    np.savez(os.path.join(HERE, "bruteforce_f32.npz"), **out)


if __name__ == "__main__":
    main()
