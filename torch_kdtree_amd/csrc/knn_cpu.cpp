// CPU KNN search for trees built on the host device ("cpu" trees, the reference's
// KDTreeCPU{d}D[F] classes, src/interface.cpp:60-77 -> src/kdtree.cpp:164-194).
//
// Same semantics as the reference's iterative search (src/kdtree.cpp:16-110):
//   * near-first depth-first traversal; the near child of a node with split (a, m) is the
//     left one iff q[a] < m;
//   * a projected point pp starts at q.  When the far child of an internal node is tested,
//     pp[a] = m first; the far child is entered iff bound(q, pp, a, m) < worst, and pp[a]
//     goes back to q[a] when the node is left.  At a leaf-parent the near leaf is scanned,
//     then the far leaf iff bound(q, pp, a, m) < worst with pp as it stands;
//   * bound = sum_i v_i^2 with v = pp - q and v[a] += m - pp[a] (src/kdtree.hpp:606-614);
//   * candidates are scanned in structured order, rejected when d >= worst, otherwise
//     inserted at the upper-bound position (src/nndistance.hpp:41-85).  A NaN distance fails
//     both comparisons: it is taken, and placed by the binary search's path through the list.
// The reference keeps a per-node tag array; here the state is the node index and the
// child we came back from, which is all the tag array encodes on the active path.
//
// A query with a NaN coordinate forms only NaN distances and bounds.  The insertion above never
// rejects a NaN (d >= worst is false) and puts it in front (d >= bd[pos] is false), and every
// bound test against worst fails, so the reference scans exactly the leaf its descent reaches (a
// NaN coordinate never compares below a split) and nothing else: its list is that leaf's last k
// records in reverse order, with their NaN distances, then (+inf, 0) entries when the leaf holds
// fewer than k points.  nan_query() writes that result directly.
//
// Extra pruning that keeps the reference's output (trees whose coordinates and split values are
// all finite; on other trees a NaN distance can enter the list, a box says nothing about it, and
// the search runs exactly the reference's steps): besides the reference's bound, a leaf or a
// far side is tested against the distance from q to the bounding box of its points.  When
// that box distance is >= worst, every point there has d >= worst and the reference would
// reject all of them; its visit would then change nothing but pp, where it leaves pp[a] = q[a]
// for each axis a of the levels it walks (it always walks the near path down to the
// leaf-parent level).  Skipping and applying exactly those resets keeps every later decision
// identical to the reference's.
#include <atomic>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include <omp.h>

#include "tree.hpp"

namespace tk {
namespace {

// Search statistics (leaves and points scanned), counted per thread when asked for.
thread_local int64_t tl_leaves = 0, tl_points = 0;

template <typename T, int D>
inline T sq_dist(const T* q, const T* r) {
    T s = T(0);
    for (int i = 0; i < D; ++i) {
        const T v = q[i] - r[i];
        s = s + v * v;
    }
    return s;
}

template <typename T, int D>
inline T proj_bound(const T* q, const T* pp, int a, T m) {
    T s = T(0);
    for (int i = 0; i < D; ++i) {
        T v = pp[i] - q[i];
        if (i == a) v = v + (m - pp[i]);
        s = s + v * v;
    }
    return s;
}

// Squared distance from q to a point bounding box (lo[D], hi[D]): per axis the gap to the box,
// one rounded difference, squared and summed in axis order like a distance, so by the
// monotonicity of rounding it never exceeds the computed distance of a point in the box.
template <typename T, int D>
inline T aabb_bound(const T* q, const T* box) {
    T s = T(0);
    for (int i = 0; i < D; ++i) {
        T v = T(0);
        if (q[i] < box[i])
            v = box[i] - q[i];
        else if (q[i] > box[D + i])
            v = q[i] - box[D + i];
        s = s + v * v;
    }
    return s;
}

// Bits l < 32 with l mod D == a.
template <int D>
inline uint32_t axis_levels(int a) {
    uint32_t m = 0;
    for (int l = a; l < 32; l += D) m |= 1u << l;
    return m;
}

template <typename T, int D>
struct Searcher {
    const Tree& t;
    const T* split;
    const T* pts;
    const T* aabb;
    int k;
    bool boxes;  // point-box pruning (t.finite)

    void scan(int32_t leaf, const T* q, T* bd, int32_t* bi) const {
        const int32_t beg = t.leaf_off[leaf], end = t.leaf_off[leaf + 1];
        ++tl_leaves;
        tl_points += end - beg;
        for (int32_t s = beg; s < end; ++s) {
            const T d = sq_dist<T, D>(q, pts + (int64_t)s * D);
            if (d >= bd[k - 1]) continue;
            int lo = 0, cnt = k;  // upper bound: first position with bd[pos] > d
            while (cnt > 0) {
                const int step = cnt / 2;
                if (d >= bd[lo + step]) {
                    lo += step + 1;
                    cnt -= step + 1;
                } else {
                    cnt = step;
                }
            }
            std::memmove(bd + lo + 1, bd + lo, sizeof(T) * (size_t)(k - 1 - lo));
            std::memmove(bi + lo + 1, bi + lo, sizeof(int32_t) * (size_t)(k - 1 - lo));
            bd[lo] = d;
            bi[lo] = s;
        }
    }

    const T* box(int32_t heap_node) const { return aabb + (size_t)heap_node * 2 * D; }

    static bool has_nan(const T* q) {
        for (int i = 0; i < D; ++i)
            if (q[i] != q[i]) return true;
        return false;
    }

    void nan_query(const T* q, T* bd, int32_t* bi) const {
        int32_t node = 0;
        for (int level = 0; level < t.levels; ++level) node = 2 * node + (q[level % D] < split[node] ? 1 : 2);
        const int32_t leaf = node - t.n_nodes;
        const int32_t beg = t.leaf_off[leaf], end = t.leaf_off[leaf + 1];
        for (int j = 0; j < k && end - 1 - j >= beg; ++j) {
            bd[j] = sq_dist<T, D>(q, pts + (int64_t)(end - 1 - j) * D);
            bi[j] = end - 1 - j;
        }
    }

#ifdef TK_DEBUG_CHECKS
    // The reference's recursive search (compQuadrDistNode, src/kdtree.cpp:112-152), for the debug
    // cross-check of src/kdtree.cpp:182-189: no point boxes, and the far side of a node searched
    // with pp = q except on that node's axis.
    void recursive(const T* q, int32_t node, int cl, const T* pp, T* bd, int32_t* bi) const {
        const int a = t.axis[node];
        const T m = split[node];
        const bool lower = q[a] < m;
        if (cl == 0) {
            const int32_t left_leaf = 2 * node + 1 - t.n_nodes;
            scan(lower ? left_leaf : left_leaf + 1, q, bd, bi);
            if (proj_bound<T, D>(q, pp, a, m) < bd[k - 1]) scan(lower ? left_leaf + 1 : left_leaf, q, bd, bi);
            return;
        }
        recursive(q, lower ? 2 * node + 1 : 2 * node + 2, cl - 1, pp, bd, bi);
        if (proj_bound<T, D>(q, pp, a, m) < bd[k - 1]) {
            T pp2[D];
            for (int i = 0; i < D; ++i) pp2[i] = q[i];
            pp2[a] = m;
            recursive(q, lower ? 2 * node + 2 : 2 * node + 1, cl - 1, pp2, bd, bi);
        }
    }
#endif

    void one(const T* q, T* bd, int32_t* bi) const {
        if (has_nan(q)) {
            nan_query(q, bd, bi);
            return;
        }
        T pp[D];
        for (int i = 0; i < D; ++i) pp[i] = q[i];
        const int L = t.levels;
        const int32_t nn = t.n_nodes;
        uint32_t far_mask = 0;  // bit l: the path goes through the far child of its level-l node
        int32_t node = 0;
        int level = 0;
        auto descend = [&]() {
            while (level < L - 1) {
                const int a = level % D;
                node = 2 * node + (q[a] < split[node] ? 1 : 2);
                ++level;
            }
        };
        auto scan_if_useful = [&](int32_t leaf) {
            if (!boxes || aabb_bound<T, D>(q, box(nn + leaf)) < bd[k - 1]) scan(leaf, q, bd, bi);
        };
        descend();
        for (;;) {
            // at a leaf-parent: near leaf, then the far leaf if the reference looks at it
            {
                const int a = level % D;
                const T m = split[node];
                const bool lower = q[a] < m;
                const int32_t left_leaf = 2 * node + 1 - nn;
                scan_if_useful(lower ? left_leaf : left_leaf + 1);
                if (proj_bound<T, D>(q, pp, a, m) < bd[k - 1]) scan_if_useful(lower ? left_leaf + 1 : left_leaf);
                if (level == 0) return;
                pp[a] = q[a];
            }
            // Climb.  Levels whose far child is on the path (far_mask) are left without a test
            // (their axis is reset in pp); levels entered through their near child get the
            // far-side test.  Bit scans jump from one such level to the next.
            bool entered = false;
            int l = level;
            for (;;) {
                const uint32_t below = (1u << l) - 1u;
                const uint32_t near_lv = below & ~far_mask;
                const int j = near_lv ? 31 - __builtin_clz(near_lv) : -1;
                const uint32_t popped = far_mask & below & ~(j >= 0 ? ((2u << j) - 1u) : 0u);
                if (popped) {
                    far_mask &= ~popped;
                    for (int ax = 0; ax < D; ++ax)
                        if (popped & axis_levels<D>(ax)) pp[ax] = q[ax];
                }
                if (j < 0) return;
                const int32_t nj = ((node + 1) >> (level - j)) - 1;
                const int a = j % D;
                const T m = split[nj];
                const bool lower = q[a] < m;
                pp[a] = m;
                if (proj_bound<T, D>(q, pp, a, m) < bd[k - 1]) {
                    const int32_t far_child = 2 * nj + (lower ? 2 : 1);
                    if (!boxes || aabb_bound<T, D>(q, box(far_child)) < bd[k - 1]) {
                        far_mask |= 1u << j;
                        node = far_child;
                        level = j + 1;
                        entered = true;
                        break;
                    }
                    // the reference enters and rejects every point; replay its pp resets
                    for (int lv = j + 1; lv < L && lv <= j + D; ++lv) pp[lv % D] = q[lv % D];
                }
                pp[a] = q[a];
                if (j == 0) return;
                l = j;
            }
            if (!entered) return;
            descend();
        }
    }
};

template <typename T, int D>
void query_typed(const Tree& t, const T* q, int64_t nq, int k, T* dist, int32_t* idx, int64_t* orig,
                 int n_threads, int64_t* stats) {
    Searcher<T, D> s{t, t.split_ptr<T>(), t.points_ptr<T>(), t.aabb_ptr<T>(), k, t.finite};
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#ifdef TK_DEBUG_CHECKS
    // Trees with a non-finite value are left out of the cross-check: there the two searches' different
    // projected points can legitimately differ (a point with a NaN coordinate behind a bound one prunes).
    std::atomic<int64_t> mismatch{-1};
#endif
    std::atomic<int64_t> leaves{0}, points{0};
#pragma omp parallel num_threads(n_threads)
    {
        tl_leaves = tl_points = 0;
        std::vector<T> bd(k);
        std::vector<int32_t> bi(k);
#ifdef TK_DEBUG_CHECKS
        std::vector<T> bd2(k);
        std::vector<int32_t> bi2(k);
#endif
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < nq; ++i) {
            std::fill(bd.begin(), bd.end(), std::numeric_limits<T>::infinity());
            std::fill(bi.begin(), bi.end(), 0);
            s.one(q + i * D, bd.data(), bi.data());
#ifdef TK_DEBUG_CHECKS
            if (t.finite) {
                std::fill(bd2.begin(), bd2.end(), std::numeric_limits<T>::infinity());
                std::fill(bi2.begin(), bi2.end(), 0);
                s.recursive(q + i * D, 0, t.levels - 1, q + i * D, bd2.data(), bi2.data());
                if (std::memcmp(bd.data(), bd2.data(), sizeof(T) * (size_t)k) != 0 ||
                    std::memcmp(bi.data(), bi2.data(), sizeof(int32_t) * (size_t)k) != 0)
                    mismatch = i;
            }
#endif
            std::memcpy(dist + i * k, bd.data(), sizeof(T) * (size_t)k);
            std::memcpy(idx + i * k, bi.data(), sizeof(int32_t) * (size_t)k);
            if (orig)
                for (int j = 0; j < k; ++j) orig[i * k + j] = t.shuffled[bi[j]];
        }
        leaves += tl_leaves;
        points += tl_points;
    }
    if (stats) {
        stats[0] = leaves.load();
        stats[1] = points.load();
    }
#ifdef TK_DEBUG_CHECKS
    if (mismatch >= 0)
        throw std::runtime_error("debug check: the iterative and the recursive search differ at query " +
                                 std::to_string((long long)mismatch.load()));
#endif
}

}  // namespace

void query_cpu(const Tree& t, const void* queries, int64_t n_queries, int k, void* dist, int32_t* idx,
               int64_t* orig_idx, int n_threads, int64_t* stats) {
    if (k < 1) throw std::runtime_error("nr_nns_searches must be >= 1");
    if (k > t.n_points) throw std::runtime_error("You requested more nearest neighbors than there are in the KD-Tree");
#define TK_Q(TT, DD) \
    query_typed<TT, DD>(t, static_cast<const TT*>(queries), n_queries, k, static_cast<TT*>(dist), idx, orig_idx, n_threads, \
                        stats)
    if (t.dtype == kF32) {
        switch (t.dims) {
            case 1: TK_Q(float, 1); break;
            case 2: TK_Q(float, 2); break;
            case 3: TK_Q(float, 3); break;
            case 8: TK_Q(float, 8); break;
        }
    } else {
        switch (t.dims) {
            case 1: TK_Q(double, 1); break;
            case 2: TK_Q(double, 2); break;
            case 3: TK_Q(double, 3); break;
            case 8: TK_Q(double, 8); break;
        }
    }
#undef TK_Q
}

}  // namespace tk
