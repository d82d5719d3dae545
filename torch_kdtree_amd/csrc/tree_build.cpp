// Host tree build.  Produces the reference's layout bit for bit (createKDTree,
// src/kdtree.hpp:418-534): median split on axis depth mod dims, lower half n/2 points,
// upper half n - n/2, median = P[n/2] for odd n and 0.5*(P[n/2]+P[n/2-1]) for even n.
//
// The reference sorts twice per node, an index array and then the points, each with
// std::sort keyed on the split coordinate (src/kdtree.hpp:435-448).  std::sort's element
// moves depend only on the outcomes of its comparisons, and both sorts see the same key
// sequence, so both apply one and the same permutation, including the order of equal keys.
// This build therefore sorts a single array of (key, position) pairs with the same
// comparator (operator< on the key) and applies that permutation to points and indices.
// The result, equal keys included, is the reference's; tests compare it with the oracle's
// faithful two-sort restatement on inputs full of duplicate coordinates.
//
// Subtrees are independent, so large ones are built as OpenMP tasks; the output does not
// depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <utility>

#include <omp.h>

#include "tree.hpp"

namespace tk {
namespace {

constexpr int32_t kTaskCutoff = 1 << 15;

template <typename T, int D>
struct Builder {
    Tree& t;
    T* pts;        // structured points being permuted in place
    int32_t* idx;  // shuffled indices being permuted in place
    T* split;

    void node(int32_t nid, int32_t beg, int32_t n, int lv, int a) {
        std::vector<std::pair<T, int32_t>> kv(n);
        for (int32_t i = 0; i < n; ++i) kv[i] = {pts[(int64_t)(beg + i) * D + a], i};
        std::sort(kv.begin(), kv.end(),
                  [](const std::pair<T, int32_t>& l, const std::pair<T, int32_t>& r) { return l.first < r.first; });
        {
            std::vector<T> tp((size_t)n * D);
            std::vector<int32_t> ti(n);
            for (int32_t i = 0; i < n; ++i) {
                const int64_t src = beg + kv[i].second;
                std::memcpy(&tp[(size_t)i * D], &pts[src * D], sizeof(T) * D);
                ti[i] = idx[src];
            }
            std::memcpy(&pts[(int64_t)beg * D], tp.data(), sizeof(T) * D * (size_t)n);
            std::memcpy(&idx[beg], ti.data(), sizeof(int32_t) * (size_t)n);
        }
        const int32_t h = n / 2;
        // src/kdtree.hpp:392-401, written with the same expression so that the rounding of
        // the even case (T sum, then the exact halving in double, then back to T) is equal.
        const T lo_key = pts[(int64_t)(beg + (h > 0 ? h - 1 : 0)) * D + a];
        const T hi_key = pts[(int64_t)(beg + h) * D + a];
        const T median = (n % 2 == 1) ? hi_key : static_cast<T>(0.5 * (hi_key + lo_key));
        split[nid] = median;
        t.axis[nid] = (uint8_t)a;
        const int32_t nlo = h, nhi = n - h;
        if (lv == 0) {
            const int32_t left_leaf = 2 * nid + 1 - t.n_nodes;
            t.leaf_off[left_leaf] = beg;
            t.leaf_off[left_leaf + 1] = beg + nlo;
            return;
        }
        const int na = (a + 1) % D;
        if (n > kTaskCutoff) {
#pragma omp task default(shared) firstprivate(nid, beg, nlo, lv, na)
            node(2 * nid + 1, beg, nlo, lv - 1, na);
#pragma omp task default(shared) firstprivate(nid, beg, nlo, nhi, lv, na)
            node(2 * nid + 2, beg + nlo, nhi, lv - 1, na);
#pragma omp taskwait
        } else {
            node(2 * nid + 1, beg, nlo, lv - 1, na);
            node(2 * nid + 2, beg + nlo, nhi, lv - 1, na);
        }
    }
};

// Point bounding boxes: leaves from their points, internal nodes as the union of their children.
// NaN coordinates are left out of the boxes (a NaN distance is never below a bound, and the
// searches that meet one, on trees that are not `finite`, use no boxes under the reference tie
// policy).  Also sets t.finite: every coordinate and every split value finite.
template <typename T, int D>
void compute_aabb(Tree& t, int n_threads) {
    const int32_t nn = t.n_nodes, nl = t.n_leaves;
    t.aabb.assign((size_t)(nn + nl) * 2 * D * sizeof(T), 0);
    T* box = reinterpret_cast<T*>(t.aabb.data());
    const T* pts = t.points_ptr<T>();
    const T inf = std::numeric_limits<T>::infinity();
    bool finite = true;
#pragma omp parallel for schedule(static) reduction(&& : finite) num_threads(n_threads)
    for (int32_t j = 0; j < nl; ++j) {
        T* b = box + (size_t)(nn + j) * 2 * D;
        for (int a = 0; a < D; ++a) {
            b[a] = inf;
            b[D + a] = -inf;
        }
        for (int32_t s = t.leaf_off[j]; s < t.leaf_off[j + 1]; ++s)
            for (int a = 0; a < D; ++a) {
                const T v = pts[(size_t)s * D + a];
                finite = finite && std::isfinite(v);
                if (std::isnan(v)) continue;
                b[a] = std::min(b[a], v);
                b[D + a] = std::max(b[D + a], v);
            }
    }
    for (int32_t i = nn - 1; i >= 0; --i) {
        T* b = box + (size_t)i * 2 * D;
        const T* l = box + (size_t)(2 * i + 1) * 2 * D;
        const T* r = box + (size_t)(2 * i + 2) * 2 * D;
        for (int a = 0; a < D; ++a) {
            b[a] = std::min(l[a], r[a]);
            b[D + a] = std::max(l[D + a], r[D + a]);
        }
    }
    const T* split = t.split_ptr<T>();
    for (int32_t i = 0; i < nn && finite; ++i) finite = std::isfinite(split[i]);
    t.finite = finite;
}

template <typename T, int D>
void build_typed(Tree& t, const T* src, int n_threads) {
    std::memcpy(t.points.data(), src, sizeof(T) * D * (size_t)t.n_points);
    Builder<T, D> b{t, t.points_ptr<T>(), t.shuffled.data(), t.split_ptr<T>()};
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#pragma omp single
    b.node(0, 0, t.n_points, t.levels - 1, 0);
    t.leaf_off[t.n_leaves] = t.n_points;
    compute_aabb<T, D>(t, n_threads);
}

template <typename T>
void compute_aabb_t(Tree& t) {
    const int nt = omp_get_max_threads();
    switch (t.dims) {
        case 1: compute_aabb<T, 1>(t, nt); break;
        case 2: compute_aabb<T, 2>(t, nt); break;
        case 3: compute_aabb<T, 3>(t, nt); break;
        case 8: compute_aabb<T, 8>(t, nt); break;
        default: throw std::runtime_error("unsupported dims");
    }
}

}  // namespace

void finish_tree(Tree& t) {
    t.axis.assign(t.n_nodes, 0);
    for (int32_t i = 0, lv = 0; i < t.n_nodes; ++i) {
        if (i + 1 >= (int32_t(1) << (lv + 1))) ++lv;  // node i is at depth floor(log2(i + 1))
        t.axis[i] = (uint8_t)(lv % t.dims);
    }
    if (t.dtype == kF32)
        compute_aabb_t<float>(t);
    else
        compute_aabb_t<double>(t);
}

Tree* build_tree(const void* points, int64_t n_points, int dims, int levels, int dtype, int n_threads) {
    if (!dims_supported(dims))
        throw std::runtime_error("Could not find the KD-Tree for dimensionality " + std::to_string(dims));
    if (dtype != kF32 && dtype != kF64) throw std::runtime_error("Unsupported dtype");
    if (levels < 1 || levels > kMaxLevels)
        throw std::runtime_error("levels must be in [1, " + std::to_string(kMaxLevels) + "]");
    if (n_points >= (int64_t(1) << 31)) throw std::runtime_error("Only 32 bit signed indexing implemented");
    // A node of depth l holds floor(N/2^l) or ceil(N/2^l) points; the reference throws when a
    // node at depth <= levels-1 is empty (src/kdtree.hpp:428-429), i.e. when N < 2^(levels-1).
    if (n_points < (int64_t(1) << (levels - 1)))
        throw std::runtime_error(
            "Error: Ran out of points while building KD-Tree. Either you required too many levels, or you "
            "used a lot of coplanar points");
    std::unique_ptr<Tree> t(new Tree());
    t->dtype = dtype;
    t->dims = dims;
    t->levels = levels;
    t->n_points = (int32_t)n_points;
    t->n_nodes = total_nodes(levels);
    t->n_leaves = (int32_t)(1u << levels);
    t->axis.assign(t->n_nodes, 0);
    t->split.assign((size_t)t->n_nodes * dtype_size(dtype), 0);
    t->leaf_off.assign((size_t)t->n_leaves + 1, 0);
    t->points.assign((size_t)n_points * dims * dtype_size(dtype), 0);
    t->shuffled.resize(n_points);
    for (int64_t i = 0; i < n_points; ++i) t->shuffled[i] = (int32_t)i;

#define TK_BUILD(TT, DD) build_typed<TT, DD>(*t, static_cast<const TT*>(points), n_threads)
    if (dtype == kF32) {
        switch (dims) {
            case 1: TK_BUILD(float, 1); break;
            case 2: TK_BUILD(float, 2); break;
            case 3: TK_BUILD(float, 3); break;
            case 8: TK_BUILD(float, 8); break;
        }
    } else {
        switch (dims) {
            case 1: TK_BUILD(double, 1); break;
            case 2: TK_BUILD(double, 2); break;
            case 3: TK_BUILD(double, 3); break;
            case 8: TK_BUILD(double, 8); break;
        }
    }
#undef TK_BUILD
#ifdef TK_DEBUG_CHECKS
    // The reference's sanity checks after createKDTree (src/kdtree.hpp:517-529): leaves contiguous,
    // each starting where the previous one ends, all 2^levels of them; plus the index permutation.
    if (t->leaf_off[0] != 0 || t->leaf_off[t->n_leaves] != t->n_points)
        throw std::runtime_error("debug check: leaf offsets do not span the points");
    for (int32_t j = 0; j < t->n_leaves; ++j)
        if (t->leaf_off[j + 1] < t->leaf_off[j]) throw std::runtime_error("debug check: leaf offsets decrease");
    std::vector<char> seen((size_t)n_points, 0);
    for (int64_t i = 0; i < n_points; ++i) {
        const int32_t o = t->shuffled[i];
        if (o < 0 || o >= n_points || seen[(size_t)o]) throw std::runtime_error("debug check: shuffled is no permutation");
        seen[(size_t)o] = 1;
        if (std::memcmp(&t->points[(size_t)i * dims * dtype_size(dtype)],
                        static_cast<const unsigned char*>(points) + (size_t)o * dims * dtype_size(dtype),
                        (size_t)dims * dtype_size(dtype)) != 0)
            throw std::runtime_error("debug check: structured point differs from its original row");
    }
#endif
    return t.release();
}

}  // namespace tk
