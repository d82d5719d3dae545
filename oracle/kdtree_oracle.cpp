// ORACLE — test infrastructure only.
//
// A CPU restatement of the reference KD-tree path of thomgrand/torch_kdtree, used as the
// checker for the HIP product path.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library; the product never does.
//
// What it restates (file:line in the reference):
//   * tree build ........ src/kdtree.hpp:392-534 (compMedian, createPartitionRecursive,
//                          createKDTree) and src/kdtree.cpp:220-229 (PartitionInfo)
//   * iterative search .. src/kdtree.cpp:16-110 (searchKDTreeIteratively), driven by the
//                          tag-array walker of src/kdtree.hpp:225-379 (TreeTraversal)
//   * recursive search .. src/kdtree.cpp:112-152 (compQuadrDistNode), the reference's own
//                          debug cross-check (src/kdtree.cpp:182-189)
//   * query driver ...... src/kdtree.cpp:164-194 (KDTreeKNNSearch): OpenMP over queries,
//                          per-query std::vector k-lists and a per-query tag array
//   * distance .......... src/kdtree.hpp:546-591 (q - r per axis, squares summed left to
//                          right in T; built with -ffp-contract=off so nothing is fused)
//   * pruning bound ..... src/kdtree.hpp:606-651 (projectionDist, partitionNecessary)
//   * leaf scan ......... src/kdtree.hpp:663-692, insertion src/nndistance.hpp:41-127
//   * brute force ....... src/nndistance_g.cu:6-72 semantics (scan in index order, insert
//                          when d < worst, :57, at the upper-bound position), i.e. ties go to
//                          the lower index and a NaN distance is never inserted; slots left
//                          empty hold (+inf, 0) (the reference leaves their index unset).
//
// Pins (see tests/test_oracle.py): brute force on every query, the reference's own
// iterative-vs-recursive equality, hand-traced known-answer vectors in tests/golden/,
// tree-layout invariants, and a numpy restatement of the distance arithmetic.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off -DNDEBUG, no -march).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>
#include <omp.h>

namespace {

thread_local std::string g_err;

enum Tag : uint8_t { kUncharted = 0, kLeftVisited = 1, kRightVisited = 2, kBothVisited = 3 };
enum Dir : uint8_t { kUp = 0, kLeft = 1, kRight = 2, kFinished = 3 };

template <typename T, int D>
struct OTree {
    int levels = 0;
    int n_nodes = 0;
    int n_leaves = 0;
    int32_t n_points = 0;
    std::vector<uint8_t> axis;      // Partition::axis_split, heap order
    std::vector<T> median;          // Partition::median, heap order
    std::vector<int32_t> leaf_n;    // PartitionLeaf::nr_points
    std::vector<int32_t> leaf_off;  // PartitionLeaf::offset
    std::vector<std::array<T, D>> structured;
    std::vector<int32_t> shuffled;
};

// kdtree.hpp:73-80
int total_nodes(int levels) {
    int s = 0;
    for (int l = 0; l < levels; ++l) s += 1 << l;
    return s;
}

// kdtree.hpp:392-401
template <typename T, int D>
T comp_median(const std::array<T, D>* p, int32_t n, int a) {
    const int32_t h = n / 2;
    if (n % 2 == 1) return p[h][a];
    return 0.5 * (p[h][a] + p[h - 1][a]);
}

// kdtree.hpp:418-482.  Two sorts per node, exactly as the reference: the index permutation
// first, then the points themselves, both keyed on the split axis with operator<.
template <typename T, int D>
void build_rec(OTree<T, D>& t, int node, std::array<T, D>* P, int32_t* I, int32_t n, int lv, int a,
               int32_t off) {
    if (n == 0)
        throw std::runtime_error(
            "Error: Ran out of points while building KD-Tree. Either you required too many levels, "
            "or you used a lot of coplanar points");
    std::vector<int32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [a, P](int32_t l, int32_t r) { return P[l][a] < P[r][a]; });
    std::vector<int32_t> old(I, I + n);
    for (int32_t e = 0; e < n; ++e) I[e] = old[idx[e]];
    std::sort(P, P + n,
              [a](const std::array<T, D>& l, const std::array<T, D>& r) { return l[a] < r[a]; });
    const T m = comp_median<T, D>(P, n, a);
    const int32_t lo = n / 2, hi = n - lo;
    if (lv == 0) {
        const int left_leaf = 2 * node + 1 - t.n_nodes;  // kdtree.hpp:300-303
        t.leaf_n[left_leaf] = lo;
        t.leaf_off[left_leaf] = off;
        t.leaf_n[left_leaf + 1] = hi;
        t.leaf_off[left_leaf + 1] = off + lo;
    } else {
        build_rec<T, D>(t, 2 * node + 1, P, I, lo, lv - 1, (a + 1) % D, off);
        build_rec<T, D>(t, 2 * node + 2, P + lo, I + lo, hi, lv - 1, (a + 1) % D, off + lo);
    }
    t.axis[node] = static_cast<uint8_t>(a);
    t.median[node] = m;
}

// kdtree.hpp:494-534
template <typename T, int D>
OTree<T, D>* build(const T* flat, int32_t n, int levels) {
    auto* t = new OTree<T, D>();
    t->levels = levels;
    t->n_nodes = total_nodes(levels);
    t->n_leaves = 1 << levels;
    t->n_points = n;
    t->axis.assign(t->n_nodes, 0xFF);
    t->median.assign(t->n_nodes, T(-1));
    t->leaf_n.assign(t->n_leaves, 0);
    t->leaf_off.assign(t->n_leaves, 0);
    t->structured.resize(n);
    std::memcpy(t->structured.data(), flat, sizeof(T) * D * static_cast<size_t>(n));
    t->shuffled.resize(n);
    std::iota(t->shuffled.begin(), t->shuffled.end(), 0);
    try {
        build_rec<T, D>(*t, 0, t->structured.data(), t->shuffled.data(), n, levels - 1, 0, 0);
    } catch (...) {
        delete t;
        throw;
    }
    return t;
}

// kdtree.hpp:546-591: diff = q - r in T, then s = s + diff*diff left to right in T.
template <typename T, int D>
inline T quadr_dist(const std::array<T, D>& q, const std::array<T, D>& r) {
    std::array<T, D> v;
    for (int i = 0; i < D; ++i) v[i] = q[i] - r[i];
    T s = T(0);
    for (int i = 0; i < D; ++i) s = s + v[i] * v[i];
    return s;
}

// kdtree.hpp:606-614: v = pp - q; v[a] += m - pp[a]; sum of squares as above.
template <typename T, int D>
inline T projection_dist(const std::array<T, D>& q, const std::array<T, D>& pp, int a, T m) {
    std::array<T, D> v;
    for (int i = 0; i < D; ++i) v[i] = pp[i] - q[i];
    v[a] = v[a] + (m - pp[a]);
    T s = T(0);
    for (int i = 0; i < D; ++i) s = s + v[i] * v[i];
    return s;
}

// nndistance.hpp:41-60
template <typename T>
inline uint32_t upper_bound_pos(const T* first, uint32_t count, T value) {
    uint32_t pos = 0;
    while (count > 0) {
        const uint32_t step = count / 2;
        if (value >= first[pos + step]) {
            pos += step + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    return pos;
}

// kdtree.hpp:663-692 with nndistance.hpp:74-85 and :115-127
template <typename T, int D>
void scan_leaf(const OTree<T, D>& t, int leaf, const std::array<T, D>& q, T* bd, int32_t* bi, int k) {
    const int32_t n = t.leaf_n[leaf], off = t.leaf_off[leaf];
    const std::array<T, D>* data = t.structured.data() + off;
    for (int32_t r = 0; r < n; ++r) {
        const T d = quadr_dist<T, D>(q, data[r]);
        if (d >= bd[k - 1]) continue;
        const uint32_t pos = upper_bound_pos<T>(bd, static_cast<uint32_t>(k), d);
        for (int j = k - 1; j > static_cast<int>(pos); --j) {
            bd[j] = bd[j - 1];
            bi[j] = bi[j - 1];
        }
        bd[pos] = d;
        bi[pos] = r + off;
    }
}

// The tag-array walker of kdtree.hpp:225-379 (per-query heap allocation and fill included,
// kdtree.hpp:250-256, because the CPU baseline is meant to cost what the reference costs).
template <typename T, int D>
struct Walker {
    const OTree<T, D>& t;
    uint8_t* tags;
    int cur = 0;
    int level = 0;
    explicit Walker(const OTree<T, D>& tr) : t(tr) {
        tags = new uint8_t[t.n_nodes];
        std::fill(tags, tags + t.n_nodes, kUncharted);
    }
    ~Walker() { delete[] tags; }
    bool leaf_parent() const { return level == t.levels - 1; }
    void walk(const std::function<Dir(int, uint8_t, int)>& f) {
        Dir d;
        while ((d = f(cur, tags[cur], level)) != kFinished) {
            if (d == kUp) {
                cur = (cur - 1) / 2;
                level -= 1;
            } else if (d == kLeft) {
                cur = 2 * cur + 1;
                level += 1;
            } else {
                cur = 2 * cur + 2;
                level += 1;
            }
        }
    }
};

// kdtree.cpp:16-110
template <typename T, int D>
void search_iterative(const OTree<T, D>& t, const std::array<T, D>& q, T* bd, int32_t* bi, int k) {
    Walker<T, D> w(t);
    std::array<T, D> pp(q);
    T worst = bd[k - 1];
    const std::function<Dir(int, uint8_t, int)> f = [&](int node, uint8_t tag, int level) -> Dir {
        const int a = t.axis[node];
        const T m = t.median[node];
        const bool lower = q[a] < m;
        if (tag == kBothVisited) {
            if (level == 0) return kFinished;
            pp[a] = q[a];
            return kUp;
        }
        if (w.leaf_parent()) {
            w.tags[node] = kBothVisited;
            return kFinished;
        }
        if (tag == kUncharted) {
            w.tags[node] = lower ? kLeftVisited : kRightVisited;
            return lower ? kLeft : kRight;
        }
        w.tags[node] = kBothVisited;
        pp[a] = m;
        if (projection_dist<T, D>(q, pp, a, m) < worst) return lower ? kRight : kLeft;
        pp[a] = q[a];
        return level != 0 ? kUp : kFinished;
    };
    w.walk(f);
    do {
        const int node = w.cur;
        const int a = t.axis[node];
        const T m = t.median[node];
        const bool lower = q[a] < m;
        const int left_leaf = 2 * node + 1 - t.n_nodes;
        scan_leaf<T, D>(t, lower ? left_leaf : left_leaf + 1, q, bd, bi, k);
        worst = bd[k - 1];
        if (projection_dist<T, D>(q, pp, a, m) < worst)
            scan_leaf<T, D>(t, lower ? left_leaf + 1 : left_leaf, q, bd, bi, k);
        worst = bd[k - 1];
        w.walk(f);
    } while (w.level != 0);
}

// kdtree.cpp:112-152
template <typename T, int D>
void search_recursive(const OTree<T, D>& t, const std::array<T, D>& q, int node, int cl,
                      const std::array<T, D>& pp, T* bd, int32_t* bi, int k) {
    const int a = t.axis[node];
    const T m = t.median[node];
    const bool lower = q[a] < m;
    if (cl == 0) {
        const int left_leaf = 2 * node + 1 - t.n_nodes;
        scan_leaf<T, D>(t, lower ? left_leaf : left_leaf + 1, q, bd, bi, k);
        if (projection_dist<T, D>(q, pp, a, m) < bd[k - 1])
            scan_leaf<T, D>(t, lower ? left_leaf + 1 : left_leaf, q, bd, bi, k);
    } else {
        const int lc = 2 * node + 1, rc = 2 * node + 2;
        search_recursive<T, D>(t, q, lower ? lc : rc, cl - 1, pp, bd, bi, k);
        if (projection_dist<T, D>(q, pp, a, m) < bd[k - 1]) {
            std::array<T, D> pp2(q);
            pp2[a] = m;
            search_recursive<T, D>(t, q, lower ? rc : lc, cl - 1, pp2, bd, bi, k);
        }
    }
}

// kdtree.cpp:164-194
template <typename T, int D>
void search_all(const OTree<T, D>& t, const T* qflat, int64_t nq, int k, T* dist, int32_t* idx,
                int nthreads, bool recursive) {
    const auto* qs = reinterpret_cast<const std::array<T, D>*>(qflat);
    // a num_threads clause, not omp_set_num_threads: the caller's default team size stays as it was
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt)
    for (int64_t i = 0; i < nq; ++i) {
        const std::array<T, D> q = qs[i];
        std::vector<T> bd(k, std::numeric_limits<T>::infinity());
        std::vector<int32_t> bi(k, 0);
        if (recursive)
            search_recursive<T, D>(t, q, 0, t.levels - 1, q, bd.data(), bi.data(), k);
        else
            search_iterative<T, D>(t, q, bd.data(), bi.data(), k);
        std::copy(bd.begin(), bd.end(), dist + i * k);
        std::copy(bi.begin(), bi.end(), idx + i * k);
    }
}

// nndistance_g.cu:6-72 semantics on the original (unpermuted) points: a candidate is inserted
// iff d < worst (nndistance_g.cu:57, 126), so NaN distances never are.
template <typename T, int D>
void bruteforce(const T* rflat, int32_t n, const T* qflat, int64_t nq, int k, T* dist, int32_t* idx,
                int nthreads) {
    const auto* rs = reinterpret_cast<const std::array<T, D>*>(rflat);
    const auto* qs = reinterpret_cast<const std::array<T, D>*>(qflat);
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < nq; ++i) {
        std::vector<T> bd(k, std::numeric_limits<T>::infinity());
        std::vector<int32_t> bi(k, 0);
        for (int32_t j = 0; j < n; ++j) {
            const T d = quadr_dist<T, D>(qs[i], rs[j]);
            if (!(d < bd[k - 1])) continue;
            const uint32_t pos = upper_bound_pos<T>(bd.data(), static_cast<uint32_t>(k), d);
            for (int s = k - 1; s > static_cast<int>(pos); --s) {
                bd[s] = bd[s - 1];
                bi[s] = bi[s - 1];
            }
            bd[pos] = d;
            bi[pos] = j;
        }
        std::copy(bd.begin(), bd.end(), dist + i * k);
        std::copy(bi.begin(), bi.end(), idx + i * k);
    }
}

struct Handle {
    int dtype;  // 0 = float32, 1 = float64
    int dims;
    void* tree;
};

template <typename T>
void* tree_ptr_build(const T* p, int32_t n, int dims, int levels) {
    switch (dims) {
        case 1: return build<T, 1>(p, n, levels);
        case 2: return build<T, 2>(p, n, levels);
        case 3: return build<T, 3>(p, n, levels);
        case 8: return build<T, 8>(p, n, levels);
        default: throw std::runtime_error("oracle: unsupported dims");
    }
}

#define ORACLE_DISPATCH(h, ...)                                               \
    do {                                                                       \
        if ((h)->dtype == 0) {                                                 \
            using T = float;                                                   \
            switch ((h)->dims) {                                               \
                case 1: { constexpr int D = 1; auto& t = *static_cast<OTree<T, D>*>((h)->tree); __VA_ARGS__; } break; \
                case 2: { constexpr int D = 2; auto& t = *static_cast<OTree<T, D>*>((h)->tree); __VA_ARGS__; } break; \
                case 3: { constexpr int D = 3; auto& t = *static_cast<OTree<T, D>*>((h)->tree); __VA_ARGS__; } break; \
                case 8: { constexpr int D = 8; auto& t = *static_cast<OTree<T, D>*>((h)->tree); __VA_ARGS__; } break; \
            }                                                                  \
        } else {                                                               \
            using T = double;                                                  \
            switch ((h)->dims) {                                               \
                case 1: { constexpr int D = 1; auto& t = *static_cast<OTree<T, D>*>((h)->tree); __VA_ARGS__; } break; \
                case 2: { constexpr int D = 2; auto& t = *static_cast<OTree<T, D>*>((h)->tree); __VA_ARGS__; } break; \
                case 3: { constexpr int D = 3; auto& t = *static_cast<OTree<T, D>*>((h)->tree); __VA_ARGS__; } break; \
                case 8: { constexpr int D = 8; auto& t = *static_cast<OTree<T, D>*>((h)->tree); __VA_ARGS__; } break; \
            }                                                                  \
        }                                                                      \
    } while (0)

}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// Returns an opaque tree handle, or NULL with oracle_last_error() set.
void* oracle_build(const void* points, int32_t n, int dims, int levels, int dtype) {
    try {
        if (dims != 1 && dims != 2 && dims != 3 && dims != 8) throw std::runtime_error("oracle: unsupported dims");
        if (levels < 1) throw std::runtime_error("oracle: levels must be >= 1");
        auto* h = new Handle{dtype, dims, nullptr};
        try {
            h->tree = dtype == 0 ? tree_ptr_build<float>(static_cast<const float*>(points), n, dims, levels)
                                 : tree_ptr_build<double>(static_cast<const double*>(points), n, dims, levels);
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// Wraps a tree given as arrays (the layout oracle_tree_export writes) so that the oracle's
// search can run on a tree built elsewhere.
void* oracle_tree_from_arrays(const void* structured, const int32_t* shuffled, const int32_t* axis,
                              const void* median, const int32_t* leaf_n, const int32_t* leaf_off,
                              int32_t n, int dims, int levels, int dtype) {
    try {
        auto* h = new Handle{dtype, dims, nullptr};
        auto fill = [&](auto* t) {
            using TreeT = std::remove_pointer_t<decltype(t)>;
            using T = typename decltype(TreeT::median)::value_type;
            t->levels = levels;
            t->n_nodes = total_nodes(levels);
            t->n_leaves = 1 << levels;
            t->n_points = n;
            t->axis.resize(t->n_nodes);
            for (int i = 0; i < t->n_nodes; ++i) t->axis[i] = static_cast<uint8_t>(axis[i]);
            t->median.assign(static_cast<const T*>(median), static_cast<const T*>(median) + t->n_nodes);
            t->leaf_n.assign(leaf_n, leaf_n + t->n_leaves);
            t->leaf_off.assign(leaf_off, leaf_off + t->n_leaves);
            t->structured.resize(n);
            std::memcpy(t->structured.data(), structured, sizeof(t->structured[0]) * static_cast<size_t>(n));
            t->shuffled.assign(shuffled, shuffled + n);
            h->tree = t;
        };
        if (dtype == 0) {
            switch (dims) {
                case 1: fill(new OTree<float, 1>()); break;
                case 2: fill(new OTree<float, 2>()); break;
                case 3: fill(new OTree<float, 3>()); break;
                case 8: fill(new OTree<float, 8>()); break;
                default: delete h; throw std::runtime_error("oracle: unsupported dims");
            }
        } else {
            switch (dims) {
                case 1: fill(new OTree<double, 1>()); break;
                case 2: fill(new OTree<double, 2>()); break;
                case 3: fill(new OTree<double, 3>()); break;
                case 8: fill(new OTree<double, 8>()); break;
                default: delete h; throw std::runtime_error("oracle: unsupported dims");
            }
        }
        return h;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void oracle_free(void* hp) {
    auto* h = static_cast<Handle*>(hp);
    if (!h) return;
    ORACLE_DISPATCH(h, delete &t);
    delete h;
}

void oracle_tree_info(const void* hp, int32_t* out /* levels, n_nodes, n_leaves, n_points */) {
    const auto* h = static_cast<const Handle*>(hp);
    ORACLE_DISPATCH(h, {
        out[0] = t.levels;
        out[1] = t.n_nodes;
        out[2] = t.n_leaves;
        out[3] = t.n_points;
    });
}

// Copies the tree out: structured points [N*dims], shuffled [N], axis [n_nodes] (as int32),
// median [n_nodes], leaf_n [n_leaves], leaf_off [n_leaves].
void oracle_tree_export(const void* hp, void* structured, int32_t* shuffled, int32_t* axis, void* median,
                        int32_t* leaf_n, int32_t* leaf_off) {
    const auto* h = static_cast<const Handle*>(hp);
    ORACLE_DISPATCH(h, {
        std::memcpy(structured, t.structured.data(), sizeof(T) * D * static_cast<size_t>(t.n_points));
        std::memcpy(shuffled, t.shuffled.data(), sizeof(int32_t) * t.n_points);
        for (int i = 0; i < t.n_nodes; ++i) axis[i] = t.axis[i];
        std::memcpy(median, t.median.data(), sizeof(T) * t.n_nodes);
        std::memcpy(leaf_n, t.leaf_n.data(), sizeof(int32_t) * t.n_leaves);
        std::memcpy(leaf_off, t.leaf_off.data(), sizeof(int32_t) * t.n_leaves);
    });
}

// KDTreeKNNSearch restated: squared distances [nq*k] and structured int32 indices [nq*k].
// recursive != 0 runs compQuadrDistNode instead of the iterative walker.
int oracle_search(const void* hp, const void* q, int64_t nq, int k, void* dist, int32_t* idx, int nthreads,
                  int recursive) {
    const auto* h = static_cast<const Handle*>(hp);
    try {
        ORACLE_DISPATCH(h, search_all<T, D>(t, static_cast<const T*>(q), nq, k, static_cast<T*>(dist), idx,
                                            nthreads, recursive != 0));
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
    return 0;
}

// Exact KNN by (distance, original index) on the original point order; NaN distances excluded.
int oracle_bruteforce(const void* ref, int32_t n, const void* q, int64_t nq, int dims, int k, void* dist,
                      int32_t* idx, int dtype, int nthreads) {
    try {
#define BF(TT, DD) bruteforce<TT, DD>(static_cast<const TT*>(ref), n, static_cast<const TT*>(q), nq, k, \
                                      static_cast<TT*>(dist), idx, nthreads)
        if (dtype == 0) {
            switch (dims) {
                case 1: BF(float, 1); break;
                case 2: BF(float, 2); break;
                case 3: BF(float, 3); break;
                case 8: BF(float, 8); break;
                default: throw std::runtime_error("oracle: unsupported dims");
            }
        } else {
            switch (dims) {
                case 1: BF(double, 1); break;
                case 2: BF(double, 2); break;
                case 3: BF(double, 3); break;
                case 8: BF(double, 8); break;
                default: throw std::runtime_error("oracle: unsupported dims");
            }
        }
#undef BF
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
    return 0;
}

}  // extern "C"
