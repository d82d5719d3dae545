// GPU tree build: the layout of createKDTree (src/kdtree.hpp:418-534) built on the device, for one
// tree or for many trees at once (a forest over one concatenated point array).
//
// The reference recurses node by node: at a node of depth l it sorts the node's points by axis
// l mod dims, takes the median of compMedian (src/kdtree.hpp:392-401) as split value, gives the
// first n/2 points to the left child and the rest to the right (src/kdtree.hpp:447-452).  Here
// every depth is one step over all nodes of all trees:
//   1. key[p] = the sortable bits of coordinate (l mod dims) of the point at position p;
//   2. a stable radix sort of (key, p) over the whole array;
//   3. a stable radix sort of the result by node (segment) number, which regroups the positions
//      node by node with each node's points in coordinate order;
//   4. the permutation is composed with that order, and each node's median is read off.
// After the last depth the structured points are gathered and copied to the host trees.
//
// Exactness: the node sizes, split values and leaf contents equal those of the reference
// whenever no node has two points with equal coordinates on its split axis.  With such ties the
// reference's two std::sort calls (one permutation, applied to both arrays: std::sort's moves
// follow its comparison outcomes alone) order the tied points in the order libstdc++'s introsort
// leaves them, which only a sequential emulation could follow; this build orders them by their
// previous position.  It is therefore opt-in; the default build stays the host restatement.
// Points with a NaN coordinate are refused: std::sort's order for them is the introsort's
// accident too.  The host build accepts them and reproduces the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpu_common.hpp"
#include "knn_gpu.hpp"
#include "tree.hpp"

namespace tk {
namespace {

constexpr int kGbBlock = 256;

__device__ __forceinline__ uint32_t sortable(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint64_t sortable(double x) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | (uint64_t(1) << 63));
}

// Tree of arena position p: the last c with base[c] <= p (bases ascending, base[nc] = total).
__device__ __forceinline__ int tree_of(const int64_t* __restrict__ base, int nc, int64_t p) {
    int lo = 0, hi = nc - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Node (segment) of local position p at depth l in a tree of n points: the reference's split
// rule, n/2 points to the left child and n - n/2 to the right.
__device__ __forceinline__ uint32_t segment_of(int64_t p, int64_t n, int depth) {
    int64_t off = 0, cnt = n;
    uint32_t s = 0;
    for (int i = 0; i < depth; ++i) {
        const int64_t h = cnt >> 1;
        if (p < off + h) {
            cnt = h;
            s = 2 * s;
        } else {
            off += h;
            cnt -= h;
            s = 2 * s + 1;
        }
    }
    return s;
}

struct ForestDev {
    const int64_t* base;      // [nc + 1] first point of each tree in the arena
    const int32_t* levels;    // [nc]
    const int64_t* split_base;  // [nc] first split value of each tree
    int nc;
};

// key[j] = word `word` (0 low, 1 high) of the sortable coordinate of position p = order[j] (or j);
// positions of trees that are complete at this depth get key 0, so they keep their order.
template <typename T>
__global__ __launch_bounds__(kGbBlock) void gb_key_kernel(const T* __restrict__ pts, int dims,
                                                          const uint32_t* __restrict__ perm,
                                                          const uint32_t* __restrict__ order, int64_t n,
                                                          ForestDev f, int depth, int axis, int word,
                                                          uint32_t* __restrict__ key, uint32_t* __restrict__ val,
                                                          int32_t* __restrict__ nan_flag) {
    const int64_t j = (int64_t)blockIdx.x * kGbBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t p = order ? order[j] : (uint32_t)j;
    const int c = tree_of(f.base, f.nc, p);
    uint32_t k = 0;
    if (f.levels[c] > depth) {
        const T x = pts[(size_t)perm[p] * dims + axis];
        if (x != x) atomicOr(nan_flag, 1);
        if constexpr (sizeof(T) == 4) {
            k = sortable(x);
        } else {
            const uint64_t u = sortable(x);
            k = word ? (uint32_t)(u >> 32) : (uint32_t)u;
        }
    }
    key[j] = k;
    val[j] = p;
}

// key[j] = node number (over all trees) of position p = order[j] at this depth.
__global__ __launch_bounds__(kGbBlock) void gb_segment_kernel(const uint32_t* __restrict__ order, int64_t n, ForestDev f,
                                                              int depth, const uint32_t* __restrict__ seg_base,
                                                              uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
    const int64_t j = (int64_t)blockIdx.x * kGbBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t p = order[j];
    const int c = tree_of(f.base, f.nc, p);
    const int64_t b = f.base[c];
    const uint32_t s = f.levels[c] > depth ? segment_of((int64_t)p - b, f.base[c + 1] - b, depth) : 0u;
    key[j] = seg_base[c] + s;
    val[j] = p;
}

__global__ __launch_bounds__(kGbBlock) void gb_compose_kernel(const uint32_t* __restrict__ perm,
                                                              const uint32_t* __restrict__ order, int64_t n,
                                                              uint32_t* __restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * kGbBlock + threadIdx.x;
    if (j >= n) return;
    out[j] = perm[order[j]];
}

__global__ __launch_bounds__(kGbBlock) void gb_iota_kernel(uint32_t* __restrict__ out, int64_t n) {
    const int64_t j = (int64_t)blockIdx.x * kGbBlock + threadIdx.x;
    if (j >= n) return;
    out[j] = (uint32_t)j;
}

// Split values of the nodes at this depth: task t is node s of tree c (task_base[c] <= t <
// task_base[c + 1]); the node's points, sorted on its axis, sit at its local offset.  The median
// is compMedian's: the middle point for an odd count, else half the sum of the two middle points
// (the sum rounded in T, the halving exact).
template <typename T>
__global__ __launch_bounds__(kGbBlock) void gb_median_kernel(const T* __restrict__ pts, int dims,
                                                             const uint32_t* __restrict__ perm, ForestDev f,
                                                             int depth, int axis,
                                                             const int64_t* __restrict__ task_base,
                                                             int64_t n_tasks, T* __restrict__ split) {
    const int64_t t = (int64_t)blockIdx.x * kGbBlock + threadIdx.x;
    if (t >= n_tasks) return;
    const int c = tree_of(task_base, f.nc, t);
    const uint32_t s = (uint32_t)(t - task_base[c]);
    const int64_t b = f.base[c];
    int64_t off = 0, cnt = f.base[c + 1] - b;
    for (int i = depth - 1; i >= 0; --i) {
        const int64_t h = cnt >> 1;
        if ((s >> i) & 1u) {
            off += h;
            cnt -= h;
        } else {
            cnt = h;
        }
    }
    const int64_t mid = b + off + (cnt >> 1);
    const T hi = pts[(size_t)perm[mid] * dims + axis];
    T m = hi;
    if ((cnt & 1) == 0) {
        const T lo = pts[(size_t)perm[mid - 1] * dims + axis];
        m = (lo + hi) * T(0.5);
    }
    split[f.split_base[c] + ((int64_t(1) << depth) - 1) + s] = m;
}

// Structured points and local original indices: out[j] = pts[perm[j]], idx[j] = perm[j] - base.
template <typename T>
__global__ __launch_bounds__(kGbBlock) void gb_gather_kernel(const T* __restrict__ pts, int dims,
                                                             const uint32_t* __restrict__ perm, int64_t n, ForestDev f,
                                                             T* __restrict__ out, int32_t* __restrict__ idx) {
    const int64_t j = (int64_t)blockIdx.x * kGbBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t p = perm[j];
    const int c = tree_of(f.base, f.nc, j);
    idx[j] = (int32_t)((int64_t)p - f.base[c]);
    for (int a = 0; a < dims; ++a) out[(size_t)j * dims + a] = pts[(size_t)p * dims + a];
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kGbBlock - 1) / kGbBlock); }

int bits_for_count(uint64_t count) {  // bits to number `count` values 0..count-1
    int b = 0;
    while (b < 32 && (uint64_t(1) << b) < count) ++b;
    return b;
}

// Leaf offsets of the reference's split rule (n/2 left, n - n/2 right at every node).
std::vector<int32_t> leaf_offsets(int64_t n, int levels) {
    std::vector<int64_t> sizes{n};
    for (int l = 0; l < levels; ++l) {
        std::vector<int64_t> next;
        next.reserve(sizes.size() * 2);
        for (int64_t s : sizes) {
            next.push_back(s / 2);
            next.push_back(s - s / 2);
        }
        sizes.swap(next);
    }
    std::vector<int32_t> off(sizes.size() + 1, 0);
    for (size_t i = 0; i < sizes.size(); ++i) off[i + 1] = off[i] + (int32_t)sizes[i];
    return off;
}

struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { (void)hipFree(p); }
    void alloc(size_t bytes) { TK_HIP_CHECK(hipMalloc(&p, bytes ? bytes : 1)); }
    template <typename U> U* as() const { return static_cast<U*>(p); }
};

template <typename T>
std::vector<Tree*> build_forest_typed(const T* pts, const int64_t* n_points, const int32_t* levels, int nc, int dims,
                                      hipStream_t s) {
    std::vector<int64_t> base(nc + 1, 0), split_base(nc, 0);
    int max_levels = 0;
    int64_t total_nodes_all = 0;
    for (int c = 0; c < nc; ++c) {
        base[c + 1] = base[c] + n_points[c];
        split_base[c] = total_nodes_all;
        total_nodes_all += total_nodes(levels[c]);
        max_levels = std::max(max_levels, (int)levels[c]);
    }
    const int64_t n = base[nc];
    // per depth: node numbering (segment bases) and median task bases
    std::vector<uint32_t> seg_base((size_t)max_levels * nc);
    std::vector<int64_t> task_base((size_t)max_levels * (nc + 1));
    std::vector<uint64_t> seg_total(max_levels);
    for (int l = 0; l < max_levels; ++l) {
        uint64_t sb = 0;
        int64_t tb = 0;
        for (int c = 0; c < nc; ++c) {
            seg_base[(size_t)l * nc + c] = (uint32_t)sb;
            task_base[(size_t)l * (nc + 1) + c] = tb;
            const bool active = levels[c] > l;
            sb += active ? (uint64_t(1) << l) : 1u;
            tb += active ? (int64_t(1) << l) : 0;
        }
        task_base[(size_t)l * (nc + 1) + nc] = tb;
        seg_total[l] = sb;
        if (sb > (uint64_t(1) << 32)) throw std::runtime_error("GPU tree build: too many nodes at one depth");
    }

    DevBuf d_base, d_levels, d_split_base, d_seg_base, d_task_base, d_perm_a, d_perm_b, d_key, d_val,
        d_key_s, d_val_s, d_split, d_flag, d_ws;
    d_base.alloc(sizeof(int64_t) * (nc + 1));
    d_levels.alloc(sizeof(int32_t) * nc);
    d_split_base.alloc(sizeof(int64_t) * nc);
    d_seg_base.alloc(sizeof(uint32_t) * seg_base.size());
    d_task_base.alloc(sizeof(int64_t) * task_base.size());
    d_perm_a.alloc(sizeof(uint32_t) * n);
    d_perm_b.alloc(sizeof(uint32_t) * n);
    d_key.alloc(sizeof(uint32_t) * n);
    d_val.alloc(sizeof(uint32_t) * n);
    d_key_s.alloc(sizeof(uint32_t) * n);
    d_val_s.alloc(sizeof(uint32_t) * n);
    d_split.alloc(sizeof(T) * total_nodes_all);
    d_flag.alloc(sizeof(int32_t));
    const size_t ws_bytes = radix_sort_workspace(n);
    d_ws.alloc(ws_bytes);
    TK_HIP_CHECK(hipMemcpyAsync(d_base.p, base.data(), sizeof(int64_t) * (nc + 1), hipMemcpyHostToDevice, s));
    TK_HIP_CHECK(hipMemcpyAsync(d_levels.p, levels, sizeof(int32_t) * nc, hipMemcpyHostToDevice, s));
    TK_HIP_CHECK(hipMemcpyAsync(d_split_base.p, split_base.data(), sizeof(int64_t) * nc, hipMemcpyHostToDevice, s));
    TK_HIP_CHECK(hipMemcpyAsync(d_seg_base.p, seg_base.data(), sizeof(uint32_t) * seg_base.size(),
                                hipMemcpyHostToDevice, s));
    TK_HIP_CHECK(hipMemcpyAsync(d_task_base.p, task_base.data(), sizeof(int64_t) * task_base.size(),
                                hipMemcpyHostToDevice, s));
    TK_HIP_CHECK(hipMemsetAsync(d_flag.p, 0, sizeof(int32_t), s));
    const ForestDev f{d_base.as<int64_t>(), d_levels.as<int32_t>(), d_split_base.as<int64_t>(), nc};

    uint32_t* perm = d_perm_a.as<uint32_t>();
    uint32_t* perm_next = d_perm_b.as<uint32_t>();
    uint32_t* key = d_key.as<uint32_t>();
    uint32_t* val = d_val.as<uint32_t>();
    uint32_t* key_s = d_key_s.as<uint32_t>();
    uint32_t* val_s = d_val_s.as<uint32_t>();
    if (n > 0) {
        hipLaunchKernelGGL(gb_iota_kernel, dim3(grid_for(n)), dim3(kGbBlock), 0, s, perm, n);
        TK_HIP_CHECK(hipGetLastError());
    }
    for (int l = 0; l < max_levels && n > 0; ++l) {
        const int axis = l % dims;
        // coordinate order (two 32-bit words for doubles: low word first, then high, both stable)
        const int words = sizeof(T) == 8 ? 2 : 1;
        for (int w = 0; w < words; ++w) {
            hipLaunchKernelGGL(gb_key_kernel<T>, dim3(grid_for(n)), dim3(kGbBlock), 0, s, pts, dims, perm,
                               w == 0 ? (const uint32_t*)nullptr : val_s, n, f, l, axis, words == 2 ? w : 0, key, val,
                               d_flag.as<int32_t>());
            TK_HIP_CHECK(hipGetLastError());
            radix_sort_pairs(key, val, key_s, val_s, n, 32, d_ws.p, ws_bytes, s);
        }
        // node order, stable: each node's points stay in coordinate order
        const int seg_bits = bits_for_count(seg_total[l]);
        if (seg_bits > 0) {
            hipLaunchKernelGGL(gb_segment_kernel, dim3(grid_for(n)), dim3(kGbBlock), 0, s, val_s, n, f, l,
                               d_seg_base.as<uint32_t>() + (size_t)l * nc, key, val);
            TK_HIP_CHECK(hipGetLastError());
            radix_sort_pairs(key, val, key_s, val_s, n, seg_bits, d_ws.p, ws_bytes, s);
        }
        hipLaunchKernelGGL(gb_compose_kernel, dim3(grid_for(n)), dim3(kGbBlock), 0, s, perm, val_s, n, perm_next);
        TK_HIP_CHECK(hipGetLastError());
        std::swap(perm, perm_next);
        const int64_t tasks = task_base[(size_t)l * (nc + 1) + nc];
        if (tasks > 0) {
            hipLaunchKernelGGL(gb_median_kernel<T>, dim3(grid_for(tasks)), dim3(kGbBlock), 0, s, pts, dims, perm, f, l,
                               axis, d_task_base.as<int64_t>() + (size_t)l * (nc + 1), tasks, d_split.as<T>());
            TK_HIP_CHECK(hipGetLastError());
        }
    }
    // structured points and local original indices
    DevBuf d_out_pts, d_out_idx;
    d_out_pts.alloc(sizeof(T) * (size_t)n * dims);
    d_out_idx.alloc(sizeof(int32_t) * (size_t)n);
    if (n > 0) {
        hipLaunchKernelGGL(gb_gather_kernel<T>, dim3(grid_for(n)), dim3(kGbBlock), 0, s, pts, dims, perm, n, f,
                           d_out_pts.as<T>(), d_out_idx.as<int32_t>());
        TK_HIP_CHECK(hipGetLastError());
    }
    int32_t nan_flag = 0;
    std::vector<T> split_h(total_nodes_all);
    std::vector<T> pts_h((size_t)n * dims);
    std::vector<int32_t> idx_h(n);
    TK_HIP_CHECK(hipMemcpyAsync(&nan_flag, d_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TK_HIP_CHECK(hipMemcpyAsync(split_h.data(), d_split.p, sizeof(T) * split_h.size(), hipMemcpyDeviceToHost, s));
    TK_HIP_CHECK(hipMemcpyAsync(pts_h.data(), d_out_pts.p, sizeof(T) * pts_h.size(), hipMemcpyDeviceToHost, s));
    TK_HIP_CHECK(hipMemcpyAsync(idx_h.data(), d_out_idx.p, sizeof(int32_t) * idx_h.size(), hipMemcpyDeviceToHost, s));
    TK_HIP_CHECK(hipStreamSynchronize(s));
    if (nan_flag)
        throw std::runtime_error(
            "GPU tree build: the points contain NaN (the host build, builder=\"cpu\", accepts them and orders "
            "them as the reference does)");

    std::vector<Tree*> out(nc, nullptr);
    try {
        for (int c = 0; c < nc; ++c) {
            std::unique_ptr<Tree> t(new Tree());
            t->dtype = sizeof(T) == 4 ? kF32 : kF64;
            t->dims = dims;
            t->levels = levels[c];
            t->n_points = (int32_t)n_points[c];
            t->n_nodes = total_nodes(levels[c]);
            t->n_leaves = (int32_t)(1u << levels[c]);
            t->split.resize(sizeof(T) * (size_t)t->n_nodes);
            std::memcpy(t->split.data(), split_h.data() + split_base[c], t->split.size());
            t->leaf_off = leaf_offsets(n_points[c], levels[c]);
            t->points.resize(sizeof(T) * (size_t)n_points[c] * dims);
            std::memcpy(t->points.data(), pts_h.data() + (size_t)base[c] * dims, t->points.size());
            t->shuffled.assign(idx_h.begin() + base[c], idx_h.begin() + base[c + 1]);
            finish_tree(*t);
            out[c] = t.release();
        }
    } catch (...) {
        for (Tree* t : out) delete t;
        throw;
    }
    return out;
}

}  // namespace

std::vector<Tree*> build_trees_gpu(const void* points, const int64_t* n_points, const int32_t* levels, int n_trees,
                                   int dims, int dtype, int device, void* stream) {
    if (n_trees < 1) throw std::runtime_error("GPU tree build: no trees");
    if (!dims_supported(dims))
        throw std::runtime_error("Could not find the KD-Tree for dimensionality " + std::to_string(dims));
    if (dtype != kF32 && dtype != kF64) throw std::runtime_error("Unsupported dtype");
    int64_t total = 0;
    for (int c = 0; c < n_trees; ++c) {
        const int L = levels[c];
        if (L < 1 || L > kMaxLevels)
            throw std::runtime_error("levels must be in [1, " + std::to_string(kMaxLevels) + "]");
        if (n_points[c] < 0 || n_points[c] >= (int64_t(1) << 31))
            throw std::runtime_error("Only 32 bit signed indexing implemented");
        if (n_points[c] < (int64_t(1) << (L - 1)))
            throw std::runtime_error(
                "Error: Ran out of points while building KD-Tree. Either you required too many levels, or you "
                "used a lot of coplanar points");
        total += n_points[c];
    }
    if (total > (int64_t(1) << 28)) throw std::runtime_error("GPU tree build: at most 2^28 points in one build");
    if (total > 0 && !points) throw std::runtime_error("invalid arguments");
    if (device < 0) TK_HIP_CHECK(hipGetDevice(&device));
    DeviceGuard g(device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == kF32) return build_forest_typed<float>(static_cast<const float*>(points), n_points, levels, n_trees, dims, s);
    return build_forest_typed<double>(static_cast<const double*>(points), n_points, levels, n_trees, dims, s);
}

}  // namespace tk
