// KD-tree KNN traversal kernels for gfx950.
//
// One query per lane.  Each lane runs the reference's near-first depth-first search
// (src/kdtree.cpp:16-110) without a stack or a tag array: the state is the node index, its
// depth and the projected point pp (D registers).  Coming back up from a child, the lane
// recomputes which side was the near one (q[a] < m) to tell whether the far side is still
// to be tested; that is exactly the information the reference's per-node tags hold for the
// nodes on the active path, and nodes off the path are never revisited.  The pruning bound,
// the pp updates and the candidate insertion repeat the reference's arithmetic, so with the
// reference tie policy every lane produces what the reference CPU search produces.
//
// Queries arrive in an order given by `order` (the leaf-bucket presort below), so the lanes
// of a wavefront descend into the same leaves and their record loads coalesce.
#pragma once

#include "gpu_common.hpp"

namespace tk {

template <typename T>
struct TreeView {
    const T* __restrict__ split;            // [n_nodes] heap order
    const unsigned char* __restrict__ recs; // [n_points] records, structured order
    const int32_t* __restrict__ leaf_off;   // [n_leaves + 1]
    const int32_t* __restrict__ shuffled;   // [n_points]
    int levels;
    int32_t n_nodes;
};

// k-best list held in LDS for k beyond the register buckets: slot j of lane t lives at
// base[j * blockDim + t] (consecutive lanes hit consecutive banks).
template <typename T, bool kLower>
struct KBestLds {
    T* d;
    int32_t* s;
    int32_t* o;
    int k;
    int stride;
    T w;        // cached worst distance
    int32_t wo; // cached worst original index

    __device__ __forceinline__ void init() {
        for (int j = 0; j < k; ++j) {
            d[j * stride] = pos_inf<T>();
            s[j * stride] = 0;
            if constexpr (kLower) o[j * stride] = 0x7fffffff;
        }
        w = pos_inf<T>();
        wo = 0x7fffffff;
    }
    __device__ __forceinline__ T worst() const { return w; }
    __device__ __forceinline__ bool accepts(T dist, int32_t orig) const {
        if constexpr (kLower)
            return dist < w || (dist == w && orig < wo);
        else
            return dist < w;
    }
    __device__ __forceinline__ bool after(int j, T dist, int32_t orig) const {
        const T dj = d[j * stride];
        if constexpr (kLower)
            return dj > dist || (dj == dist && o[j * stride] > orig);
        else
            return dj > dist;
    }
    __device__ __forceinline__ void insert(T dist, int32_t sidx, int32_t orig) {
        int j = k - 1;
        while (j > 0 && after(j - 1, dist, orig)) {
            d[j * stride] = d[(j - 1) * stride];
            s[j * stride] = s[(j - 1) * stride];
            if constexpr (kLower) o[j * stride] = o[(j - 1) * stride];
            --j;
        }
        d[j * stride] = dist;
        s[j * stride] = sidx;
        if constexpr (kLower) o[j * stride] = orig;
        w = d[(k - 1) * stride];
        if constexpr (kLower) wo = o[(k - 1) * stride];
    }
};

template <typename T, int D, typename List>
__device__ __forceinline__ void scan_leaf(const TreeView<T>& tv, int32_t leaf, const T (&q)[D], List& kb) {
    const int32_t beg = tv.leaf_off[leaf];
    const int32_t end = tv.leaf_off[leaf + 1];
    for (int32_t s = beg; s < end; ++s) {
        const Rec<T, D> r = load_rec<T, D>(tv.recs, s);
        const T dd = sq_dist<T, D>(q, r.c);
        if (kb.accepts(dd, r.o)) kb.insert(dd, s, r.o);
    }
}

template <typename T, int D, bool kLower>
__device__ __forceinline__ bool far_side_needed(const T (&q)[D], const T (&pp)[D], int a, T m, T worst) {
    if constexpr (kLower)
        return clean_bound<T, D>(q, pp, a, m) <= worst;
    else
        return proj_bound<T, D>(q, pp, a, m) < worst;
}

// The search of one query; `kb` must be initialised.
template <typename T, int D, bool kLower, typename List>
__device__ __forceinline__ void knn_search_one(const TreeView<T>& tv, const T (&q)[D], List& kb) {
    T pp[D];
#pragma unroll
    for (int i = 0; i < D; ++i) pp[i] = q[i];
    const int L = tv.levels;
    const int32_t nn = tv.n_nodes;
    int32_t node = 0;
    int level = 0;
    while (level < L - 1) {
        const int a = level % D;
        node = 2 * node + (q[a] < tv.split[node] ? 1 : 2);
        ++level;
    }
    // phase 0: the near leaf of leaf-parent `node` is next; phase 1: the far leaf is to be
    // tested; phase 2: both leaves are done, climb.
    int phase = 0;
    for (;;) {
        int32_t leaf;
        const int a = level % D;
        const T m = tv.split[node];
        const bool lower = q[a] < m;
        if (phase == 0) {
            leaf = 2 * node + (lower ? 1 : 2) - nn;
            phase = 1;
        } else if (phase == 1 && far_side_needed<T, D, kLower>(q, pp, a, m, kb.worst())) {
            leaf = 2 * node + (lower ? 2 : 1) - nn;
            phase = 2;
        } else {
            if (level == 0) break;
            pp[a] = q[a];
            bool entered = false;
            while (level > 0) {
                const int32_t child = node;
                node = (node - 1) >> 1;
                --level;
                const int a2 = level % D;
                const T m2 = tv.split[node];
                const bool lower2 = q[a2] < m2;
                if (child == 2 * node + (lower2 ? 1 : 2)) {
                    pp[a2] = m2;
                    if (far_side_needed<T, D, kLower>(q, pp, a2, m2, kb.worst())) {
                        node = 2 * node + (lower2 ? 2 : 1);
                        ++level;
                        entered = true;
                        break;
                    }
                }
                pp[a2] = q[a2];
            }
            if (!entered) break;
            while (level < L - 1) {
                const int a3 = level % D;
                node = 2 * node + (q[a3] < tv.split[node] ? 1 : 2);
                ++level;
            }
            const int a4 = level % D;
            leaf = 2 * node + (q[a4] < tv.split[node] ? 1 : 2) - nn;
            phase = 1;
        }
        scan_leaf<T, D>(tv, leaf, q, kb);
    }
}

template <typename T, int D>
__device__ __forceinline__ void load_query(const T* __restrict__ queries, int64_t qi, T (&q)[D]) {
#pragma unroll
    for (int i = 0; i < D; ++i) q[i] = queries[qi * D + i];
}

constexpr int kTreeBlock = 256;

template <typename T, int D, int K, bool kLower>
__global__ __launch_bounds__(kTreeBlock) void knn_tree_kernel(TreeView<T> tv, const T* __restrict__ queries,
                                                               int64_t nq, const int32_t* __restrict__ order, int k,
                                                               T* __restrict__ out_d, int32_t* __restrict__ out_i,
                                                               int64_t* __restrict__ out_o) {
    const int64_t t = (int64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (t >= nq) return;
    const int64_t qi = order ? (int64_t)order[t] : t;
    T q[D];
    load_query<T, D>(queries, qi, q);
    KBest<T, K, kLower> kb;
    kb.init(k);
    knn_search_one<T, D, kLower>(tv, q, kb);
    const int pad = K - k;
    T* od = out_d + qi * k;
    int32_t* oi = out_i + qi * k;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j >= pad) {
            od[j - pad] = kb.d[j];
            oi[j - pad] = kb.s[j];
        }
    }
    if (out_o) {
        int64_t* oo = out_o + qi * k;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (j >= pad) {
                if constexpr (kLower)
                    oo[j - pad] = kb.o[j];
                else
                    oo[j - pad] = tv.shuffled[kb.s[j]];
            }
        }
    }
}

constexpr int kLdsBlock = 64;

template <typename T, int D, bool kLower>
__global__ __launch_bounds__(kLdsBlock) void knn_tree_lds_kernel(TreeView<T> tv, const T* __restrict__ queries,
                                                                  int64_t nq, const int32_t* __restrict__ order,
                                                                  int k, T* __restrict__ out_d,
                                                                  int32_t* __restrict__ out_i,
                                                                  int64_t* __restrict__ out_o) {
    extern __shared__ unsigned char smem[];
    const int64_t t = (int64_t)blockIdx.x * kLdsBlock + threadIdx.x;
    if (t >= nq) return;
    const int64_t qi = order ? (int64_t)order[t] : t;
    T* sd = reinterpret_cast<T*>(smem);
    int32_t* ss = reinterpret_cast<int32_t*>(sd + (size_t)k * kLdsBlock);
    int32_t* so = ss + (size_t)k * kLdsBlock;
    KBestLds<T, kLower> kb{sd + threadIdx.x, ss + threadIdx.x, so + threadIdx.x, k, kLdsBlock, T(0), 0};
    kb.init();
    T q[D];
    load_query<T, D>(queries, qi, q);
    knn_search_one<T, D, kLower>(tv, q, kb);
    for (int j = 0; j < k; ++j) {
        out_d[qi * k + j] = kb.d[j * kLdsBlock];
        out_i[qi * k + j] = kb.s[j * kLdsBlock];
        if (out_o) out_o[qi * k + j] = kLower ? kb.o[j * kLdsBlock] : tv.shuffled[kb.s[j * kLdsBlock]];
    }
}

inline size_t lds_list_bytes(int k, size_t tsize, bool lower) {
    return (size_t)k * kLdsBlock * (tsize + 4 + (lower ? 4 : 0));
}

// ---------------------------------------------------------------------------------------------
// Leaf-bucket presort: each query descends to the leaf on its near side (no pruning) and is
// counted into that leaf's bucket; an exclusive scan of the counts plus the per-query rank
// gives a permutation that groups queries by leaf.  Leaves are numbered left to right, so
// neighbouring buckets are neighbouring cells of space.
// ---------------------------------------------------------------------------------------------
constexpr int kPresortBlock = 256;

template <typename T, int D>
__global__ __launch_bounds__(kPresortBlock) void leaf_bucket_kernel(TreeView<T> tv, const T* __restrict__ queries,
                                                                     int64_t nq, int shift,
                                                                     int32_t* __restrict__ bucket,
                                                                     int32_t* __restrict__ rank,
                                                                     int32_t* __restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * kPresortBlock + threadIdx.x;
    if (i >= nq) return;
    T q[D];
    load_query<T, D>(queries, i, q);
    int32_t node = 0;
    for (int level = 0; level < tv.levels; ++level) {
        const int a = level % D;
        node = 2 * node + (q[a] < tv.split[node] ? 1 : 2);
    }
    const int32_t b = (node - tv.n_nodes) >> shift;
    bucket[i] = b;
    rank[i] = atomicAdd(&counts[b], 1);
}

__global__ void bucket_scatter_kernel(int64_t nq, const int32_t* __restrict__ bucket, const int32_t* __restrict__ rank,
                                      const int32_t* __restrict__ offsets, int32_t* __restrict__ order);

// Exclusive scan of n int32 counts (n <= kScanChunk * kScanChunk).
constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;
constexpr int kScanChunk = kScanBlock * kScanItems;
__global__ void scan_block_sums_kernel(const int32_t* __restrict__ in, int32_t n, int32_t* __restrict__ sums);
__global__ void scan_sums_kernel(int32_t* __restrict__ sums, int32_t nsums);
__global__ void scan_apply_kernel(const int32_t* __restrict__ in, int32_t n, const int32_t* __restrict__ sums,
                                  int32_t* __restrict__ out);

}  // namespace tk
