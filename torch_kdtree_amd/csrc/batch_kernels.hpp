// Kernels of the batched many-cloud mode and their per-dtype launchers, included once per dtype
// (batch_f32.hip, batch_f64.hip).  The traversal is the single-tree one (knn_tree.hpp) run on a
// TreeView whose per-lane offsets select the lane's tree in the arena.
#pragma once

#include "batch.hpp"
#include "knn_gpu.hpp"
#include "knn_tree.hpp"

namespace tk {
// Internal linkage: each dtype's translation unit has its own copy.
namespace {

constexpr int kBatchBlock = 256;

template <typename T>
__global__ __launch_bounds__(kBatchBlock) void batch_zero_kernel(T* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
    if (i < n) p[i] = T(0);
}

// Presort key of every query: the cloud's first cell plus the Morton code of the query's cell
// over the cloud's root box (the bounding box of its points), with the cloud's canonical grid.
template <typename T, int D>
__global__ __launch_bounds__(kBatchBlock) void batch_bucket_kernel(const unsigned char* __restrict__ boxes,
                                                                    const CloudDesc* __restrict__ clouds,
                                                                    const int32_t* __restrict__ qcloud,
                                                                    const T* __restrict__ queries, int64_t total_q,
                                                                    uint32_t* __restrict__ key,
                                                                    uint32_t* __restrict__ row) {
    const int64_t i = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
    if (i >= total_q) return;
    const CloudDesc& cd = clouds[qcloud[i]];
    using L = BoxLayout<T, D>;
    const T* box = reinterpret_cast<const T*>(boxes + (size_t)cd.box_base * L::kBytes);
    float lo[D], hi[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        lo[a] = (float)box[a];
        hi[a] = (float)box[L::kPad + a];
    }
    key[i] = (uint32_t)cd.bucket_base + morton_cell<T, D>(lo, hi, cd.morton_bits, queries + i * D);
    row[i] = (uint32_t)i;
}

// One lane per query of the k bucket K; `order` holds the presorted query rows.
template <typename T, int D, int K, bool kLower>
__global__ __launch_bounds__(kTreeBlock) void knn_batch_kernel(TreeView<T, true> arena,
                                                                const CloudDesc* __restrict__ clouds,
                                                                const int32_t* __restrict__ qcloud,
                                                                const T* __restrict__ queries,
                                                                const int32_t* __restrict__ order, int64_t begin,
                                                                int64_t count, T* __restrict__ out_d,
                                                                int32_t* __restrict__ out_i, int64_t* __restrict__ out_o) {
    constexpr int CH = StageBytes<T, K, kLower>::kChunk;
    constexpr int LD = StageBytes<T, K, kLower>::kLoads;
    static_assert(LD <= kRecPad, "a load group may read at most kRecPad records past a leaf's end");
    __shared__ T stage_d[CH * kTreeBlock];
    __shared__ int32_t stage_o[kLower ? CH * kTreeBlock : 1];
    const int64_t t = (int64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (t >= count) return;
    const int64_t qi = order[begin + t];
    const CloudDesc cd = clouds[qcloud[qi]];
    TreeView<T, true> tv = arena;
    tv.node_base = cd.node_base;
    tv.box_base = cd.box_base;
    tv.leaf_base = cd.leaf_base;
    tv.rec_base = cd.rec_base;
    tv.levels = cd.levels;
    tv.n_nodes = cd.n_nodes;
    const int k = cd.k;
    Pt<T, D> q, pp;
    KBest<T, K, kLower> kb;
    uint32_t far_mask;
    int32_t node;
    int level;
    load_query<T, D>(queries, qi, q);
    lane_start<T, D, K, kLower>(tv, k, q, pp, kb, far_mask, node, level);
    while (!lane_step<T, D, K, kLower, CH, LD>(tv, q, pp, kb, far_mask, node, level, stage_d + threadIdx.x,
                                                stage_o + (kLower ? threadIdx.x : 0))) {
    }
    // keys: reference policy = arena record position, lower-index policy = the cloud's original index
    const int pad = K - k;
    const int64_t ob = cd.out_base + (qi - cd.q_base) * k;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j >= pad) {
            const int32_t x = kb.x[j];
            out_d[ob + j - pad] = kb.d[j];
            out_i[ob + j - pad] = kLower ? load_i32_at(arena.inv_shuffled, cd.rec_base + x) : x - cd.rec_base;
            if (out_o) out_o[ob + j - pad] = kLower ? x : load_i32_at(arena.shuffled, x);
        }
    }
}

// dL/dq, and for dL/dr the canonical copies and pairs.  Lane t takes query order[t] (the sorted
// order of the forward, the canonical order within each cloud); with q_sorted it writes the query
// to position t and, per neighbour j, the upstream gradient and the pair (arena record, slot) to
// the canonical slot sorted_out_base + (t - sorted_q_base) * k + j.  dL/dq adds the terms in
// neighbour order (the single-tree grad_q_kernel's arithmetic and order).
template <typename T, int D>
__global__ __launch_bounds__(kBatchBlock) void batch_grad_q_pairs_kernel(
    const unsigned char* __restrict__ recs, const CloudDesc* __restrict__ clouds, const int32_t* __restrict__ qcloud,
    const int32_t* __restrict__ order, const T* __restrict__ queries, int64_t total_q, const int32_t* __restrict__ idx,
    const T* __restrict__ g, T* __restrict__ dq, T* __restrict__ q_sorted, T* __restrict__ g_sorted,
    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t t = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
    if (t >= total_q) return;
    const int64_t i = order[t];
    const CloudDesc cd = clouds[qcloud[i]];
    const int64_t ob = cd.out_base + (i - cd.q_base) * cd.k;
    const int64_t sb = cd.sorted_out_base + (t - cd.sorted_q_base) * cd.k;
    T qi[D], acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        qi[a] = queries[i * D + a];
        acc[a] = T(0);
    }
    if (q_sorted) {
#pragma unroll
        for (int a = 0; a < D; ++a) q_sorted[t * D + a] = qi[a];
    }
    for (int j = 0; j < cd.k; ++j) {
        const int32_t s = cd.rec_base + idx[ob + j];
        const Rec<T, D> r = load_rec<T, D>(recs, s);
        const T gij = g[ob + j];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const T diff = qi[a] - r.c[a];
            acc[a] = acc[a] + gij * (diff + diff);
        }
        if (q_sorted) {
            g_sorted[sb + j] = gij;
            keys[sb + j] = (uint32_t)s;
            vals[sb + j] = (uint32_t)(sb + j);
        }
    }
    if (dq) {
#pragma unroll
        for (int a = 0; a < D; ++a) dq[i * D + a] = acc[a];
    }
}

// dL/dr: one lane per run of equal records in the sorted pairs; the run is in canonical order.
template <typename T, int D>
__global__ __launch_bounds__(kBatchBlock) void batch_grad_r_kernel(const unsigned char* __restrict__ recs,
                                                                    const int32_t* __restrict__ rec_cloud,
                                                                    const int32_t* __restrict__ shuffled,
                                                                    const CloudDesc* __restrict__ clouds, int64_t m,
                                                                    const uint32_t* __restrict__ keys,
                                                                    const uint32_t* __restrict__ vals,
                                                                    const T* __restrict__ q_sorted,
                                                                    const T* __restrict__ g_sorted,
                                                                    T* __restrict__ dr) {
    const int64_t p = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
    if (p >= m) return;
    const uint32_t key = keys[p];
    if (p > 0 && keys[p - 1] == key) return;  // not the head of its run
    const CloudDesc cd = clouds[rec_cloud[key]];
    const Rec<T, D> r = load_rec<T, D>(recs, (int32_t)key);
    T acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = T(0);
    for (int64_t e = p; e < m && keys[e] == key; ++e) {
        const int64_t v = vals[e];
        const int64_t t = cd.sorted_q_base + (v - cd.sorted_out_base) / cd.k;
        const T gij = g_sorted[v];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const T diff = q_sorted[t * D + a] - r.c[a];
            acc[a] = acc[a] - gij * (diff + diff);
        }
    }
    const int64_t row = (int64_t)cd.rec_base + shuffled[key];
#pragma unroll
    for (int a = 0; a < D; ++a) dr[row * D + a] = acc[a];
}

template <typename T>
TreeView<T, true> batch_arena_view(const BatchPlan& p) {
    TreeView<T, true> v{static_cast<const T*>(p.split), p.recs, p.leaf_off, p.shuffled, p.inv_shuffled, p.boxes, 0, 0};
    return v;
}

inline dim3 batch_grid(int64_t n) { return dim3((unsigned)((n + kBatchBlock - 1) / kBatchBlock)); }

template <typename T>
void batch_presort_t(const BatchPlan& p, const T* q, hipStream_t s) {
#define TK_DIMS(DD)                                                                                                  \
    case DD:                                                                                                         \
        hipLaunchKernelGGL((batch_bucket_kernel<T, DD>), batch_grid(p.total_q), dim3(kBatchBlock), 0, s, p.boxes,   \
                           p.d_clouds, p.qcloud, q, p.total_q, p.key, p.row);                                        \
        break;
    switch (p.dims) {
        TK_DIMS(1)
        TK_DIMS(2)
        TK_DIMS(3)
        TK_DIMS(8)
        default: throw std::runtime_error("unsupported dims");
    }
#undef TK_DIMS
    TK_HIP_CHECK(hipGetLastError());
    radix_sort_pairs(p.key, p.row, p.key_s, reinterpret_cast<uint32_t*>(p.order), p.total_q, p.key_bits, p.radix_ws,
                     p.radix_ws_bytes, s);
}

template <typename T, int D, bool kLower>
void batch_traverse_td(const BatchPlan& p, const T* q, T* dist, int32_t* idx, int64_t* orig, hipStream_t s) {
    const TreeView<T, true> arena = batch_arena_view<T>(p);
    for (const BatchPlan::Range& r : p.ranges) {
        if (r.count == 0) continue;
        const dim3 grid((unsigned)((r.count + kTreeBlock - 1) / kTreeBlock));
#define TK_LAUNCH_K(KK)                                                                                              \
    case KK:                                                                                                         \
        hipLaunchKernelGGL((knn_batch_kernel<T, D, KK, kLower>), grid, dim3(kTreeBlock), 0, s, arena, p.d_clouds,   \
                           p.qcloud, q, p.order, r.begin, r.count, dist, idx, orig);                                 \
        break;
        switch (r.K) {
            TK_LAUNCH_K(1)
            TK_LAUNCH_K(2)
            TK_LAUNCH_K(4)
            TK_LAUNCH_K(8)
            TK_LAUNCH_K(16)
            TK_LAUNCH_K(32)
            default: throw std::runtime_error("batch: k bucket out of range");
        }
#undef TK_LAUNCH_K
        TK_HIP_CHECK(hipGetLastError());
    }
}

template <typename T>
void batch_traverse_t(const BatchPlan& p, const T* q, T* dist, int32_t* idx, int64_t* orig, hipStream_t s) {
#define TK_DIMS(DD)                                                         \
    case DD:                                                                \
        if (p.tie == kTieLowerIndex)                                        \
            batch_traverse_td<T, DD, true>(p, q, dist, idx, orig, s);       \
        else                                                                \
            batch_traverse_td<T, DD, false>(p, q, dist, idx, orig, s);      \
        break;
    switch (p.dims) {
        TK_DIMS(1)
        TK_DIMS(2)
        TK_DIMS(3)
        TK_DIMS(8)
        default: throw std::runtime_error("unsupported dims");
    }
#undef TK_DIMS
}

inline int batch_key_bits(int64_t n) {
    int b = 1;
    while (b < 32 && (int64_t(1) << b) < n) ++b;
    return b;
}

template <typename T>
void batch_backward_t(const BatchPlan& p, const T* q, const int32_t* idx, const T* g, T* dq, T* dr, hipStream_t s) {
    const bool pairs = dr && p.total_out > 0;
    T* qs = pairs ? static_cast<T*>(p.q_sorted) : nullptr;
    T* gs = pairs ? static_cast<T*>(p.g_sorted) : nullptr;
#define TK_DIMS(DD)                                                                                                  \
    case DD:                                                                                                         \
        if (dr) {                                                                                                    \
            hipLaunchKernelGGL((batch_zero_kernel<T>), batch_grid(p.total_ref * DD), dim3(kBatchBlock), 0, s, dr,    \
                               p.total_ref * DD);                                                                    \
            TK_HIP_CHECK(hipGetLastError());                                                                         \
        }                                                                                                            \
        if ((dq || pairs) && p.total_q > 0) {                                                                        \
            hipLaunchKernelGGL((batch_grad_q_pairs_kernel<T, DD>), batch_grid(p.total_q), dim3(kBatchBlock), 0, s,   \
                               p.recs, p.d_clouds, p.qcloud, p.order, q, p.total_q, idx, g, dq, qs, gs, p.keys,      \
                               p.vals);                                                                              \
            TK_HIP_CHECK(hipGetLastError());                                                                         \
        }                                                                                                            \
        if (pairs) {                                                                                                 \
            radix_sort_pairs(p.keys, p.vals, p.keys_s, p.vals_s, p.total_out, batch_key_bits(p.total_ref),           \
                             p.radix_ws, p.radix_ws_bytes, s);                                                       \
            hipLaunchKernelGGL((batch_grad_r_kernel<T, DD>), batch_grid(p.total_out), dim3(kBatchBlock), 0, s,       \
                               p.recs, p.rec_cloud, p.shuffled, p.d_clouds, p.total_out, p.keys_s, p.vals_s, qs, gs, \
                               dr);                                                                                  \
            TK_HIP_CHECK(hipGetLastError());                                                                         \
        }                                                                                                            \
        break;
    switch (p.dims) {
        TK_DIMS(1)
        TK_DIMS(2)
        TK_DIMS(3)
        TK_DIMS(8)
        default: throw std::runtime_error("unsupported dims");
    }
#undef TK_DIMS
}

}  // namespace
}  // namespace tk
