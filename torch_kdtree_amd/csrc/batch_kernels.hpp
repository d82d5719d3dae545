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
// Gathers in flight per lane in the backward kernels (measured: 8 against 4 took the fwd+bwd
// workload's backward from 1.007 to 0.969 ms and the 256-cloud replay from 20.82 to 20.60 ms;
// 2 was slower than 4).
#ifndef TK_BW_UNROLL
#define TK_BW_UNROLL 8
#endif
constexpr int kBwUnroll = TK_BW_UNROLL;  // backward loads issued ahead per lane

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

// Key of a (record, slot) pair for the dL/dr sort: the arena record position, or, with a tag shift,
// the cloud index above the tag shift and the record's position within its cloud below it (the
// sort then needs only the bits below the shift; equal low bits of different clouds stay apart,
// because each cloud's slots are one contiguous range and the sort is stable).
__device__ __forceinline__ uint32_t pair_key(int32_t cloud, int32_t local, int32_t rec_base, int tag_shift) {
    return tag_shift ? (((uint32_t)cloud << tag_shift) | (uint32_t)local) : (uint32_t)(rec_base + local);
}

// Arguments of the batched traversal shared by its per-bucket and persistent kernels.
template <typename T>
struct BatchArgs {
    TreeView<T, true> arena;
    const CloudDesc* __restrict__ clouds;
    const int32_t* __restrict__ qcloud;
    const T* __restrict__ queries;
    const int32_t* __restrict__ order;
    T* __restrict__ out_d;
    int32_t* __restrict__ out_i;
    int64_t* __restrict__ out_o;
    uint32_t* __restrict__ pair_keys;
    uint32_t* __restrict__ pair_vals;
    T* __restrict__ q_sorted;
    int tag_shift;
    const T* __restrict__ fuse_g;
    T* __restrict__ fuse_dq;
    T* __restrict__ g_sorted;
};

// The search of the query at sorted position tpos (in the k bucket K), with its outputs.
template <typename T, int D, int K, bool kLower, int CH, int LD>
__device__ __forceinline__ void batch_lane(const BatchArgs<T>& ba, int64_t tpos, T* __restrict__ st_d,
                                           int32_t* __restrict__ st_o) {
    const TreeView<T, true>& arena = ba.arena;
    const CloudDesc* __restrict__ clouds = ba.clouds;
    const T* __restrict__ queries = ba.queries;
    T* __restrict__ out_d = ba.out_d;
    int32_t* __restrict__ out_i = ba.out_i;
    int64_t* __restrict__ out_o = ba.out_o;
    uint32_t* __restrict__ pair_keys = ba.pair_keys;
    uint32_t* __restrict__ pair_vals = ba.pair_vals;
    T* __restrict__ q_sorted = ba.q_sorted;
    const int tag_shift = ba.tag_shift;
    const T* __restrict__ fuse_g = ba.fuse_g;
    T* __restrict__ fuse_dq = ba.fuse_dq;
    T* __restrict__ g_sorted = ba.g_sorted;
    const int64_t qi = ba.order[tpos];
    const int32_t cloud = ba.qcloud[qi];
    const CloudDesc cd = clouds[cloud];
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
    if (has_nan<T, D>(q))
        nan_fill<T, D, K, kLower>(tv, q, k, kb);
    else
        while (!lane_step<T, D, K, kLower, CH, LD>(tv, q, pp, kb, far_mask, node, level, st_d, st_o)) {
        }
    // keys: reference policy = arena record position, lower-index policy = the cloud's original index.
    // Every gather (index maps, and the neighbour records of the fused dL/dq) is issued for all K
    // slots before the first store or sum, so the K latencies overlap; padding slots read a valid
    // entry and are never stored.  A slot left empty (a query with a NaN coordinate) holds key 0:
    // under the reference policy that is taken as the cloud's first record (structured index 0,
    // as for a single tree), so that every index and pair key stays inside the cloud.
    const int pad = K - k;
    const int64_t ob = cd.out_base + (qi - cd.q_base) * k;
    // training outputs: this query at its sorted position, its pairs at their canonical slots
    const int64_t sb = cd.sorted_out_base + (tpos - cd.sorted_q_base) * k;
    if (q_sorted) {
#pragma unroll
        for (int a = 0; a < D; ++a) q_sorted[tpos * D + a] = q[a];
    }
    int32_t x[K], loc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        x[j] = j >= pad ? kb.key(j) : (kLower ? 0 : cd.rec_base);
        if (!kLower) x[j] = max(x[j], cd.rec_base);
        loc[j] = kLower ? load_i32_at(arena.inv_shuffled, cd.rec_base + x[j]) : x[j] - cd.rec_base;
    }
    int32_t orig[K] = {};
    if (!kLower && out_o) {
#pragma unroll
        for (int j = 0; j < K; ++j) orig[j] = load_i32_at(arena.shuffled, x[j]);
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j >= pad) {
            out_d[ob + j - pad] = kb.dist(j);
            out_i[ob + j - pad] = loc[j];
            if (out_o) out_o[ob + j - pad] = kLower ? x[j] : orig[j];
            if (pair_keys) {
                pair_keys[sb + j - pad] = pair_key(cloud, loc[j], cd.rec_base, tag_shift);
                pair_vals[sb + j - pad] = (uint32_t)(sb + j - pad);
            }
        }
    }
    // dL/dq of a training step whose upstream gradient is already in place (a captured step):
    // the terms of batch_grad_q_sorted_kernel, in neighbour order, without a second pass
    if (fuse_dq) {
        Rec<T, D> r[K];
        T gij[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            r[j] = load_rec<T, D>(arena.recs, cd.rec_base + loc[j]);
            gij[j] = j >= pad ? fuse_g[ob + j - pad] : T(0);
        }
        T acc[D];
#pragma unroll
        for (int a = 0; a < D; ++a) acc[a] = T(0);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (j >= pad) {
                if (g_sorted) g_sorted[sb + j - pad] = gij[j];
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const T diff = q[a] - r[j].c[a];
                    acc[a] = acc[a] + gij[j] * (diff + diff);
                }
            }
        }
#pragma unroll
        for (int a = 0; a < D; ++a) fuse_dq[qi * D + a] = acc[a];
    }
}

// One lane per query of the k bucket K (sorted positions begin .. begin + count).
template <typename T, int D, int K, bool kLower>
__global__ __launch_bounds__(kTreeBlock) TK_BATCH_OCCUPANCY(T, D, K, kLower) void knn_batch_kernel(
    BatchArgs<T> ba, int64_t begin, int64_t count) {
    constexpr int CH = StageBytes<T, D, K, kLower, true>::kChunk;
    constexpr int LD = StageBytes<T, D, K, kLower, true>::kLoads;
    static_assert(CH <= kRecPad, "a chunk may read at most kRecPad records past a leaf's end");
    __shared__ T stage_d[CH * kTreeBlock];
    __shared__ int32_t stage_o[kLower ? CH * kTreeBlock : 1];
    const int64_t t = xcd_tile(blockIdx.x, gridDim.x) * kTreeBlock + threadIdx.x;
    if (t >= count) return;
    batch_lane<T, D, K, kLower, CH, LD>(ba, begin + t, stage_d + threadIdx.x, stage_o + (kLower ? threadIdx.x : 0));
}

// ---------------------------------------------------------------------------------------------
// Persistent traversal: one launch for every k bucket.  The sorted order is cut into units of 64
// queries (one wavefront), each inside one bucket, largest lists first.  Units are dealt to the
// kXcdQueues XCDs (workgroup b runs on XCD b % 8) in runs of kUnitRun consecutive units, run r
// going to XCD r % 8, so every XCD gets a share of every bucket and neighbouring units, which read
// the same leaves, share one L2.  Each XCD has a counter over its own units; a wavefront takes the
// next one with one atomic add, and once its XCD's units are exhausted those of the other XCDs in
// turn.  Every wavefront loops until all are exhausted, so the launch drains.  The bucket of a
// unit is wavefront-uniform.
// ---------------------------------------------------------------------------------------------
constexpr int kXcdQueues = 8;
constexpr int kBatchMaxRanges = 6;
constexpr int64_t kUnitRun = 256;

// The c-th unit of XCD x.
__host__ __device__ __forceinline__ int64_t xcd_unit(int x, int64_t c) {
    return ((c / kUnitRun) * kXcdQueues + x) * kUnitRun + c % kUnitRun;
}

struct BatchUnits {
    int n_ranges;
    int K[kBatchMaxRanges];
    int64_t begin[kBatchMaxRanges];  // sorted position of the range's first query
    int64_t count[kBatchMaxRanges];
    int64_t unit0[kBatchMaxRanges];  // first unit of the range
    int64_t n_units;
};

// Work counters of the persistent traversal (one per XCD) back to zero.
__global__ void batch_units_reset_kernel(uint32_t* __restrict__ counters) {
    if (threadIdx.x < kXcdQueues) counters[threadIdx.x] = 0;
}

template <typename T, int D, int KMAX, bool kLower>
__global__ __launch_bounds__(kTreeBlock) __attribute__((amdgpu_waves_per_eu(TK_BATCH_WAVES_K16 > 0 ? TK_BATCH_WAVES_K16 : 1))) void knn_batch_persistent_kernel(
    BatchArgs<T> ba, BatchUnits u, uint32_t* __restrict__ counters) {
    constexpr int CHMAX = StageBytes<T, D, KMAX, kLower, true>::kChunk > StageBytes<T, D, 1, kLower, true>::kChunk
                              ? StageBytes<T, D, KMAX, kLower, true>::kChunk
                              : StageBytes<T, D, 1, kLower, true>::kChunk;
    __shared__ T stage_d[CHMAX * kTreeBlock];
    __shared__ int32_t stage_o[kLower ? CHMAX * kTreeBlock : 1];
    const int lane = threadIdx.x & 63;
    const int home = blockIdx.x % kXcdQueues;
    T* st_d = stage_d + threadIdx.x;
    int32_t* st_o = stage_o + (kLower ? threadIdx.x : 0);
    for (int probe = 0; probe < kXcdQueues;) {
        const int x = (home + probe) % kXcdQueues;
        uint32_t got = 0;
        if (lane == 0) got = atomicAdd(&counters[x], 1u);
        const int64_t unit = xcd_unit(x, (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)got));
        if (unit >= u.n_units) {
            ++probe;  // this XCD's units are exhausted: take the next XCD's
            continue;
        }
        int r = 0;
        while (r + 1 < u.n_ranges && unit >= u.unit0[r + 1]) ++r;
        const int64_t t = (unit - u.unit0[r]) * 64 + lane;
        if (t < u.count[r]) {
            const int64_t tpos = u.begin[r] + t;
#define TK_PERSIST_K(KK)                                                                                          \
    case KK:                                                                                                      \
        if constexpr (KK <= KMAX)                                                                                 \
            batch_lane<T, D, KK, kLower, StageBytes<T, D, KK, kLower, true>::kChunk,                              \
                       StageBytes<T, D, KK, kLower, true>::kLoads>(ba, tpos, st_d, st_o);                         \
        break;
            switch (u.K[r]) {
                TK_PERSIST_K(1)
                TK_PERSIST_K(2)
                TK_PERSIST_K(4)
                TK_PERSIST_K(8)
                TK_PERSIST_K(16)
                TK_PERSIST_K(32)
                default: break;
            }
#undef TK_PERSIST_K
        }
    }
}

__global__ __launch_bounds__(kBatchBlock) void batch_rank_kernel(const int32_t* __restrict__ order, int64_t n,
                                                                  uint32_t* __restrict__ rank) {
    const int64_t t = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
    if (t < n) rank[order[t]] = (uint32_t)t;
}

// dL/dq, one lane per query row, terms added in neighbour order (the single-tree grad_q_kernel's
// arithmetic and order).  With q_sorted (for dL/dr) it also copies the query to its sorted
// position t = rank[i] and, per neighbour j, the pair (arena record, slot) to the canonical slot
// sorted_out_base + (t - sorted_q_base) * k + j.
template <typename T, int D>
__global__ __launch_bounds__(kBatchBlock) void batch_grad_q_pairs_kernel(
    const unsigned char* __restrict__ recs, const CloudDesc* __restrict__ clouds, const int32_t* __restrict__ qcloud,
    const uint32_t* __restrict__ rank, const T* __restrict__ queries, int64_t total_q, const int32_t* __restrict__ idx,
    const T* __restrict__ g, T* __restrict__ dq, T* __restrict__ q_sorted, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ vals, int tag_shift) {
    const int64_t i = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
    if (i >= total_q) return;
    const int32_t cloud = qcloud[i];
    const CloudDesc cd = clouds[cloud];
    const int64_t ob = cd.out_base + (i - cd.q_base) * cd.k;
    T qi[D], acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        qi[a] = queries[i * D + a];
        acc[a] = T(0);
    }
    const int64_t t = q_sorted ? (int64_t)rank[i] : 0;
    const int64_t sb = cd.sorted_out_base + (t - cd.sorted_q_base) * cd.k;
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
            keys[sb + j] = pair_key(cloud, idx[ob + j], cd.rec_base, tag_shift);
            vals[sb + j] = (uint32_t)(sb + j);
        }
    }
    if (dq) {
#pragma unroll
        for (int a = 0; a < D; ++a) dq[i * D + a] = acc[a];
    }
}

// dL/dq after a forward that wrote the training outputs: one lane per sorted position t (query
// order[t], read from q_sorted), terms added in neighbour order.
template <typename T, int D>
__global__ __launch_bounds__(kBatchBlock) void batch_grad_q_sorted_kernel(
    const unsigned char* __restrict__ recs, const CloudDesc* __restrict__ clouds, const int32_t* __restrict__ qcloud,
    const int32_t* __restrict__ order, const T* __restrict__ q_sorted, int64_t total_q,
    const int32_t* __restrict__ idx, const T* __restrict__ g, T* __restrict__ dq) {
    const int64_t t = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
    if (t >= total_q) return;
    const int64_t i = order[t];
    const CloudDesc cd = clouds[qcloud[i]];
    const int64_t ob = cd.out_base + (i - cd.q_base) * cd.k;
    T qi[D], acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        qi[a] = q_sorted[t * D + a];
        acc[a] = T(0);
    }
    // loads kBwUnroll neighbours ahead; the terms are still added in neighbour order
    for (int j0 = 0; j0 < cd.k; j0 += kBwUnroll) {
        int32_t id[kBwUnroll];
        T gg[kBwUnroll];
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
            const bool ok = j0 + u < cd.k;
            id[u] = ok ? idx[ob + j0 + u] : 0;
            gg[u] = ok ? g[ob + j0 + u] : T(0);
        }
        Rec<T, D> r[kBwUnroll];
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) r[u] = load_rec<T, D>(recs, cd.rec_base + id[u]);
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
            if (j0 + u < cd.k) {
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const T diff = qi[a] - r[u].c[a];
                    acc[a] = acc[a] + gg[u] * (diff + diff);
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) dq[i * D + a] = acc[a];
}

__global__ __launch_bounds__(kBatchBlock) void batch_zero_runs_kernel(uint32_t* __restrict__ a,
                                                                       uint32_t* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
    if (i < n) {
        a[i] = 0;
        b[i] = 0;
    }
}

// Bounds of each record's run in the sorted pairs (zero for records without pairs).  A run is a
// maximal range of equal keys (whole keys, tag included).
constexpr int kRunItems = 4;
__global__ __launch_bounds__(kBatchBlock) void batch_run_bounds_kernel(const uint32_t* __restrict__ keys, int64_t m,
                                                                        const CloudDesc* __restrict__ clouds,
                                                                        int tag_shift,
                                                                        uint32_t* __restrict__ run_start,
                                                                        uint32_t* __restrict__ run_end) {
    // each lane looks at kRunItems consecutive pairs, with one key on either side
    const int64_t p0 = ((int64_t)blockIdx.x * kBatchBlock + threadIdx.x) * kRunItems;
    if (p0 >= m) return;
    uint32_t k[kRunItems + 2];
#pragma unroll
    for (int u = 0; u < kRunItems + 2; ++u) {
        const int64_t p = p0 - 1 + u;
        k[u] = (p >= 0 && p < m) ? keys[p] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kRunItems; ++u) {
        const int64_t p = p0 + u;
        if (p >= m) break;
        const uint32_t key = k[u + 1];
        const bool first = p == 0 || k[u] != key;
        const bool last = p == m - 1 || k[u + 2] != key;
        if (!first && !last) continue;
        const uint32_t rec = tag_shift ? (uint32_t)clouds[key >> tag_shift].rec_base + (key & ((1u << tag_shift) - 1u))
                                       : key;
        if (first) run_start[rec] = (uint32_t)p;
        if (last) run_end[rec] = (uint32_t)(p + 1);
    }
}

// dL/dr, one lane per record: its run of sorted pairs (canonical order), written to the record's
// original row.  A slot's upstream gradient is read from the caller's layout through the order, or,
// with kSlots, from g_sorted, where the captured step's forward stored it by canonical slot.
template <typename T, int D, bool kSlots>
__global__ __launch_bounds__(kBatchBlock) void batch_grad_r_kernel(const unsigned char* __restrict__ recs,
                                                                    const int32_t* __restrict__ rec_cloud,
                                                                    const int32_t* __restrict__ shuffled,
                                                                    const CloudDesc* __restrict__ clouds,
                                                                    int64_t total_ref,
                                                                    const uint32_t* __restrict__ run_start,
                                                                    const uint32_t* __restrict__ run_end,
                                                                    const uint32_t* __restrict__ vals,
                                                                    const T* __restrict__ q_sorted,
                                                                    const int32_t* __restrict__ order,
                                                                    const T* __restrict__ g, T* __restrict__ dr) {
    // g: the caller's upstream gradient, or with kSlots the gradient by canonical slot
    const int64_t s = (int64_t)blockIdx.x * kBatchBlock + threadIdx.x;
    if (s >= total_ref) return;
    const CloudDesc cd = clouds[rec_cloud[s]];
    const Rec<T, D> r = load_rec<T, D>(recs, (int32_t)s);
    const uint32_t e0 = run_start[s], e1 = run_end[s];
    T acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] = T(0);
    // loads kBwUnroll pairs ahead; the terms are still subtracted in run order
    for (uint32_t e0u = e0; e0u < e1; e0u += kBwUnroll) {
        int64_t t[kBwUnroll], jj[kBwUnroll];
        uint32_t slot[kBwUnroll];
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
            const bool ok = e0u + u < e1;
            slot[u] = ok ? vals[e0u + u] : (uint32_t)cd.sorted_out_base;
            const uint32_t off = (uint32_t)((int64_t)slot[u] - cd.sorted_out_base);
            t[u] = cd.sorted_q_base + off / (uint32_t)cd.k;
            jj[u] = off % (uint32_t)cd.k;
        }
        int64_t i[kBwUnroll];
        if constexpr (!kSlots) {
#pragma unroll
            for (int u = 0; u < kBwUnroll; ++u) i[u] = e0u + u < e1 ? (int64_t)order[t[u]] : cd.q_base;
        }
        T gg[kBwUnroll], qq[kBwUnroll][D];
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
            const bool ok = e0u + u < e1;
            if constexpr (kSlots)
                gg[u] = ok ? g[slot[u]] : T(0);
            else
                gg[u] = ok ? g[cd.out_base + (i[u] - cd.q_base) * cd.k + jj[u]] : T(0);
#pragma unroll
            for (int a = 0; a < D; ++a) qq[u][a] = ok ? q_sorted[t[u] * D + a] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
            if (e0u + u < e1) {
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const T diff = qq[u][a] - r.c[a];
                    acc[a] = acc[a] - gg[u] * (diff + diff);
                }
            }
        }
    }
    const int64_t row = (int64_t)cd.rec_base + shuffled[s];
#pragma unroll
    for (int a = 0; a < D; ++a) dr[row * D + a] = acc[a];
}

template <typename T>
TreeView<T, true> batch_arena_view(const BatchPlan& p) {
    TreeView<T, true> v{static_cast<const T*>(p.split), p.recs, p.leaf_off, p.shuffled, p.inv_shuffled, p.boxes, 0, 0};
    v.split_bounds = p.all_finite;
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
    hipLaunchKernelGGL(batch_rank_kernel, batch_grid(p.total_q), dim3(kBatchBlock), 0, s, p.order, p.total_q, p.rank);
    TK_HIP_CHECK(hipGetLastError());
}

template <typename T, int D, bool kLower>
void batch_traverse_td(const BatchPlan& p, const T* q, T* dist, int32_t* idx, int64_t* orig, bool emit,
                       const T* g, T* dq, hipStream_t s) {
    // the captured training step also stores the upstream gradient by canonical slot
    T* gs = (emit && g && dq && p.g_sorted) ? static_cast<T*>(p.g_sorted) : nullptr;
    uint32_t* pk = emit ? p.keys : nullptr;
    uint32_t* pv = emit ? p.vals : nullptr;
    T* qs = emit ? static_cast<T*>(p.q_sorted) : nullptr;
    const BatchArgs<T> ba{batch_arena_view<T>(p), p.d_clouds, p.qcloud, q, p.order, dist, idx, orig, pk, pv, qs,
                          p.pair_tag_shift, g, dq, gs};
    if (p.persistent) {
        // units of 64 queries per range, largest lists first (the longest work starts first)
        BatchUnits u{};
        int kmax = 1;
        for (int i = (int)p.ranges.size() - 1; i >= 0; --i) {
            const BatchPlan::Range& r = p.ranges[(size_t)i];
            if (r.count == 0) continue;
            u.K[u.n_ranges] = r.K;
            u.begin[u.n_ranges] = r.begin;
            u.count[u.n_ranges] = r.count;
            u.unit0[u.n_ranges] = u.n_units;
            u.n_units += (r.count + 63) / 64;
            kmax = std::max(kmax, r.K);
            ++u.n_ranges;
        }
        if (u.n_units == 0) return;
        if (u.n_units >= (int64_t(1) << 31)) throw std::runtime_error("batch: too many units for the persistent launch");
        hipLaunchKernelGGL(batch_units_reset_kernel, dim3(1), dim3(64), 0, s, p.work);
        int dev = 0, cus = 0;
        TK_HIP_CHECK(hipGetDevice(&dev));
        TK_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        // a resident grid: 8 workgroups of 128 threads per CU (4 waves per SIMD), a multiple of 8
        const unsigned blocks = (unsigned)std::max(8, (cus * 8 + 7) / 8 * 8);
#define TK_PERSIST_LAUNCH(KM)                                                                                     \
    if (kmax <= KM) {                                                                                             \
        hipLaunchKernelGGL((knn_batch_persistent_kernel<T, D, KM, kLower>), dim3(blocks), dim3(kTreeBlock), 0, s,  \
                           ba, u, p.work);                                                                        \
        TK_HIP_CHECK(hipGetLastError());                                                                          \
        return;                                                                                                   \
    }
        TK_PERSIST_LAUNCH(4)
        TK_PERSIST_LAUNCH(16)
        TK_PERSIST_LAUNCH(32)
#undef TK_PERSIST_LAUNCH
        throw std::runtime_error("batch: k bucket out of range");
    }
    for (const BatchPlan::Range& r : p.ranges) {
        if (r.count == 0) continue;
        const dim3 grid((unsigned)((r.count + kTreeBlock - 1) / kTreeBlock));
#define TK_LAUNCH_K(KK)                                                                                              \
    case KK:                                                                                                         \
        hipLaunchKernelGGL((knn_batch_kernel<T, D, KK, kLower>), grid, dim3(kTreeBlock), 0, s, ba, r.begin, r.count); \
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
void batch_traverse_t(const BatchPlan& p, const T* q, T* dist, int32_t* idx, int64_t* orig, bool emit,
                      const T* g, T* dq, hipStream_t s) {
#define TK_DIMS(DD)                                                         \
    case DD:                                                                \
        if (p.tie == kTieLowerIndex)                                        \
            batch_traverse_td<T, DD, true>(p, q, dist, idx, orig, emit, g, dq, s); \
        else                                                                \
            batch_traverse_td<T, DD, false>(p, q, dist, idx, orig, emit, g, dq, s);\
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

// from_forward: the forward wrote the training outputs (pairs, q_sorted) for these queries and
// the order is current; otherwise the order is recomputed by the caller and the pairs are made
// here.
template <typename T>
void batch_backward_t(const BatchPlan& p, const T* q, const int32_t* idx, const T* g, T* dq, T* dr, bool from_forward,
                      hipStream_t s) {
    const bool pairs = dr && p.total_out > 0;
    T* qs = static_cast<T*>(p.q_sorted);
#define TK_DIMS(DD)                                                                                                  \
    case DD:                                                                                                         \
        if (dr && !pairs) {                                                                                          \
            hipLaunchKernelGGL((batch_zero_kernel<T>), batch_grid(p.total_ref * DD), dim3(kBatchBlock), 0, s, dr,    \
                               p.total_ref * DD);                                                                    \
            TK_HIP_CHECK(hipGetLastError());                                                                         \
        }                                                                                                            \
        if (p.total_q > 0 && from_forward && dq) {                                                                   \
            hipLaunchKernelGGL((batch_grad_q_sorted_kernel<T, DD>), batch_grid(p.total_q), dim3(kBatchBlock), 0, s,  \
                               p.recs, p.d_clouds, p.qcloud, p.order, qs, p.total_q, idx, g, dq);                    \
            TK_HIP_CHECK(hipGetLastError());                                                                         \
        } else if (p.total_q > 0 && !from_forward && (dq || pairs)) {                                                \
            hipLaunchKernelGGL((batch_grad_q_pairs_kernel<T, DD>), batch_grid(p.total_q), dim3(kBatchBlock), 0, s,   \
                               p.recs, p.d_clouds, p.qcloud, p.rank, q, p.total_q, idx, g, dq,                       \
                               pairs ? qs : nullptr, p.keys, p.vals, p.pair_tag_shift);                              \
            TK_HIP_CHECK(hipGetLastError());                                                                         \
        }                                                                                                            \
        if (pairs) {                                                                                                 \
            radix_sort_pairs(p.keys, p.vals, p.keys_s, p.vals_s, p.total_out, p.pair_bits,                           \
                             p.radix_ws, p.radix_ws_bytes, s);                                                       \
            hipLaunchKernelGGL(batch_zero_runs_kernel, batch_grid(p.total_ref), dim3(kBatchBlock), 0, s,             \
                               p.run_start, p.run_end, p.total_ref);                                                 \
            hipLaunchKernelGGL(batch_run_bounds_kernel, batch_grid((p.total_out + kRunItems - 1) / kRunItems),      \
                               dim3(kBatchBlock), 0, s, p.keys_s,                                                    \
                               p.total_out, p.d_clouds, p.pair_tag_shift, p.run_start, p.run_end);                   \
            if (from_forward && p.g_sorted && p.g_sorted_valid)                                                      \
                hipLaunchKernelGGL((batch_grad_r_kernel<T, DD, true>), batch_grid(p.total_ref), dim3(kBatchBlock), 0, \
                                   s, p.recs, p.rec_cloud, p.shuffled, p.d_clouds, p.total_ref, p.run_start,         \
                                   p.run_end, p.vals_s, qs, p.order, static_cast<const T*>(p.g_sorted), dr);         \
            else                                                                                                     \
                hipLaunchKernelGGL((batch_grad_r_kernel<T, DD, false>), batch_grid(p.total_ref), dim3(kBatchBlock),  \
                                   0, s, p.recs, p.rec_cloud, p.shuffled, p.d_clouds, p.total_ref, p.run_start,      \
                                   p.run_end, p.vals_s, qs, p.order, g, dr);                                         \
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
