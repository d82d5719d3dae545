// Batched many-cloud mode: many trees of one dims/dtype queried together.
//
// A plan packs the trees into one device arena (records, split values, point boxes, leaf offsets,
// permutations) and describes each cloud by a CloudDesc.  Queries of all clouds are one row-major
// [sum Q, dims] array (cloud c owns rows q_base .. q_base + n_queries); outputs are one array per
// kind (distances, structured int32 indices, original int64 indices), cloud c owning the
// n_queries x k elements from out_base.  Per query the plan knows its cloud (qcloud).
//
// Forward: one presort over all queries: stable sort by Morton cell over the cloud's root box
// (the bounding box of its points) with canonical_morton_bits(Q_c, d) bits, each cloud's cells
// placed so that the order groups clouds by their k bucket (K = next power of two >= k); then
// one traversal launch per k bucket over its contiguous range of the order.  Within a cloud this
// order is the single-tree backward's canonical order.  Backward: dL/dq per query, and dL/dr
// through one stable sort of (reference record, canonical slot) pairs and per-run sums, which is
// exactly the summation order of the single-tree backward.  All scratch memory is allocated with
// the plan, so a forward + backward sequence can be captured in a HIP graph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "tree.hpp"

namespace tk {

struct CloudDesc {
    int32_t rec_base;   // first record (= first reference row) of the cloud in the arena
    int32_t node_base;  // first split value
    int32_t box_base;   // first box (heap node 0)
    int32_t leaf_base;  // first leaf offset entry
    int32_t levels;
    int32_t n_nodes;
    int32_t n_points;
    int32_t k;
    int64_t q_base;     // first query row
    int64_t out_base;   // first output element
    int32_t n_queries;
    int32_t bucket_base;  // first Morton bucket of the cloud
    int32_t morton_bits;  // bits per axis of the cloud's Morton grid (canonical_morton_bits)
    int32_t kbucket;      // K: the register list size its queries run with
    int64_t sorted_q_base;    // first position of the cloud's queries in the sorted order
    int64_t sorted_out_base;  // first canonical slot of the cloud's (query, neighbour) pairs
};

constexpr int kBatchMaxK = 32;

struct BatchPlan {
    int device = -1;
    int dtype = 0, dims = 0, tie = 0;
    int n_clouds = 0;
    bool all_finite = true;  // every cloud's coordinates and split values are finite
    std::vector<CloudDesc> clouds;
    int64_t total_q = 0, total_out = 0, total_ref = 0;
    int32_t total_buckets = 0;
    struct Range {
        int K;
        int64_t begin, count;
    };
    std::vector<Range> ranges;  // contiguous ranges of the sorted order, one per k bucket present

    // arena
    void* split = nullptr;
    unsigned char* recs = nullptr;
    int32_t* leaf_off = nullptr;
    int32_t* shuffled = nullptr;
    int32_t* inv_shuffled = nullptr;
    unsigned char* boxes = nullptr;
    CloudDesc* d_clouds = nullptr;
    int32_t* qcloud = nullptr;     // [total_q] cloud of each query row
    int32_t* rec_cloud = nullptr;  // [total_ref] cloud of each record
    // scratch
    uint32_t* key = nullptr;    // [total_q] presort keys
    uint32_t* row = nullptr;    // [total_q] query rows
    uint32_t* key_s = nullptr;  // [total_q] sorted keys
    int32_t* order = nullptr;   // [total_q] query rows in sorted order
    uint32_t* rank = nullptr;   // [total_q] position of each query row in the sorted order
    uint32_t* run_start = nullptr;  // [total_ref] runs of each record in the sorted pairs
    uint32_t* run_end = nullptr;
    void* q_sorted = nullptr;   // [total_q * dims] queries in sorted order
    void* g_sorted = nullptr;   // [total_out] upstream gradient by canonical slot (captured training steps)
    bool g_sorted_valid = false;  // the captured graph's forward writes g_sorted before its backward
    uint32_t* keys = nullptr;   // [total_out] pairs (record, canonical slot)
    uint32_t* vals = nullptr;
    uint32_t* keys_s = nullptr;
    uint32_t* vals_s = nullptr;
    bool persistent = false;    // one work-queue traversal launch for every k bucket
    uint32_t* work = nullptr;   // [8] work counters of the persistent traversal
    int key_bits = 1;           // bits of the presort keys
    int pair_bits = 1;          // bits the dL/dr pair sort orders by
    int pair_tag_shift = 0;     // pair keys: 0 = arena record; else cloud << shift | record within the cloud
    void* radix_ws = nullptr;
    size_t radix_ws_bytes = 0;
    // captured forward + backward
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

// Builds the plan from host trees (their device copies, if any, are not used).  ks[c] in
// [1, min(32, n_points)]; n_queries[c] >= 0.
BatchPlan* batch_create(const Tree* const* trees, int n_clouds, const int64_t* n_queries, const int32_t* ks, int tie,
                        int device);
void batch_free(BatchPlan* p) noexcept;
void batch_forward(BatchPlan& p, const void* queries, void* dist, int32_t* idx, int64_t* orig, hipStream_t s);
// idx: the structured indices the forward wrote; either gradient may be null.  The canonical
// order is recomputed from the queries.
void batch_backward(BatchPlan& p, const void* queries, const int32_t* idx, const void* grad_dist, void* grad_q,
                    void* grad_ref, hipStream_t s);
void batch_capture(BatchPlan& p, const void* queries, void* dist, int32_t* idx, int64_t* orig, const void* grad_dist,
                   void* grad_q, void* grad_ref, hipStream_t s);
void batch_launch(BatchPlan& p, hipStream_t s);

// Per-dtype kernel launchers (batch_f32.hip / batch_f64.hip).
void batch_init_qcloud(const BatchPlan& p, hipStream_t s);
void batch_presort_f32(const BatchPlan& p, const float* q, hipStream_t s);
void batch_presort_f64(const BatchPlan& p, const double* q, hipStream_t s);
// emit: also write the training outputs (the queries at their sorted positions and the
// (record, canonical slot) pairs) for a following batch_backward_*(from_forward = true).
// dq (with g, the upstream gradient of the distances, already in place): also write dL/dq.
void batch_traverse_f32(const BatchPlan& p, const float* q, float* dist, int32_t* idx, int64_t* orig, bool emit,
                        const float* g, float* dq, hipStream_t s);
void batch_traverse_f64(const BatchPlan& p, const double* q, double* dist, int32_t* idx, int64_t* orig, bool emit,
                        const double* g, double* dq, hipStream_t s);
void batch_backward_f32(const BatchPlan& p, const float* q, const int32_t* idx, const float* g, float* dq, float* dr,
                        bool from_forward, hipStream_t s);
void batch_backward_f64(const BatchPlan& p, const double* q, const int32_t* idx, const double* g, double* dq,
                        double* dr, bool from_forward, hipStream_t s);

}  // namespace tk
