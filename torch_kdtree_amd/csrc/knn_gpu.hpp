// Host-side interface of the GPU subsystem (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tree.hpp"

namespace tk {

// Device copy of a tree (the reference's copyPartitionToGPU, src/kdtree_g.cu:63-98, with a
// different layout: split values in heap order, one record per point, leaf offsets).
struct DeviceTree {
    int device = -1;
    void* split = nullptr;            // T[n_nodes]
    unsigned char* recs = nullptr;    // records, RecLayout<T, dims>::kBytes each
    int32_t* leaf_off = nullptr;      // [n_leaves + 1]
    int32_t* shuffled = nullptr;      // [n_points] original index of each structured position
    int32_t* inv_shuffled = nullptr;  // [n_points] structured position of each original index
    unsigned char* boxes = nullptr;   // [n_nodes + n_leaves] point boxes, BoxLayout<T, dims>::kBytes each
    // point boxes of the record groups: each leaf's records in runs of kGroupRecs (knn_tree.hpp)
    // from its first (the last run may be shorter); group_base[leaf] = index of the leaf's first
    // group box.  Null for trees holding a non-finite value (their searches use no boxes) and for
    // trees whose leaves hold fewer than kGroupMinLeaf points on average.
    unsigned char* group_boxes = nullptr;
    int32_t* group_base = nullptr;    // [n_leaves + 1]
    int rec_bytes = 0;
    int box_bytes = 0;
};

int rec_bytes_for(int dtype, int dims);
int box_bytes_for(int dtype, int dims);
void upload_tree(Tree& t, int device);
void free_device_tree(Tree& t) noexcept;

enum TiePolicy : int { kTieReference = 0, kTieLowerIndex = 1 };
// Query presort: auto = Morton buckets when there are enough queries to matter.
enum PresortMode : int { kPresortAuto = -1, kPresortOff = 0, kPresortMorton = 1, kPresortLeaf = 2 };

// Largest k served on the GPU (the reference's limit, src/kdtree_g.cu:283,421-422).
constexpr int kMaxGpuK = 128;

struct QueryOpts {
    int tie = kTieReference;
    int presort = kPresortAuto;
};

// Workspace (bytes) the query needs for the presort; 0 when no presort runs.
size_t query_gpu_workspace(const Tree& t, int64_t nq, const QueryOpts& o);

// dist [nq*k] T, idx [nq*k] int32 structured indices, orig [nq*k] int64 original indices
// (may be null).  `workspace` may be null: then it is allocated stream-ordered.
void query_gpu(const Tree& t, const void* queries, int64_t nq, int k, void* dist, int32_t* idx, int64_t* orig,
               const QueryOpts& o, hipStream_t stream, void* workspace, size_t workspace_bytes);

// query_gpu with the presort workspace taken from a static device array instead of an allocation
// (knn_gpu.hip): at most kNoAllocMaxQueries queries per call; calls that overlap on one GPU share
// the array and race.
constexpr int64_t kNoAllocMaxQueries = int64_t(1) << 21;
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
void query_gpu_noalloc(const Tree& t, const void* queries, int64_t nq, int k, void* dist, int32_t* idx, int64_t* orig,
                       const QueryOpts& o, hipStream_t stream);

// Prebuilt launch variants of the traversal kernel that an autotuner picks among: leaf-scan chunk
// (records staged in LDS per pass), occupancy target (waves per SIMD) and threads per block.
// Variant 0 is the built-in choice (StageBytes / TreeOccupancy in knn_tree.hpp, blocks of 128).
// The others exist for float searches at d = 3 with lists of 8 or 16 entries (k = 5..16); any
// other search runs variant 0 whatever is asked.  Results never depend on the variant.
struct TreeVariant {
    int chunk, waves, block;
};
constexpr TreeVariant kTreeVariants[] = {{0, 0, 0},    {4, 7, 128},  {8, 7, 128},  {16, 7, 128}, {8, 5, 128},
                                         {16, 5, 128}, {16, 7, 64},  {16, 7, 256}, {8, 7, 256}};
constexpr int kNumTreeVariants = (int)(sizeof(kTreeVariants) / sizeof(kTreeVariants[0]));
// Launches variant `variant` (> 0) for a float d = 3 search when it is prebuilt for list width kb
// (knn_tree_variants.hip); returns false when it is not, and the caller launches variant 0.
bool launch_tree_variant_f32_d3(const Tree& t, const float* q, int64_t nq, const int32_t* order, int k, int kb,
                                float* dist, int32_t* idx, int64_t* orig, bool lower, int variant, hipStream_t s);

// Per-dtype kernel launchers (knn_tree_f32.hip / knn_tree_f64.hip).
void launch_tree_search_f32(const Tree& t, const float* q, int64_t nq, const int32_t* order, int k, float* dist,
                            int32_t* idx, int64_t* orig, int tie, hipStream_t s);
void launch_tree_search_f64(const Tree& t, const double* q, int64_t nq, const int32_t* order, int k, double* dist,
                            int32_t* idx, int64_t* orig, int tie, hipStream_t s);
void launch_leaf_bucket_f32(const Tree& t, const float* q, int64_t nq, int shift, uint32_t* key, uint32_t* row,
                            hipStream_t s);
void launch_leaf_bucket_f64(const Tree& t, const double* q, int64_t nq, int shift, uint32_t* key, uint32_t* row,
                            hipStream_t s);
void launch_morton_bucket_f32(const Tree& t, const float* q, int64_t nq, void* bbox, uint32_t* key, uint32_t* row,
                              hipStream_t s);
void launch_morton_bucket_f64(const Tree& t, const double* q, int64_t nq, void* bbox, uint32_t* key, uint32_t* row,
                              hipStream_t s);
int morton_bucket_count(int dims);
// Sets a 2*dims ordered-integer bounding box to (all ones | zero), the start of a min/max reduction.
void init_bbox(void* bbox, int dtype, int dims, hipStream_t stream);
// Bits per axis of the Morton grid of the backward's canonical query order: about 8 queries per
// cell, at most the grid of the forward presort.
int canonical_morton_bits(int64_t nq, int dims);

// Streaming device copy (16-byte loads and stores) of `bytes` (a multiple of 16, aligned pointers).
void stream_copy_gpu(void* dst, const void* src, size_t bytes, hipStream_t stream);
// VALU issue probe over `threads` lanes (a multiple of 256), writing one float per lane to `out`;
// returns the wave instructions of its loop (16 per iteration per wavefront, 24 in mode 2).
int64_t valu_probe_gpu(float* out, int64_t threads, int iters, int mode, hipStream_t stream);

// Brute force (bruteforce.hip).
void bruteforce_gpu(const void* ref, int32_t n, const void* queries, int64_t nq, int dims, int dtype, int k,
                    void* dist, int32_t* idx, hipStream_t stream, void* workspace, size_t workspace_bytes);
size_t bruteforce_gpu_workspace(int32_t n, int64_t nq, int dims, int dtype, int k);

// Backward (backward.hip).
void knn_backward_gpu(const void* queries, const void* ref, int64_t nq, int32_t n, int dims, int dtype, int k,
                      const int64_t* idx, const void* grad_dist, void* grad_q, void* grad_ref, hipStream_t stream,
                      void* workspace, size_t workspace_bytes);
size_t knn_backward_workspace(int64_t nq, int32_t n, int k, int dims, int dtype);
// The same backward with the neighbours given by structured position (sidx[nq*k], int32) and the
// tree's structured-to-original map shuffled[n] (device memory): the same results, with the
// reference points gathered in structured order (spatially coherent reads).  Always takes a
// workspace (the reordered points); workspace may be null (then allocated stream-ordered).
void knn_backward_structured_gpu(const void* queries, const void* ref, int64_t nq, int32_t n, int dims, int dtype,
                                 int k, const int32_t* sidx, const int32_t* shuffled, const void* grad_dist,
                                 void* grad_q, void* grad_ref, hipStream_t stream, void* workspace,
                                 size_t workspace_bytes);
size_t knn_backward_structured_workspace(int64_t nq, int32_t n, int k, int dims, int dtype);

// Exclusive prefix sum of n int32 values (n <= 2^24) into out; sums needs scan_sums_count(n)
// int32 of scratch.  Kernels only (no memset, no host synchronisation).
int32_t scan_sums_count(int32_t n);
void exclusive_scan_i32(const int32_t* in, int32_t n, int32_t* sums, int32_t* out, hipStream_t stream);

// Stable LSD radix sort of (key, value) pairs of uint32 by the low key_bits bits of the keys
// (radix_sort.hip).  Reads keys_in / vals_in, writes keys_out / vals_out (distinct from the
// inputs); workspace of radix_sort_workspace(n) bytes.  Kernels only, so it can be captured in
// a graph.  n <= 2^28.
size_t radix_sort_workspace(int64_t n);
void radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      int64_t n, int key_bits, void* workspace, size_t workspace_bytes, hipStream_t stream);

// RAII stream-ordered scratch allocation used when the caller passes no workspace.
struct StreamScratch {
    void* ptr = nullptr;
    hipStream_t stream = nullptr;
    StreamScratch(size_t bytes, hipStream_t s);
    ~StreamScratch();
    StreamScratch(const StreamScratch&) = delete;
    StreamScratch& operator=(const StreamScratch&) = delete;
};

// Optional HIP-event timing of the kernels a query launches (off by default).  When on, each
// stage ("presort", "traverse", "bruteforce", "backward") is bracketed by two events on the
// launch stream; prof_read synchronises those events and returns the summed duration.
void prof_enable(bool on);
void prof_reset();
bool prof_read(const char* stage, double* total_ms, int64_t* launches);
struct ProfScope {
    const char* stage = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr;
    ProfScope(const char* stage, hipStream_t s);
    ~ProfScope();
};

// Makes the tree's device current for the lifetime of the guard.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int device);
    ~DeviceGuard();
};

}  // namespace tk
