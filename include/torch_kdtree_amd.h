/*
 * torch_kdtree_amd — C ABI of the MI355X (gfx950) KD-tree KNN operator.
 *
 * Drop-in boundary for the hot path of thomgrand/torch_kdtree:
 *     build_kd_tree(points).query(q, nr_nns_searches=k)
 * The reference binds this path to Python through pybind11 (module `torch_knn`,
 * src/interface.cpp); each entry point below names the reference interface it replaces.
 * This library's own pybind11 module `torch_knn` is a thin layer over these functions.
 *
 * Conventions
 *   - Plain pointers and sizes only.  Point arrays are row-major [n, dims] of float
 *     (TK_F32) or double (TK_F64); KNN outputs are row-major [n_queries, k].
 *   - "structured" indices are positions in the tree's reordered point array (what the
 *     reference's C++ layer writes, src/kdtree.cpp:191-192); "original" indices are rows of
 *     the input point array (what the reference's Python layer returns,
 *     torch_kdtree/nn_distance.py:109).
 *   - GPU entry points take device pointers and a hipStream_t passed as void* (NULL = the
 *     legacy default stream, the stream the reference launches on).  They are asynchronous
 *     with respect to the host, like the reference's launches (src/kdtree_g.cu:425,444).
 *   - Every function returns 0 on success and nonzero on failure; tk_last_error() then
 *     returns the message (thread-local).  No function calls exit() (the reference does,
 *     src/cutils.cuh:34-43).
 *   - `workspace` arguments may be NULL: the library then takes stream-ordered scratch
 *     memory (hipMallocAsync/hipFreeAsync) on the given stream.  The *_workspace functions
 *     return the byte count a caller-provided workspace must have.
 */
#ifndef TORCH_KDTREE_AMD_H
#define TORCH_KDTREE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tk_tree tk_tree; /* opaque: host tree (+ its GPU copy once uploaded) */

enum { TK_F32 = 0, TK_F64 = 1 };

/* Exact-distance tie policies of the GPU query.
 * TK_TIE_REFERENCE: the reference CPU search's order (ties keep the near-first depth-first
 *   scan order, src/kdtree.cpp:16-110 with src/nndistance.hpp:41-85); output is identical to
 *   KDTreeKNNSearch on the same tree.
 * TK_TIE_LOWER_INDEX: the k smallest (distance, original index) pairs. */
enum { TK_TIE_REFERENCE = 0, TK_TIE_LOWER_INDEX = 1 };

/* Query presort before the traversal, so that the lanes of a wavefront get nearby queries:
 * TK_PRESORT_MORTON orders queries along a space-filling curve over a grid on the queries'
 * bounding box (the Hilbert curve for d <= 3, the Morton code for d = 8; the name is kept),
 * TK_PRESORT_LEAF by the tree leaf they fall into; AUTO = MORTON (LEAF at d = 8) from 4096 queries up.  The
 * presort changes the order of work only, never the results. */
enum { TK_PRESORT_AUTO = -1, TK_PRESORT_OFF = 0, TK_PRESORT_MORTON = 1, TK_PRESORT_LEAF = 2 };

/* Thread-local message of the last failure on this thread. */
const char* tk_last_error(void);

/* Library version string, "1.0" (torch_kdtree/__init__.py:4, CMakeLists.txt:15-16,151). */
const char* tk_version(void);

/* Replaces check_for_gpu() (src/interface.cpp:105-112): 1 when built with GPU support. */
int tk_gpu_available(void);

/* Number of visible GPUs (0 when none). */
int tk_device_count(int* count);

/* Replaces the KDTree<T,dims,use_gpu> constructor's host part (src/interface.cpp:25-36)
 * and createKDTree (src/kdtree.hpp:494-534).  Builds the tree on the host with the
 * reference's layout.  dims in {1,2,3,8}; 1 <= levels <= 24; n_threads <= 0 = all cores.
 * Fails with the reference's "Ran out of points" message when n_points < 2^(levels-1). */
int tk_tree_build(const void* points, int64_t n_points, int dims, int dtype, int levels, int n_threads,
                  tk_tree** out);

/* Builds n_trees trees in parallel (one tree per thread, largest first), each as tk_tree_build
 * would: out[i] from points[i][n_points[i]*dims] with levels[i].  On failure no tree is left
 * allocated.  (No counterpart in the reference, whose trees are built one by one.) */
int tk_tree_build_many(const void* const* points, const int64_t* n_points, int n_trees, int dims, int dtype,
                       const int32_t* levels, int n_threads, tk_tree** out);

/* GPU tree build (no counterpart in the reference, whose createKDTree, src/kdtree.hpp:494-534,
 * runs on the host only).  points: device array [n_points*dims] on `device` (-1 = current);
 * the kernels run on `stream` (a hipStream_t; NULL = default stream).  Same node sizes, axes,
 * split values and leaf ranges as tk_tree_build whenever no node has two points with equal
 * coordinates on its split axis; with such ties the tied points may be ordered (and split)
 * differently, because the reference orders them by an unstable std::sort.  Fails when a
 * coordinate it sorts on is NaN.  The tree is uploaded to `device`.  At most 2^28 points. */
int tk_tree_build_gpu(const void* points, int64_t n_points, int dims, int dtype, int levels, int device, void* stream,
                      tk_tree** out);

/* Builds n_trees trees in one GPU build: tree i from the n_points[i] rows that follow the rows of
 * trees 0..i-1 in the device array `points`, with levels[i] levels.  The trees are returned on
 * the host (not uploaded), as tk_tree_build_many returns them.  At most 2^28 points in all. */
int tk_tree_build_many_gpu(const void* points, const int64_t* n_points, int n_trees, int dims, int dtype,
                           const int32_t* levels, int device, void* stream, tk_tree** out);

/* Tree serialization (no counterpart in the reference, whose trees exist only in memory; it
 * extends the read-only getters of src/interface.cpp:45-58 so that a static cloud is built
 * once).  The image is versioned and checksummed: split values, leaf offsets, structured points
 * and shuffled indices.  tk_tree_serialize with buf == NULL stores the image size in *bytes;
 * otherwise *bytes is the buffer size on entry and the image size on return.  Loading checks
 * the checksum and the structure before returning a host tree (upload it with tk_tree_upload). */
int tk_tree_serialize(const tk_tree* tree, void* buf, size_t* bytes);
int tk_tree_deserialize(const void* buf, size_t bytes, tk_tree** out);
int tk_tree_save(const tk_tree* tree, const char* path);
int tk_tree_load(const char* path, tk_tree** out);

/* Replaces copyPartitionToGPU (src/kdtree_g.cu:63-98), called from the constructor at
 * src/interface.cpp:38-41.  Copies the tree to GPU `device` (-1 = current device). */
int tk_tree_upload(tk_tree* tree, int device);

/* Replaces ~KDTree (src/interface.cpp:88-94) and freePartitionFromGPU
 * (src/kdtree_g.cu:140-157).  Frees host and device copies; NULL is ignored. */
void tk_tree_free(tk_tree* tree);

/* Tree facts: out[0]=dims, out[1]=dtype, out[2]=levels, out[3]=n_points, out[4]=device
 * (-1 when not uploaded), out[5]=1 when every coordinate and split value is finite.  A tree with a
 * non-finite value can produce NaN distances, which the reference's insertion takes into the list
 * (src/nndistance.hpp:41-85); its reference-policy queries then follow the reference's walk step
 * for step, without point-box pruning. */
int tk_tree_info(const tk_tree* tree, int32_t out[6]);

/* Launch variants of the GPU traversal (no counterpart in the reference, whose launch shape is
 * fixed at compile time, src/kdtree_g.cu:6,281-283): prebuilt instantiations that differ in the
 * leaf-scan chunk staged in LDS, the occupancy target and the threads per block, for an autotuner
 * to measure and pick among.  Variant 0 is the built-in choice; variants 1..n-1 exist for float
 * searches at d = 3 with k in 5..16 on trees whose values are finite, and any other search runs
 * variant 0.  Results never depend on the variant.
 * tk_tree_variant_info: out = {chunk, waves per SIMD (0 = no target), threads per block} of a
 * variant (zeros for variant 0), *count = number of variants; fails for an unknown variant.
 * tk_tree_set_variant: the variant the tree's later GPU queries launch. */
int tk_tree_variant_info(int variant, int32_t out[3], int* count);
int tk_tree_set_variant(tk_tree* tree, int variant);
int tk_tree_get_variant(const tk_tree* tree, int* variant);

/* The list of the tree's lower-index float GPU searches with k in 5..32 (no counterpart in the
 * reference, whose GPU kernel keeps one LDS list, src/kdtree_g.cu:285-407): mode 0 = chosen at
 * upload from a sample of the cloud, 1 = the reference policy's register list with tie flags
 * (each query that meets an exact distance tie is searched again with the (distance, original
 * index) list), 2 = the (distance, original index) list from the start.  Mode 0 takes 2 for a
 * cloud whose sample shows duplicated points or coordinates on few distinct values, and for a tree
 * with a non-finite value; 1 otherwise.  Results never depend on the mode.
 * tk_tree_get_lower_list: out[0] = mode, out[1] = 1 when mode 0 takes list 2 for this tree. */
int tk_tree_set_lower_list(tk_tree* tree, int mode);
int tk_tree_get_lower_list(const tk_tree* tree, int32_t out[2]);

/* Replaces KDTree::get_shuffled_inds (src/interface.cpp:45-50): out[n_points] receives the
 * original index of each structured position. */
int tk_tree_get_shuffled_inds(const tk_tree* tree, int32_t* out);

/* Replaces KDTree::get_structured_points (src/interface.cpp:52-58): out[n_points*dims]. */
int tk_tree_get_structured_points(const tk_tree* tree, void* out);

/* Node and leaf arrays of the tree (the reference's PartitionInfo, src/kdtree.hpp:165-206):
 * axis[2^levels-1], split[2^levels-1] (dtype), leaf_off[2^levels+1].  Any may be NULL. */
int tk_tree_get_nodes(const tk_tree* tree, int32_t* axis, void* split, int32_t* leaf_off);

/* Replaces KDTree::query / query_recast for CPU trees (src/interface.cpp:60-86) ->
 * KDTreeKNNSearch (src/kdtree.cpp:164-194).  Host pointers.  dist[nq*k], idx[nq*k]
 * structured; orig_idx[nq*k] (may be NULL) original indices. */
int tk_knn_query_cpu(const tk_tree* tree, const void* queries, int64_t n_queries, int k, void* dist, int32_t* idx,
                     int64_t* orig_idx, int n_threads);

/* Search statistics (no counterpart in the reference): runs the host search of the n_queries host
 * queries on the tree's host arrays (CPU or GPU tree) and returns out[0] = leaves scanned and
 * out[1] = points scanned, summed over the queries.  The host search makes the same pruning
 * decisions as the GPU traversal under the reference tie policy, so these are the GPU's work too. */
int tk_knn_search_stats(const tk_tree* tree, const void* queries, int64_t n_queries, int k, int64_t out[2],
                        int n_threads);

/* Replaces KDTree::query / query_recast for GPU trees (src/interface.cpp:60-86) ->
 * KDTreeKNNGPUSearch (src/kdtree_g.cu:409-457).  Device pointers.  Writes squared
 * distances and structured int32 indices like the reference; orig_idx (may be NULL) also
 * receives the original int64 indices, which the reference computes afterwards in torch
 * (nn_distance.py:109).  1 <= k <= min(n_points, 128) (src/kdtree_g.cu:421-422). */
int tk_knn_query_gpu(const tk_tree* tree, const void* queries, int64_t n_queries, int k, void* dist, int32_t* idx,
                     int64_t* orig_idx, int tie, int presort, void* workspace, size_t workspace_bytes, void* stream);
int tk_knn_query_gpu_workspace(const tk_tree* tree, int64_t n_queries, int k, int tie, int presort, size_t* bytes);

/* No counterpart in the reference.  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  Same arguments,
 * checks and results as tk_knn_query_gpu with workspace == NULL, but the presort workspace is a
 * static device array of 64 MiB per GPU (allocated with the library's code object and held until
 * the process ends) instead of an allocation: no hipMalloc, hipFree or allocating library call.
 * At most TK_NOALLOC_MAX_QUERIES queries per call; more fail with an error.  Every call on one GPU
 * shares that GPU's array, so calls that overlap in time on one GPU (different streams without
 * ordering, or several host threads) race: serialise them. */
#define TK_NOALLOC_MAX_QUERIES 2097152
int tk_knn_query_gpu_noalloc(const tk_tree* tree, const void* queries, int64_t n_queries, int k, void* dist,
                             int32_t* idx, int64_t* orig_idx, int tie, int presort, void* stream);

/* Replaces NmDistanceKernelLauncher[Dynamic] (src/nndistance_g.cu:142-220), which the
 * reference compiles but never binds.  Exhaustive KNN over ref[n_ref*dims] in original
 * order; ties go to the lower index; a NaN distance is never inserted.  idx receives original
 * int32 indices.  1 <= k <= 128 (lists beyond 32 entries are held in LDS). */
int tk_knn_bruteforce_gpu(const void* ref, int32_t n_ref, const void* queries, int64_t n_queries, int dims, int dtype,
                          int k, void* dist, int32_t* idx, void* workspace, size_t workspace_bytes, void* stream);
int tk_knn_bruteforce_gpu_workspace(int32_t n_ref, int64_t n_queries, int dims, int dtype, int k, size_t* bytes);

/* Replaces the autograd recomputation of torch_kdtree/nn_distance.py:111-112.  Given the
 * original-order queries[nq*dims], ref[n_ref*dims], neighbour indices idx[nq*k] (original,
 * int64) and the upstream gradient grad_dist[nq*k] of the squared distances, writes
 * grad_queries[nq*dims] and grad_ref[n_ref*dims] (either may be NULL).  grad_ref is
 * deterministic: no atomics, a fixed summation order (each reference point's terms subtracted in
 * a canonical query order).  grad_queries alone has no size limit and needs no workspace;
 * grad_ref takes at most 2^28 queries, any k: its (reference, slot) pairs are sorted in slices of
 * at most 2^28 pairs, consecutive in the canonical order, each slice continuing the sums of the
 * previous ones, which gives the same bits as one pass. */
int tk_knn_backward_gpu(const void* queries, const void* ref, int64_t n_queries, int32_t n_ref, int dims, int dtype,
                        int k, const int64_t* idx, const void* grad_dist, void* grad_queries, void* grad_ref,
                        void* workspace, size_t workspace_bytes, void* stream);
int tk_knn_backward_gpu_workspace(int64_t n_queries, int32_t n_ref, int dims, int dtype, int k, size_t* bytes);

/* The same backward for the neighbours of a GPU query on `tree` given by their structured indices
 * sidx[nq*k] (int32, the query's idx output) instead of original ones; ref[n_ref*dims] holds the
 * current coordinates of the tree's points in original order (n_ref = the tree's point count,
 * dims and dtype the tree's).  Same outputs, bit for bit, as tk_knn_backward_gpu with the matching
 * original indices: the points are first gathered into the tree's structured order, so the
 * gathers of spatially close queries read close rows.  Replaces the same recomputation
 * (nn_distance.py:111-112); always takes a workspace (NULL: allocated stream-ordered). */
int tk_knn_backward_tree_gpu(const tk_tree* tree, const void* queries, const void* ref, int64_t n_queries, int k,
                             const int32_t* sidx, const void* grad_dist, void* grad_queries, void* grad_ref,
                             void* workspace, size_t workspace_bytes, void* stream);
int tk_knn_backward_tree_gpu_workspace(const tk_tree* tree, int64_t n_queries, int k, size_t* bytes);

/* Batched many-cloud mode (no counterpart in the reference, which queries one tree per call;
 * it replaces a loop of KDTree::query calls, src/interface.cpp:60-86, over many clouds).
 * A plan packs n_trees trees (same dims and dtype; built, not necessarily uploaded) with their
 * query counts n_queries[c] and k values ks[c] (1..32) into one device arena.  The trees may be
 * freed afterwards.
 *   Queries: one row-major [sum n_queries, dims] array; cloud c owns rows
 *     query_off[c] .. query_off[c+1] (tk_batch_offsets).
 *   Outputs: dist (dtype), idx (int32, structured indices of the cloud's tree) and orig_idx
 *     (int64, original indices within the cloud's point array; may be NULL) are flat arrays of
 *     sum(n_queries[c] * ks[c]) elements; cloud c owns out_off[c] .. out_off[c+1], row-major
 *     [n_queries[c], ks[c]].  Results equal those of tk_knn_query_gpu on each tree.
 *   Backward: grad_dist has the layout of dist; grad_queries the layout of the queries;
 *     grad_ref is [sum n_points, dims], cloud c owning rows ref_off[c] .. ref_off[c+1] in its
 *     original point order.  Either gradient may be NULL.  Deterministic, and equal to
 *     tk_knn_backward_gpu on each cloud.
 * All scratch memory is allocated with the plan; a plan serves one stream at a time.  At most
 * 2^27 queries and 2^28 output elements per plan.
 * tk_batch_capture records forward (+ backward when grad_dist and a gradient are given) with
 * the given buffers into a HIP graph; tk_batch_launch replays it on a stream. */
typedef struct tk_batch tk_batch;
int tk_batch_create(const tk_tree* const* trees, int n_trees, const int64_t* n_queries, const int32_t* ks, int tie,
                    int device, tk_batch** out);
void tk_batch_free(tk_batch* batch);
/* out[0] = total queries, out[1] = total output elements, out[2] = total reference points,
 * out[3] = number of clouds. */
int tk_batch_sizes(const tk_batch* batch, int64_t out[4]);
/* Each array (any may be NULL) receives n_clouds + 1 prefix offsets. */
int tk_batch_offsets(const tk_batch* batch, int64_t* query_off, int64_t* out_off, int64_t* ref_off);
int tk_batch_forward(tk_batch* batch, const void* queries, void* dist, int32_t* idx, int64_t* orig_idx, void* stream);
int tk_batch_backward(tk_batch* batch, const void* queries, const int32_t* idx, const void* grad_dist,
                      void* grad_queries, void* grad_ref, void* stream);
int tk_batch_capture(tk_batch* batch, const void* queries, void* dist, int32_t* idx, int64_t* orig_idx,
                     const void* grad_dist, void* grad_queries, void* grad_ref, void* stream);
int tk_batch_launch(tk_batch* batch, void* stream);
/* Traversal launch shape of the plan's later forwards and captures: 0 (default) one launch per k
 * bucket; 1 one persistent launch for every bucket, whose wavefronts take units of 64 presorted
 * queries from per-XCD work queues until all are done.  Results are identical either way. */
int tk_batch_set_persistent(tk_batch* batch, int on);

/* Stable radix sort of n (key, value) uint32 pairs on the GPU by the low key_bits bits of the
 * keys (no counterpart in the reference: it is the sort behind the deterministic grad_ref of
 * tk_knn_backward_gpu, exported for callers that need the same order).  Inputs and outputs
 * are distinct device arrays; n <= 2^28.  Launches kernels only (no memset, no host wait), so
 * it may be captured in a HIP graph. */
int tk_sort_pairs_gpu(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      int64_t n, int key_bits, void* workspace, size_t workspace_bytes, void* stream);
int tk_sort_pairs_gpu_workspace(int64_t n, size_t* bytes);

/* Kernel timing with HIP events (no counterpart in the reference, whose only timing is
 * commented out, src/kdtree_g.cu:435-456).  When enabled, every GPU entry point brackets its
 * kernels with events on its stream, per stage: "presort", "traverse", "bruteforce",
 * "backward".  tk_profile_read waits for the recorded events and returns the summed
 * milliseconds and the number of bracketed calls of that stage. */
int tk_profile_enable(int on);
int tk_profile_reset(void);
int tk_profile_read(const char* stage, double* total_ms, int64_t* calls);

/* A streaming device-to-device copy (16-byte loads and stores; aligned pointers, bytes a multiple
 * of 16), bracketed as stage "copy" when timing is on.  No counterpart in the reference: its rate
 * is the achievable HBM bandwidth that the traversal's roofline is quoted against. */
int tk_copy_gpu(void* dst, const void* src, size_t bytes, void* stream);

/* A VALU issue probe (no counterpart in the reference): `threads` lanes (a positive multiple of
 * 256) each run `iters` iterations of 8 independent chains of 32-bit VALU instructions (mode 0: a
 * multiply and an add, mode 1: two med3, mode 2: a compare, a select and an add per chain) and write
 * one float to out[lane].  *wave_insts receives the VALU wave
 * instructions of the loop.  Bracketed as stage "valu_probe" when timing is on: instructions per
 * second over all CUs is the issue ceiling the traversal's instruction count is quoted against. */
int tk_valu_probe_gpu(float* out, int64_t threads, int iters, int mode, void* stream, int64_t* wave_insts);

#ifdef __cplusplus
}
#endif

#endif /* TORCH_KDTREE_AMD_H */
