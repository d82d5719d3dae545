// Host-side KD-tree object shared by the CPU build, the CPU search and the GPU upload.
//
// Layout (identical to the reference's flat layout, src/kdtree.hpp:126-206):
//   * nodes in heap order: node i has children 2i+1 and 2i+2; a tree of `levels` levels has
//     2^levels - 1 nodes and 2^levels leaves (src/kdtree.hpp:73-93);
//   * the split axis of a node at depth l is l mod dims (root axis 0, src/kdtree.hpp:473-477),
//     its split value is the median of src/kdtree.hpp:392-401;
//   * leaf j of leaf-parent p is 2p+1-(2^levels-1) (left) and +1 (right), src/kdtree.hpp:300-303;
//   * leaves are contiguous runs of the structured point array; leaf_off[j] .. leaf_off[j+1].
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace tk {

enum DType : int { kF32 = 0, kF64 = 1 };

inline size_t dtype_size(int dtype) { return dtype == kF64 ? 8 : 4; }

inline int32_t total_nodes(int levels) { return (int32_t)((1u << levels) - 1u); }

struct DeviceTree;  // defined by the GPU translation unit

struct Tree {
    int dtype = kF32;
    int dims = 0;
    int levels = 0;
    int32_t n_points = 0;
    int32_t n_nodes = 0;
    int32_t n_leaves = 0;
    std::vector<uint8_t> axis;          // [n_nodes]
    std::vector<unsigned char> split;   // [n_nodes] values of type T
    std::vector<int32_t> leaf_off;      // [n_leaves + 1]
    std::vector<unsigned char> points;  // [n_points * dims] values of type T, structured order
    std::vector<int32_t> shuffled;      // [n_points] original index of each structured position
    // Bounding box of the points below each node, heap order over internal nodes and leaves
    // (leaf j is heap node n_nodes + j): [(n_nodes + n_leaves) * 2 * dims] values of type T,
    // per node lo[dims] then hi[dims]; an empty leaf has lo = +inf, hi = -inf.
    std::vector<unsigned char> aabb;
    // Every coordinate and every split value is finite.  Otherwise a distance or a pruning bound
    // can be NaN, and the reference's insertion (src/nndistance.hpp:41-85) takes a NaN distance
    // into the list; the searches then follow the reference step by step, without point boxes.
    bool finite = true;
    // Launch variant of the GPU traversal (kTreeVariants in knn_gpu.hpp; 0 = the built-in choice).
    int variant = 0;
    // List of the GPU's lower-index float searches with 8 to 32 entries: 0 = chosen from the cloud
    // (tie_prone), 1 = the reference policy's list with tie flags, each unsure lane searched again,
    // 2 = the list keyed by (distance, original index) from the start.  Results never depend on it.
    int lower_list = 0;
    // Set at upload: exact distance ties look common among this cloud's neighbours (duplicated
    // points, coordinates on few distinct values), or the tree holds a non-finite value.
    bool tie_prone = false;
    DeviceTree* dev = nullptr;          // GPU copy, owned

    template <typename T> T* split_ptr() { return reinterpret_cast<T*>(split.data()); }
    template <typename T> const T* split_ptr() const { return reinterpret_cast<const T*>(split.data()); }
    template <typename T> T* points_ptr() { return reinterpret_cast<T*>(points.data()); }
    template <typename T> const T* points_ptr() const { return reinterpret_cast<const T*>(points.data()); }
    template <typename T> const T* aabb_ptr() const { return reinterpret_cast<const T*>(aabb.data()); }
};

// Largest tree depth accepted by the build (the reference caps `levels` at 13 in Python,
// nn_distance.py:155, because of its GPU kernel's LDS tag array; this build has no tag array).
constexpr int kMaxLevels = 24;

// Dimensions this library is instantiated for (the reference: 1, 2, 3; README.md:83-95
// describes adding more; 8 is added here).
inline bool dims_supported(int d) { return d == 1 || d == 2 || d == 3 || d == 8; }

// Builds the tree on the host.  Throws std::runtime_error with the reference's messages.
Tree* build_tree(const void* points, int64_t n_points, int dims, int levels, int dtype, int n_threads);

// Builds trees on the GPU (tree_build_gpu.hip): tree c from points[base_c .. base_c + n_points[c])
// of the concatenated device array `points` ([sum n_points, dims] values of dtype, on `device`),
// with levels[c] levels; the work runs on `stream` (a hipStream_t).  Returns host trees (not
// uploaded).  Equal to build_tree whenever no node has two points with equal coordinates on its
// split axis; otherwise equal up to the order of such points.
std::vector<Tree*> build_trees_gpu(const void* points, const int64_t* n_points, const int32_t* levels, int n_trees,
                                   int dims, int dtype, int device, void* stream);

// Derived arrays of a tree whose split values, leaf offsets, points and indices are set: the
// split axes and the point boxes.
void finish_tree(Tree& t);

// Serialization (tree_io.cpp): a versioned, checksummed byte image of the tree's arrays.
std::vector<unsigned char> serialize_tree(const Tree& t);
Tree* deserialize_tree(const unsigned char* buf, size_t n);
void save_tree(const Tree& t, const char* path);
Tree* load_tree(const char* path);

// CPU KNN search with the reference's semantics (iterative near-first DFS, upper-bound
// insertion; src/kdtree.cpp:16-194).
// stats (may be null): stats[0] = leaves scanned, stats[1] = points scanned, over all queries.
void query_cpu(const Tree& t, const void* queries, int64_t n_queries, int k, void* dist, int32_t* idx,
               int64_t* orig_idx, int n_threads, int64_t* stats = nullptr);

}  // namespace tk
