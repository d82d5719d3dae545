// Launch helpers for the traversal kernels, instantiated once per dtype
// (knn_tree_f32.hip, knn_tree_f64.hip) so the two compile in parallel.
#pragma once

#include "knn_gpu.hpp"
#include "knn_tree.hpp"

namespace tk {

template <typename T>
TreeView<T> tree_view(const Tree& t) {
    TreeView<T> v{static_cast<const T*>(t.dev->split), t.dev->recs, t.dev->leaf_off, t.dev->shuffled,
                  t.dev->inv_shuffled, t.dev->boxes, t.levels, t.n_nodes};
    v.split_bounds = t.finite;
    v.lower_direct = t.lower_list == 2 || (t.lower_list == 0 && t.tie_prone);
    v.group_boxes = t.dev->group_boxes;
    v.group_base = t.dev->group_base;
    return v;
}

// One register-list kernel; a lower-index search on a tree whose cloud ties often
// (tv.lower_direct) takes the 64-bit keyed list's instantiation.
template <typename T, int D, int KK, bool kLower, bool kGroups>
void launch_tree_kernel(const TreeView<T>& tv, dim3 grid, dim3 block, hipStream_t s, const T* q, int64_t nq,
                        const int32_t* order, int k, T* dist, int32_t* idx, int64_t* orig) {
    if constexpr (kLower && TreeList<T, KK, true>::kTie) {
        if (tv.lower_direct) {
            hipLaunchKernelGGL((knn_tree_kernel<T, D, KK, true, kGroups, false>), grid, block, 0, s, tv, q, nq, order, k,
                               dist, idx, orig);
            return;
        }
    }
    hipLaunchKernelGGL((knn_tree_kernel<T, D, KK, kLower, kGroups>), grid, block, 0, s, tv, q, nq, order, k, dist, idx,
                       orig);
}

template <typename T, int D, bool kLower>
void launch_tree_td(const TreeView<T>& tv, const T* q, int64_t nq, const int32_t* order, int k, T* dist, int32_t* idx,
                    int64_t* orig, hipStream_t s) {
    const dim3 grid((unsigned)((nq + kTreeBlock - 1) / kTreeBlock));
    const dim3 block(kTreeBlock);
    int kb = 1;
    while (kb < k) kb <<= 1;
    if (kb > TK_REG_LIST_MAX_K) kb = 0;  // the LDS-list kernel
#define TK_LAUNCH_K(KK)                                                                                         \
    case KK:                                                                                                    \
        if (tv.group_boxes)                                                                                     \
            launch_tree_kernel<T, D, KK, kLower, true>(tv, grid, block, s, q, nq, order, k, dist, idx, orig);   \
        else                                                                                                    \
            launch_tree_kernel<T, D, KK, kLower, false>(tv, grid, block, s, q, nq, order, k, dist, idx, orig);  \
        break;
    switch (kb) {
        TK_LAUNCH_K(1)
        TK_LAUNCH_K(2)
        TK_LAUNCH_K(4)
        TK_LAUNCH_K(8)
        TK_LAUNCH_K(16)
        TK_LAUNCH_K(32)
        default: {
            const size_t shmem = lds_list_bytes(k, sizeof(T), kLower);
            const void* fn = reinterpret_cast<const void*>(&knn_tree_lds_kernel<T, D, kLower>);
            if (shmem > 65536)
                TK_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
            hipLaunchKernelGGL((knn_tree_lds_kernel<T, D, kLower>), dim3((unsigned)((nq + kLdsBlock - 1) / kLdsBlock)),
                               dim3(kLdsBlock), shmem, s, tv, q, nq, order, k, dist, idx, orig);
        }
    }
#undef TK_LAUNCH_K
    TK_HIP_CHECK(hipGetLastError());
}

// The reference's own walk and list (knn_tree_exact_kernel), for trees holding a non-finite value.
template <typename T, int D>
void launch_tree_exact(const TreeView<T>& tv, const T* q, int64_t nq, const int32_t* order, int k, T* dist,
                       int32_t* idx, int64_t* orig, hipStream_t s) {
    const size_t shmem = exact_list_bytes(k, sizeof(T));
    const void* fn = reinterpret_cast<const void*>(&knn_tree_exact_kernel<T, D>);
    if (shmem > 65536) TK_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    hipLaunchKernelGGL((knn_tree_exact_kernel<T, D>), dim3((unsigned)((nq + kLdsBlock - 1) / kLdsBlock)),
                       dim3(kLdsBlock), shmem, s, tv, q, nq, order, k, dist, idx, orig);
    TK_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_tree_search_t(const Tree& t, const T* q, int64_t nq, const int32_t* order, int k, T* dist, int32_t* idx,
                          int64_t* orig, int tie, hipStream_t s) {
    const TreeView<T> tv = tree_view<T>(t);
#define TK_DIMS(DD)                                                            \
    case DD:                                                                   \
        if (tie != kTieLowerIndex && !t.finite)                                \
            launch_tree_exact<T, DD>(tv, q, nq, order, k, dist, idx, orig, s); \
        else if (tie == kTieLowerIndex)                                        \
            launch_tree_td<T, DD, true>(tv, q, nq, order, k, dist, idx, orig, s); \
        else                                                                   \
            launch_tree_td<T, DD, false>(tv, q, nq, order, k, dist, idx, orig, s); \
        break;
    switch (t.dims) {
        TK_DIMS(1)
        TK_DIMS(2)
        TK_DIMS(3)
        TK_DIMS(8)
        default: throw std::runtime_error("unsupported dims");
    }
#undef TK_DIMS
}

#ifndef TK_HILBERT_MAX_D
#define TK_HILBERT_MAX_D 3
#endif
// bbox: 2*dims ordered integers of sizeof(T) bytes each, initialised by the caller (init_bbox).
template <typename T>
void launch_morton_bucket_t(const Tree& t, const T* q, int64_t nq, void* bbox, uint32_t* key, uint32_t* row,
                            hipStream_t s) {
    auto* box = static_cast<OrderedT<T>*>(bbox);
    const unsigned box_blocks = (unsigned)std::min<int64_t>(1024, (nq + kBoxBlock - 1) / kBoxBlock);
    const dim3 grid((unsigned)((nq + kPresortBlock - 1) / kPresortBlock));
#define TK_DIMS(DD)                                                                                              \
    case DD:                                                                                                     \
        hipLaunchKernelGGL((bbox_kernel<T, DD>), dim3(box_blocks), dim3(kBoxBlock), 0, s, q, nq, box);          \
        hipLaunchKernelGGL((morton_bucket_kernel<T, DD>), grid, dim3(kPresortBlock), 0, s, q, nq, box,           \
                           MortonBits<DD>::kBits, key, row, DD <= TK_HILBERT_MAX_D);                                                     \
        break;
    switch (t.dims) {
        TK_DIMS(1)
        TK_DIMS(2)
        TK_DIMS(3)
        TK_DIMS(8)
        default: throw std::runtime_error("unsupported dims");
    }
#undef TK_DIMS
    TK_HIP_CHECK(hipGetLastError());
}

inline int morton_buckets(int dims) {
    switch (dims) {
        case 1: return MortonBits<1>::kBuckets;
        case 2: return MortonBits<2>::kBuckets;
        case 3: return MortonBits<3>::kBuckets;
        case 8: return MortonBits<8>::kBuckets;
    }
    throw std::runtime_error("unsupported dims");
}

template <typename T>
void launch_leaf_bucket_t(const Tree& t, const T* q, int64_t nq, int shift, uint32_t* key, uint32_t* row,
                          hipStream_t s) {
    const TreeView<T> tv = tree_view<T>(t);
    const dim3 grid((unsigned)((nq + kPresortBlock - 1) / kPresortBlock));
#define TK_DIMS(DD)                                                                                              \
    case DD:                                                                                                     \
        hipLaunchKernelGGL((leaf_bucket_kernel<T, DD>), grid, dim3(kPresortBlock), 0, s, tv, q, nq, shift, key,  \
                           row);                                                                                 \
        break;
    switch (t.dims) {
        TK_DIMS(1)
        TK_DIMS(2)
        TK_DIMS(3)
        TK_DIMS(8)
        default: throw std::runtime_error("unsupported dims");
    }
#undef TK_DIMS
    TK_HIP_CHECK(hipGetLastError());
}

}  // namespace tk
