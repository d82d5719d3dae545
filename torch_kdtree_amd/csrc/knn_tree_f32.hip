// float32 instantiations of the traversal and presort kernels.
#include <algorithm>

#include "knn_tree_launch.hpp"

namespace tk {

void launch_tree_search_f32(const Tree& t, const float* q, int64_t nq, const int32_t* order, int k, float* dist,
                            int32_t* idx, int64_t* orig, int tie, hipStream_t s) {
    int kb = 1;
    while (kb < k) kb <<= 1;
    // a tuned launch variant, where one is prebuilt (finite trees without record-group boxes: the
    // others run the exact kernel or the group-testing one)
    if (t.variant > 0 && t.finite && !t.dev->group_boxes &&
        launch_tree_variant_f32_d3(t, q, nq, order, k, kb, dist, idx, orig, tie == kTieLowerIndex, t.variant, s))
        return;
    launch_tree_search_t<float>(t, q, nq, order, k, dist, idx, orig, tie, s);
}

void launch_leaf_bucket_f32(const Tree& t, const float* q, int64_t nq, int shift, uint32_t* key, uint32_t* row,
                            hipStream_t s) {
    launch_leaf_bucket_t<float>(t, q, nq, shift, key, row, s);
}

void launch_morton_bucket_f32(const Tree& t, const float* q, int64_t nq, void* bbox, uint32_t* key, uint32_t* row,
                              hipStream_t s) {
    launch_morton_bucket_t<float>(t, q, nq, bbox, key, row, s);
}

int morton_bucket_count(int dims) { return morton_buckets(dims); }

}  // namespace tk
