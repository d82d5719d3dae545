// Prebuilt launch variants of the float d = 3 traversal kernel for lists of 8 and 16 entries (the
// table kTreeVariants, knn_gpu.hpp): the choices an autotuner measures and picks among.  Separate
// from knn_tree_f32.hip so the instantiations compile in parallel.
#include "knn_tree_launch.hpp"

namespace tk {
namespace {

template <int K, bool kLower, int CH, int BLK, int WAVES>
void launch_v(const TreeView<float>& tv, const float* q, int64_t nq, const int32_t* order, int k, float* dist,
              int32_t* idx, int64_t* orig, hipStream_t s) {
    constexpr int LD = CH < TK_SCAN_LOADS_F32 ? CH : TK_SCAN_LOADS_F32;
    hipLaunchKernelGGL((knn_tree_kernel_v<float, 3, K, kLower, CH, LD, BLK, WAVES>),
                       dim3((unsigned)((nq + BLK - 1) / BLK)), dim3(BLK), 0, s, tv, q, nq, order, k, dist, idx, orig);
    TK_HIP_CHECK(hipGetLastError());
}

template <int K, bool kLower>
bool launch_k(const TreeView<float>& tv, const float* q, int64_t nq, const int32_t* order, int k, float* dist,
              int32_t* idx, int64_t* orig, int variant, hipStream_t s) {
    switch (variant) {
        case 1: launch_v<K, kLower, 4, 128, 7>(tv, q, nq, order, k, dist, idx, orig, s); return true;
        case 2: launch_v<K, kLower, 8, 128, 7>(tv, q, nq, order, k, dist, idx, orig, s); return true;
        case 3: launch_v<K, kLower, 16, 128, 7>(tv, q, nq, order, k, dist, idx, orig, s); return true;
        case 4: launch_v<K, kLower, 8, 128, 5>(tv, q, nq, order, k, dist, idx, orig, s); return true;
        case 5: launch_v<K, kLower, 16, 128, 5>(tv, q, nq, order, k, dist, idx, orig, s); return true;
        case 6: launch_v<K, kLower, 16, 64, 7>(tv, q, nq, order, k, dist, idx, orig, s); return true;
        case 7: launch_v<K, kLower, 16, 256, 7>(tv, q, nq, order, k, dist, idx, orig, s); return true;
        case 8: launch_v<K, kLower, 8, 256, 7>(tv, q, nq, order, k, dist, idx, orig, s); return true;
        default: return false;
    }
}

}  // namespace

static_assert(kNumTreeVariants == 9, "launch_k covers every variant of the table");

bool launch_tree_variant_f32_d3(const Tree& t, const float* q, int64_t nq, const int32_t* order, int k, int kb,
                                float* dist, int32_t* idx, int64_t* orig, bool lower, int variant, hipStream_t s) {
    if (t.dims != 3 || variant <= 0 || (kb != 8 && kb != 16)) return false;
    const TreeView<float> tv = tree_view<float>(t);
    if (lower && tv.lower_direct) return false;  // the 64-bit keyed list's built-in kernel
    if (kb == 8)
        return lower ? launch_k<8, true>(tv, q, nq, order, k, dist, idx, orig, variant, s)
                     : launch_k<8, false>(tv, q, nq, order, k, dist, idx, orig, variant, s);
    return lower ? launch_k<16, true>(tv, q, nq, order, k, dist, idx, orig, variant, s)
                 : launch_k<16, false>(tv, q, nq, order, k, dist, idx, orig, variant, s);
}

}  // namespace tk
