// Batched-mode kernels for float (see batch_kernels.hpp).
#include "batch_kernels.hpp"

namespace tk {

void batch_presort_f32(const BatchPlan& p, const float* q, hipStream_t s) { batch_presort_t<float>(p, q, s); }

void batch_traverse_f32(const BatchPlan& p, const float* q, float* dist, int32_t* idx, int64_t* orig, bool emit,
                        const float* g, float* dq, hipStream_t s) {
    batch_traverse_t<float>(p, q, dist, idx, orig, emit, g, dq, s);
}

void batch_backward_f32(const BatchPlan& p, const float* q, const int32_t* idx, const float* g, float* dq, float* dr,
                        bool from_forward, hipStream_t s) {
    batch_backward_t<float>(p, q, idx, g, dq, dr, from_forward, s);
}

}  // namespace tk
