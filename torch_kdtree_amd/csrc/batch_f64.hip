// Batched-mode kernels for double (see batch_kernels.hpp).
#include "batch_kernels.hpp"

namespace tk {

void batch_presort_f64(const BatchPlan& p, const double* q, hipStream_t s) { batch_presort_t<double>(p, q, s); }

void batch_traverse_f64(const BatchPlan& p, const double* q, double* dist, int32_t* idx, int64_t* orig, bool emit,
                        const double* g, double* dq, hipStream_t s) {
    batch_traverse_t<double>(p, q, dist, idx, orig, emit, g, dq, s);
}

void batch_backward_f64(const BatchPlan& p, const double* q, const int32_t* idx, const double* g, double* dq, double* dr,
                        bool from_forward, hipStream_t s) {
    batch_backward_t<double>(p, q, idx, g, dq, dr, from_forward, s);
}

}  // namespace tk
