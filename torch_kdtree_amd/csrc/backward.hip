// Backward of the KNN squared distances (no counterpart kernel in the reference: there the
// gradient comes from torch autograd through the recomputation at nn_distance.py:111-112).
//
// For d_ij = sum_a (q_ia - r_{idx_ij, a})^2 and an upstream gradient g_ij:
//   dL/dq_i = sum_j g_ij * 2 (q_i - r_idx_ij)
//   dL/dr_m = -sum_{(i,j): idx_ij = m} g_ij * 2 (q_i - r_m)
// Each term is computed as g * (diff + diff) with diff = q - r rounded in T (the product
// autograd forms for pow(diff, 2)), and nothing from the forward is stored: the differences
// are recomputed from the coordinates.
//
// dL/dq: one lane per query, terms added in neighbour order j = 0..k-1.
// dL/dr: deterministic, no atomics.  The (neighbour index, flat position) pairs are sorted
// by neighbour index with a stable LSD radix sort (radix_sort.hip), so each reference point's contributions
// form a contiguous run in increasing flat position; one lane per run adds them in that
// order.  The result is bit-identical from run to run.
#include <memory>

#include "gpu_common.hpp"
#include "knn_gpu.hpp"

namespace tk {
namespace {

template <typename T, int D>
__global__ __launch_bounds__(256) void grad_q_kernel(const T* __restrict__ q, const T* __restrict__ ref, int64_t nq,
                                                     int k, const int64_t* __restrict__ idx,
                                                     const T* __restrict__ g, T* __restrict__ dq) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nq) return;
    T qi[D], acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        qi[a] = q[i * D + a];
        acc[a] = T(0);
    }
    for (int j = 0; j < k; ++j) {
        const int64_t m = idx[i * k + j];
        const T gij = g[i * k + j];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const T diff = qi[a] - ref[m * D + a];
            acc[a] = acc[a] + gij * (diff + diff);
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) dq[i * D + a] = acc[a];
}

__global__ __launch_bounds__(256) void pairs_kernel(const int64_t* __restrict__ idx, int64_t m,
                                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (f >= m) return;
    keys[f] = (uint32_t)idx[f];
    vals[f] = (uint32_t)f;
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = T(0);
}

template <typename T, int D>
__global__ __launch_bounds__(256) void grad_r_kernel(const T* __restrict__ q, const T* __restrict__ ref, int64_t m,
                                                     int k, const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ vals, const T* __restrict__ g,
                                                     T* __restrict__ dr) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= m) return;
    const uint32_t key = keys[p];
    if (p > 0 && keys[p - 1] == key) return;  // not the head of its run
    T r[D], acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        r[a] = ref[(int64_t)key * D + a];
        acc[a] = T(0);
    }
    for (int64_t e = p; e < m && keys[e] == key; ++e) {
        const uint32_t f = vals[e];
        const int64_t i = f / (uint32_t)k;
        const T gij = g[f];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const T diff = q[i * D + a] - r[a];
            acc[a] = acc[a] - gij * (diff + diff);
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) dr[(int64_t)key * D + a] = acc[a];
}

inline size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

int key_bits(int32_t n) {
    int b = 1;
    while (b < 32 && (int64_t(1) << b) < n) ++b;
    return b;
}

struct BwLayout {
    size_t keys, vals, keys_s, vals_s, temp, temp_bytes, total;
};

BwLayout bw_layout(int64_t nq, int32_t n, int k) {
    (void)n;
    const int64_t m = nq * k;
    BwLayout l{};
    size_t off = 0;
    l.keys = off; off += a256(4 * (size_t)m);
    l.vals = off; off += a256(4 * (size_t)m);
    l.keys_s = off; off += a256(4 * (size_t)m);
    l.vals_s = off; off += a256(4 * (size_t)m);
    l.temp = off;
    l.temp_bytes = radix_sort_workspace(m);
    off += a256(l.temp_bytes);
    l.total = off;
    return l;
}

template <typename T, int D>
void backward_typed(const T* q, const T* ref, int64_t nq, int32_t n, int k, const int64_t* idx, const T* g, T* dq,
                    T* dr, hipStream_t s, unsigned char* ws, const BwLayout& l) {
    const int64_t m = nq * k;
    if (dq) {
        hipLaunchKernelGGL((grad_q_kernel<T, D>), dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, q, ref, nq, k,
                           idx, g, dq);
        TK_HIP_CHECK(hipGetLastError());
    }
    if (!dr) return;
    hipLaunchKernelGGL((fill_kernel<T>), dim3((unsigned)(((int64_t)n * D + 255) / 256)), dim3(256), 0, s, dr,
                       (int64_t)n * D);
    TK_HIP_CHECK(hipGetLastError());
    if (m == 0) return;
    uint32_t* keys = reinterpret_cast<uint32_t*>(ws + l.keys);
    uint32_t* vals = reinterpret_cast<uint32_t*>(ws + l.vals);
    uint32_t* keys_s = reinterpret_cast<uint32_t*>(ws + l.keys_s);
    uint32_t* vals_s = reinterpret_cast<uint32_t*>(ws + l.vals_s);
    hipLaunchKernelGGL(pairs_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, idx, m, keys, vals);
    TK_HIP_CHECK(hipGetLastError());
    radix_sort_pairs(keys, vals, keys_s, vals_s, m, key_bits(n), ws + l.temp, l.temp_bytes, s);
    hipLaunchKernelGGL((grad_r_kernel<T, D>), dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, q, ref, m, k, keys_s,
                       vals_s, g, dr);
    TK_HIP_CHECK(hipGetLastError());
}

}  // namespace

size_t knn_backward_workspace(int64_t nq, int32_t n, int k) { return bw_layout(nq, n, k).total; }

void knn_backward_gpu(const void* queries, const void* ref, int64_t nq, int32_t n, int dims, int dtype, int k,
                      const int64_t* idx, const void* grad_dist, void* grad_q, void* grad_ref, hipStream_t stream,
                      void* workspace, size_t workspace_bytes) {
    if (!dims_supported(dims)) throw std::runtime_error("unsupported dims");
    if (k < 1 || n < 1) throw std::runtime_error("invalid backward shape");
    if (nq * (int64_t)k > (int64_t(1) << 28)) throw std::runtime_error("backward: nr_query * k must be <= 2^28");
    const BwLayout l = bw_layout(nq, n, k);
    std::unique_ptr<StreamScratch> own;
    unsigned char* ws = static_cast<unsigned char*>(workspace);
    if (grad_ref) {
        if (!ws) {
            own.reset(new StreamScratch(l.total, stream));
            ws = static_cast<unsigned char*>(own->ptr);
        } else if (workspace_bytes < l.total) {
            throw std::runtime_error("backward workspace too small");
        }
    }
    ProfScope prof("backward", stream);
#define TK_BW(TT, DD)                                                                                                 \
    backward_typed<TT, DD>(static_cast<const TT*>(queries), static_cast<const TT*>(ref), nq, n, k, idx,              \
                           static_cast<const TT*>(grad_dist), static_cast<TT*>(grad_q), static_cast<TT*>(grad_ref), \
                           stream, ws, l)
    if (dtype == kF32) {
        switch (dims) {
            case 1: TK_BW(float, 1); break;
            case 2: TK_BW(float, 2); break;
            case 3: TK_BW(float, 3); break;
            case 8: TK_BW(float, 8); break;
        }
    } else {
        switch (dims) {
            case 1: TK_BW(double, 1); break;
            case 2: TK_BW(double, 2); break;
            case 3: TK_BW(double, 3); break;
            case 8: TK_BW(double, 8); break;
        }
    }
#undef TK_BW
}

}  // namespace tk
