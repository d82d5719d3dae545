// Backward of the KNN squared distances (no counterpart kernel in the reference: there the
// gradient comes from torch autograd through the recomputation at nn_distance.py:111-112).
//
// For d_ij = sum_a (q_ia - r_{idx_ij, a})^2 and an upstream gradient g_ij:
//   dL/dq_i = sum_j g_ij * 2 (q_i - r_idx_ij)
//   dL/dr_m = -sum_{(i,j): idx_ij = m} g_ij * 2 (q_i - r_m)
// Each term is computed as g * (diff + diff) with diff = q - r rounded in T (the product
// autograd forms for pow(diff, 2)), and nothing from the forward is stored: the differences
// are recomputed from the coordinates (once, in the dL/dq pass, which also keeps each term for
// the dL/dr sums).
//
// dL/dq: one lane per query, terms added in neighbour order j = 0..k-1.
// dL/dr: deterministic, no atomics.  Queries are put in a canonical order: stably sorted by
// (Morton cell over the reference points' bounding box, row), with canonical_morton_bits(Q, d)
// bits per axis.  In that order each query's coordinates and upstream gradients are copied to
// consecutive slots and its (reference index, slot) pairs are written; a stable radix sort of the
// pairs by reference index then makes each reference point's contributions a contiguous run in
// canonical order, and one lane per reference point subtracts its run in that order.  Spatially close queries
// are close in the canonical order, so a run's reads are close together.  The result is
// bit-identical from run to run, and equal to the batched backward's on the same cloud.
//
// Structured form (knn_backward_structured_gpu, for the neighbours of a GPU tree query given by
// structured index): the points are first gathered into the tree's structured order, the pairs are
// keyed by structured position, and each position's sums are written to its original row.  The
// runs hold the same pairs in the same canonical order, so the results are the same bits; the
// gathers of spatially close queries then read close rows.
//
// Size: the pairs of one sort are at most 2^28 (the radix sort's limit).  With more, the canonical
// order is cut into consecutive slices of at most 2^28 / k queries, and each slice's runs are
// subtracted from the sums the previous slices left (the first slice starts from zero).  Every
// reference point's terms are then subtracted in the same canonical order as with one slice, one
// after another into the same accumulator, so the sums are bit-identical to an unsliced pass.
// dL/dq alone needs no sort and no workspace, at any size.
#include <algorithm>
#include <memory>

#include "gpu_common.hpp"
#include "knn_gpu.hpp"
#include "knn_tree.hpp"

namespace tk {
namespace {

// Gathers in flight per lane in the backward kernels (measured: 8 against 4 took the fwd+bwd
// workload's backward from 1.007 to 0.969 ms and the 256-cloud replay from 20.82 to 20.60 ms;
// 2 was slower than 4).
#ifndef TK_BW_UNROLL
#define TK_BW_UNROLL 8
#endif
constexpr int kBwUnroll = TK_BW_UNROLL;  // loads issued ahead per lane

// dL/dq, terms added in neighbour order.  Without q_sorted one lane per query row (nq rows).  With
// q_sorted (for dL/dr) lane t takes the query i = order[t0 + t] of canonical position t0 + t (a
// slice of nq positions), so that neighbouring lanes gather neighbouring reference rows, and also
// copies the query to q_sorted[t] and, per neighbour j, the upstream gradient and the pair
// (reference index, slot) to the slice's slot t * k + j, written contiguously across lanes.
template <typename T, int D, typename IdxT>
__global__ __launch_bounds__(256) void grad_q_kernel(const T* __restrict__ q, const T* __restrict__ ref, int64_t nq,
                                                     int k, const IdxT* __restrict__ idx,
                                                     const T* __restrict__ g, const uint32_t* __restrict__ order,
                                                     int64_t t0, T* __restrict__ dq, T* __restrict__ q_sorted,
                                                     T* __restrict__ g_sorted, uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals) {
    const int64_t lane = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (lane >= nq) return;
    const int64_t t = q_sorted ? lane : 0;
    const int64_t i = q_sorted ? (int64_t)order[t0 + lane] : lane;
    T qi[D], acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        qi[a] = q[i * D + a];
        acc[a] = T(0);
    }
    if (q_sorted) {
#pragma unroll
        for (int a = 0; a < D; ++a) q_sorted[t * D + a] = qi[a];
    }
    // loads kBwUnroll neighbours ahead; the terms are still added in neighbour order
    for (int j0 = 0; j0 < k; j0 += kBwUnroll) {
        int64_t m[kBwUnroll];
        T gg[kBwUnroll], rr[kBwUnroll][D];
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
            const bool ok = j0 + u < k;
            m[u] = ok ? idx[i * k + j0 + u] : 0;
            gg[u] = ok ? g[i * k + j0 + u] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
#pragma unroll
            for (int a = 0; a < D; ++a) rr[u][a] = j0 + u < k ? ref[m[u] * D + a] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
            const int j = j0 + u;
            if (j < k) {
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const T diff = qi[a] - rr[u][a];
                    acc[a] = acc[a] + gg[u] * (diff + diff);
                }
                if (q_sorted) {
                    const int64_t v = t * k + j;
                    g_sorted[v] = gg[u];
                    keys[v] = (uint32_t)m[u];
                    vals[v] = (uint32_t)v;
                }
            }
        }
    }
    if (dq) {
#pragma unroll
        for (int a = 0; a < D; ++a) dq[i * D + a] = acc[a];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = T(0);
}

__global__ __launch_bounds__(256) void zero_u32_kernel(uint32_t* __restrict__ a, uint32_t* __restrict__ b,
                                                       int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        a[i] = 0;
        b[i] = 0;
    }
}

// Bounds of each key's run in the sorted pairs: [run_start[key], run_end[key]) (both zero for keys
// that have no pairs).
__global__ __launch_bounds__(256) void run_bounds_kernel(const uint32_t* __restrict__ keys, int64_t m,
                                                         uint32_t* __restrict__ run_start,
                                                         uint32_t* __restrict__ run_end) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= m) return;
    const uint32_t key = keys[p];
    if (p == 0 || keys[p - 1] != key) run_start[key] = (uint32_t)p;
    if (p == m - 1 || keys[p + 1] != key) run_end[key] = (uint32_t)(p + 1);
}

// dL/dr, one lane per reference point: its run of sorted pairs, subtracted in canonical order,
// from zero (the first slice) or from the sums in dr (later slices).  With out_row the points are
// in structured order (lane s takes the point of tree position s) and the lane's sums go to row
// out_row[s] of dr; else lane m takes row m.
template <typename T, int D>
__global__ __launch_bounds__(256) void grad_r_kernel(const T* __restrict__ ref, int32_t n, int k,
                                                     const uint32_t* __restrict__ run_start,
                                                     const uint32_t* __restrict__ run_end,
                                                     const uint32_t* __restrict__ vals,
                                                     const T* __restrict__ q_sorted, const T* __restrict__ g_sorted,
                                                     bool accumulate, const int32_t* __restrict__ out_row,
                                                     T* __restrict__ dr) {
    const int64_t mi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (mi >= n) return;
    const int64_t row = out_row ? (int64_t)out_row[mi] : mi;
    const uint32_t e0 = run_start[mi], e1 = run_end[mi];
    T r[D], acc[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        r[a] = ref[mi * D + a];
        acc[a] = accumulate ? dr[row * D + a] : T(0);
    }
    // loads kBwUnroll pairs ahead; the terms are still subtracted in run order
    for (uint32_t e0u = e0; e0u < e1; e0u += kBwUnroll) {
        int64_t v[kBwUnroll];
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) v[u] = e0u + u < e1 ? (int64_t)vals[e0u + u] : 0;
        T gg[kBwUnroll], qq[kBwUnroll][D];
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
            const bool ok = e0u + u < e1;
            gg[u] = ok ? g_sorted[v[u]] : T(0);
#pragma unroll
            for (int a = 0; a < D; ++a) qq[u][a] = ok ? q_sorted[(v[u] / k) * D + a] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kBwUnroll; ++u) {
            if (e0u + u < e1) {
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const T diff = qq[u][a] - r[a];
                    acc[a] = acc[a] - gg[u] * (diff + diff);
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) dr[row * D + a] = acc[a];
}

// out[s] = in[row[s]] for rows of D coordinates: the reference points in the tree's structured
// order, so that the gathers of neighbouring queries' neighbours read neighbouring rows.
template <typename T, int D>
__global__ __launch_bounds__(256) void gather_rows_kernel(const T* __restrict__ in, const int32_t* __restrict__ row,
                                                          int32_t n, T* __restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const int64_t r = row[s];
#pragma unroll
    for (int a = 0; a < D; ++a) out[s * D + a] = in[r * D + a];
}

inline size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

int key_bits(int64_t n) {
    int b = 1;
    while (b < 32 && (int64_t(1) << b) < n) ++b;
    return b;
}

// Pairs per sort (the radix sort's limit); queries per slice = kMaxPairs / k.
constexpr int64_t kMaxPairs = int64_t(1) << 28;

int64_t slice_queries(int64_t nq, int k) { return std::min<int64_t>(nq, kMaxPairs / k); }

struct BwLayout {
    size_t qkey, qrow, qkey_s, order, run_start, run_end, bbox, q_sorted, g_sorted, keys, vals, keys_s, vals_s,
        ref_s, temp, temp_bytes, total;
};

// Workspace of dL/dr: the canonical order of all nq queries, and the buffers of one slice; with
// `structured` also the reference points in structured order.
BwLayout bw_layout(int64_t nq, int32_t n, int k, int dims, int dtype, bool structured = false) {
    const int64_t sq = slice_queries(nq, k);
    const int64_t m = sq * k;
    const size_t ts = dtype_size(dtype);
    BwLayout l{};
    size_t off = 0;
    l.qkey = off; off += a256(4 * (size_t)nq);
    l.qrow = off; off += a256(4 * (size_t)nq);
    l.qkey_s = off; off += a256(4 * (size_t)nq);
    l.order = off; off += a256(4 * (size_t)nq);
    l.run_start = off; off += a256(4 * (size_t)n);
    l.run_end = off; off += a256(4 * (size_t)n);
    l.bbox = off; off += a256(16 * (size_t)dims);
    l.q_sorted = off; off += a256(ts * (size_t)dims * (size_t)sq);
    l.g_sorted = off; off += a256(ts * (size_t)m);
    l.keys = off; off += a256(4 * (size_t)m);
    l.vals = off; off += a256(4 * (size_t)m);
    l.keys_s = off; off += a256(4 * (size_t)m);
    l.vals_s = off; off += a256(4 * (size_t)m);
    l.ref_s = off; off += structured ? a256(ts * (size_t)dims * (size_t)n) : 0;
    l.temp = off;
    l.temp_bytes = std::max(radix_sort_workspace(m), radix_sort_workspace(nq));
    off += a256(l.temp_bytes);
    l.total = off;
    return l;
}

// idx: the neighbours' rows of `ref` (original indices, int64), or with `shuffled` their
// structured positions (int32), shuffled[s] being the original row of structured position s.  In
// the structured form the points are first gathered into structured order, every pair is keyed by
// its structured position and dL/dr is summed per structured position and written to its original
// row: the same terms in the same order, so the same bits, with the gathers of spatially close
// queries reading close rows.
template <typename T, int D, typename IdxT>
void backward_typed(const T* q, const T* ref, int64_t nq, int32_t n, int k, const IdxT* idx,
                    const int32_t* shuffled, const T* g, T* dq, T* dr, hipStream_t s, unsigned char* ws,
                    const BwLayout& l) {
    const int64_t m = nq * k;
    const bool pairs = dr && m > 0;
    const dim3 qgrid((unsigned)((nq + 255) / 256));
    if (dr && !pairs) {
        hipLaunchKernelGGL((fill_kernel<T>), dim3((unsigned)(((int64_t)n * D + 255) / 256)), dim3(256), 0, s, dr,
                           (int64_t)n * D);
        TK_HIP_CHECK(hipGetLastError());
    }
    if (nq == 0 || !(dq || pairs)) return;
    const T* rows = ref;
    if (shuffled) {
        T* ref_s = reinterpret_cast<T*>(ws + l.ref_s);
        hipLaunchKernelGGL((gather_rows_kernel<T, D>), dim3((unsigned)(((int64_t)n + 255) / 256)), dim3(256), 0, s, ref,
                           shuffled, n, ref_s);
        TK_HIP_CHECK(hipGetLastError());
        rows = ref_s;
    }
    if (!pairs) {
        hipLaunchKernelGGL((grad_q_kernel<T, D, IdxT>), qgrid, dim3(256), 0, s, q, rows, nq, k, idx, g, nullptr,
                           (int64_t)0, dq, nullptr, nullptr, nullptr, nullptr);
        TK_HIP_CHECK(hipGetLastError());
        return;
    }
    // canonical query order
    uint32_t* qkey = reinterpret_cast<uint32_t*>(ws + l.qkey);
    uint32_t* qrow = reinterpret_cast<uint32_t*>(ws + l.qrow);
    uint32_t* qkey_s = reinterpret_cast<uint32_t*>(ws + l.qkey_s);
    uint32_t* order = reinterpret_cast<uint32_t*>(ws + l.order);
    auto* box = reinterpret_cast<OrderedT<T>*>(ws + l.bbox);
    init_bbox(box, sizeof(T) == 4 ? kF32 : kF64, D, s);
    const unsigned box_blocks = (unsigned)std::min<int64_t>(1024, ((int64_t)n + kBoxBlock - 1) / kBoxBlock);
    hipLaunchKernelGGL((bbox_kernel<T, D>), dim3(box_blocks), dim3(kBoxBlock), 0, s, ref, (int64_t)n, box);
    const int bits = canonical_morton_bits(nq, D);
    hipLaunchKernelGGL((morton_bucket_kernel<T, D>), qgrid, dim3(256), 0, s, q, nq, box, bits, qkey, qrow, false);
    TK_HIP_CHECK(hipGetLastError());
    radix_sort_pairs(qkey, qrow, qkey_s, order, nq, bits * D, ws + l.temp, l.temp_bytes, s);
    // dL/dq, canonical copies and pairs; pairs sorted by reference index; dL/dr per reference point
    T* q_sorted = reinterpret_cast<T*>(ws + l.q_sorted);
    T* g_sorted = reinterpret_cast<T*>(ws + l.g_sorted);
    uint32_t* keys = reinterpret_cast<uint32_t*>(ws + l.keys);
    uint32_t* vals = reinterpret_cast<uint32_t*>(ws + l.vals);
    uint32_t* keys_s = reinterpret_cast<uint32_t*>(ws + l.keys_s);
    uint32_t* vals_s = reinterpret_cast<uint32_t*>(ws + l.vals_s);
    uint32_t* run_start = reinterpret_cast<uint32_t*>(ws + l.run_start);
    uint32_t* run_end = reinterpret_cast<uint32_t*>(ws + l.run_end);
    const dim3 rgrid((unsigned)(((int64_t)n + 255) / 256));
    const int64_t sq = slice_queries(nq, k);
    for (int64_t t0 = 0; t0 < nq; t0 += sq) {
        const int64_t cnt = std::min(sq, nq - t0);
        const int64_t mm = cnt * k;
        hipLaunchKernelGGL((grad_q_kernel<T, D, IdxT>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, q, rows,
                           cnt, k, idx, g, order, t0, dq, q_sorted, g_sorted, keys, vals);
        TK_HIP_CHECK(hipGetLastError());
        radix_sort_pairs(keys, vals, keys_s, vals_s, mm, key_bits(n), ws + l.temp, l.temp_bytes, s);
        hipLaunchKernelGGL(zero_u32_kernel, rgrid, dim3(256), 0, s, run_start, run_end, (int64_t)n);
        hipLaunchKernelGGL(run_bounds_kernel, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, s, keys_s, mm,
                           run_start, run_end);
        hipLaunchKernelGGL((grad_r_kernel<T, D>), rgrid, dim3(256), 0, s, rows, n, k, run_start, run_end, vals_s,
                           q_sorted, g_sorted, t0 > 0, shuffled, dr);
        TK_HIP_CHECK(hipGetLastError());
    }
}

}  // namespace

size_t knn_backward_workspace(int64_t nq, int32_t n, int k, int dims, int dtype) {
    return bw_layout(nq, n, k, dims, dtype).total;
}

void knn_backward_gpu(const void* queries, const void* ref, int64_t nq, int32_t n, int dims, int dtype, int k,
                      const int64_t* idx, const void* grad_dist, void* grad_q, void* grad_ref, hipStream_t stream,
                      void* workspace, size_t workspace_bytes) {
    if (!dims_supported(dims)) throw std::runtime_error("unsupported dims");
    if (k < 1 || n < 1) throw std::runtime_error("invalid backward shape");
    if (nq < 0) throw std::runtime_error("invalid backward shape");
    if (grad_ref && nq > kMaxPairs)
        throw std::runtime_error("backward: the gradient of the reference points takes at most 2^28 queries");
    const BwLayout l = bw_layout(nq, n, k, dims, dtype);
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
#define TK_BW(TT, DD)                                                                                           \
    backward_typed<TT, DD, int64_t>(static_cast<const TT*>(queries), static_cast<const TT*>(ref), nq, n, k, idx,   \
                                    nullptr, static_cast<const TT*>(grad_dist), static_cast<TT*>(grad_q),           \
                                    static_cast<TT*>(grad_ref), stream, ws, l)
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

size_t knn_backward_structured_workspace(int64_t nq, int32_t n, int k, int dims, int dtype) {
    return bw_layout(nq, n, k, dims, dtype, true).total;
}

void knn_backward_structured_gpu(const void* queries, const void* ref, int64_t nq, int32_t n, int dims, int dtype,
                                 int k, const int32_t* sidx, const int32_t* shuffled, const void* grad_dist,
                                 void* grad_q, void* grad_ref, hipStream_t stream, void* workspace,
                                 size_t workspace_bytes) {
    if (!dims_supported(dims)) throw std::runtime_error("unsupported dims");
    if (k < 1 || n < 1) throw std::runtime_error("invalid backward shape");
    if (nq < 0) throw std::runtime_error("invalid backward shape");
    if (!shuffled) throw std::runtime_error("backward: no structured-to-original map");
    if (grad_ref && nq > kMaxPairs)
        throw std::runtime_error("backward: the gradient of the reference points takes at most 2^28 queries");
    const BwLayout l = bw_layout(nq, n, k, dims, dtype, true);
    std::unique_ptr<StreamScratch> own;
    unsigned char* ws = static_cast<unsigned char*>(workspace);
    if (!ws) {
        own.reset(new StreamScratch(l.total, stream));
        ws = static_cast<unsigned char*>(own->ptr);
    } else if (workspace_bytes < l.total) {
        throw std::runtime_error("backward workspace too small");
    }
    ProfScope prof("backward", stream);
#define TK_BW(TT, DD)                                                                                           \
    backward_typed<TT, DD, int32_t>(static_cast<const TT*>(queries), static_cast<const TT*>(ref), nq, n, k, sidx,  \
                                    shuffled, static_cast<const TT*>(grad_dist), static_cast<TT*>(grad_q),          \
                                    static_cast<TT*>(grad_ref), stream, ws, l)
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
