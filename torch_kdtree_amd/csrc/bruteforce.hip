// Exhaustive KNN (the reference's NmDistanceKernel[Dynamic], src/nndistance_g.cu:6-140,
// generalised to any supported d and run over the whole GPU instead of one block).
//
// Semantics: reference points are scanned in index order; a candidate enters the k-list iff
// d < worst (src/nndistance_g.cu:57,126), at the upper-bound position; so ties go to the lower
// reference index, a NaN distance never enters, and the result is the k smallest (distance,
// index) pairs among the non-NaN distances.  Slots left empty (a query with a NaN coordinate, or
// fewer than k non-NaN distances) hold (+inf, 0); the reference leaves their index unset.  Distances use the CPU path's arithmetic
// (q - r per axis, squares summed left to right, no fused multiply-add).
//
// Layout: blockIdx.x picks a tile of kBfBlock queries (one per lane), blockIdx.y a chunk of
// the reference set.  Reference points stream through LDS in tiles of kBfTile points and are
// read back as wave-wide broadcasts.  With more than one chunk, every chunk writes its partial
// k-list and a merge kernel inserts the chunk lists in chunk order, which keeps the
// (distance, index) order because chunk c only holds indices above those of chunk c-1.
#include <algorithm>
#include <memory>

#include "gpu_common.hpp"
#include "knn_gpu.hpp"

namespace tk {
namespace {

#ifndef TK_BF_BLOCK
#define TK_BF_BLOCK 256
#endif
#ifndef TK_BF_CHUNK
#define TK_BF_CHUNK 16
#endif
constexpr int kBfBlock = TK_BF_BLOCK;
constexpr int kBfTile = 512;
constexpr int kBfTargetBlocks = 4096;

constexpr int kBfChunk = TK_BF_CHUNK;  // candidates tested against one stale worst before their insertions

// Elements per point in the LDS tile: d = 3 is padded to 4 so a point is one 16-byte read.
template <int D>
constexpr int bf_stride() { return D == 3 ? 4 : D; }

template <typename T, int D, int K>
__global__ __launch_bounds__(kBfBlock) void bf_kernel(const T* __restrict__ ref, int32_t n, const T* __restrict__ queries,
                                                      int64_t nq, int k, int32_t chunk, T* __restrict__ out_d,
                                                      int32_t* __restrict__ out_i) {
    constexpr int S = bf_stride<D>();
    __shared__ T tile[kBfTile * S];
    __shared__ T stage[kBfChunk * kBfBlock];
    const int64_t qi = (int64_t)blockIdx.x * kBfBlock + threadIdx.x;
    const bool active = qi < nq;
    T q[D];
#pragma unroll
    for (int i = 0; i < D; ++i) q[i] = active ? queries[qi * D + i] : T(0);
    KBest<T, K, false> kb;
    kb.init(k);
    T* st = stage + threadIdx.x;
    const int32_t cbeg = (int32_t)blockIdx.y * chunk;
    const int32_t cend = (int32_t)min((int64_t)n, (int64_t)cbeg + chunk);
    for (int32_t t0 = cbeg; t0 < cend; t0 += kBfTile) {
        const int32_t cnt = min(kBfTile, cend - t0);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * D; e += kBfBlock) tile[(e / D) * S + e % D] = ref[(int64_t)t0 * D + e];
        __syncthreads();
        if (active) {
            // chunks of kBfChunk candidates: distances against the worst at the chunk's start, then
            // the accepted ones inserted in index order (a candidate the stale worst rejects is
            // rejected by every later worst too, so the list ends as a one-by-one scan leaves it)
            for (int32_t j0 = 0; j0 < cnt; j0 += kBfChunk) {
                const T w = kb.worst();
                uint32_t mask = 0;
#pragma unroll
                for (int c = 0; c < kBfChunk; ++c) {
                    const int32_t j = min(j0 + c, cnt - 1);
                    T r[D];
#pragma unroll
                    for (int i = 0; i < D; ++i) r[i] = tile[j * S + i];
                    const T dd = sq_dist<T, D>(q, r);
                    mask |= (dd < w && j0 + c < cnt) ? (1u << c) : 0u;
                    st[c * kBfBlock] = dd;
                }
                while (mask) {
                    const int c = __builtin_ctz(mask);
                    mask &= mask - 1u;
                    const T dd = st[c * kBfBlock];
                    const bool ok = dd < kb.worst();
                    kb.insert(ok ? dd : pos_inf<T>(), ok ? t0 + j0 + c : 0x7fffffff);
                }
            }
        }
    }
    if (!active) return;
    const int pad = K - k;
    const int64_t row = ((int64_t)blockIdx.y * nq + qi) * k;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j >= pad) {
            out_d[row + j - pad] = kb.d[j];
            out_i[row + j - pad] = kb.x[j];
        }
    }
}

template <typename T, int K>
__global__ __launch_bounds__(256) void bf_merge_kernel(const T* __restrict__ part_d, const int32_t* __restrict__ part_i,
                                                       int nchunks, int64_t nq, int k, T* __restrict__ out_d,
                                                       int32_t* __restrict__ out_i) {
    const int64_t qi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (qi >= nq) return;
    KBest<T, K, false> kb;
    kb.init(k);
    for (int c = 0; c < nchunks; ++c) {
        const int64_t row = ((int64_t)c * nq + qi) * k;
        for (int j = 0; j < k; ++j) {
            const T dd = part_d[row + j];
            if (!(dd < kb.worst())) break;  // the chunk list is sorted: nothing further enters
            kb.insert(dd, part_i[row + j]);
        }
    }
    const int pad = K - k;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j >= pad) {
            out_d[qi * k + j - pad] = kb.d[j];
            out_i[qi * k + j - pad] = kb.x[j];
        }
    }
}

// Lists of 33..128 entries (nndistance_g.cu:74-140 takes any k, its lists in global memory): the
// list of each lane lives in LDS, slot j of lane t at [j * kBfLdsBlock + t], as a max-heap on
// (distance, index), so the worst entry is the root.  The same scan as bf_kernel: tiles of
// reference points through LDS, chunks of kBfChunk distances tested against the worst at the
// chunk's start, then offered in index order.  An offered candidate is taken iff d < worst (the
// reference's test): it replaces the root, which then sifts down.  Since candidates come in
// increasing index order, a candidate never ties the root's (distance, index), so the heap holds
// exactly the k smallest (distance, index) pairs a sorted list would; a final heapsort orders them.
// Empty slots are (+inf, 0).
constexpr int kBfLdsBlock = 64;

template <typename T>
struct BfLdsList {
    T* d;
    int32_t* x;
    int k;
    T w;  // cached worst (the root's distance)

    __device__ __forceinline__ void init() {
        for (int j = 0; j < k; ++j) {
            d[j * kBfLdsBlock] = pos_inf<T>();
            x[j * kBfLdsBlock] = 0;
        }
        w = pos_inf<T>();
    }
    static __device__ __forceinline__ bool after(T da, int32_t xa, T db, int32_t xb) {
        return da > db || (da == db && xa > xb);
    }
    // Puts (dd, idx) at slot i of the heap [0, size) and restores the heap order below it.
    __device__ __forceinline__ void sift_down(int i, int size, T dd, int32_t idx) {
        for (;;) {
            const int c = 2 * i + 1;
            if (c >= size) break;
            int m = c;
            T dm = d[c * kBfLdsBlock];
            int32_t xm = x[c * kBfLdsBlock];
            if (c + 1 < size) {
                const T d2 = d[(c + 1) * kBfLdsBlock];
                const int32_t x2 = x[(c + 1) * kBfLdsBlock];
                if (after(d2, x2, dm, xm)) {
                    m = c + 1;
                    dm = d2;
                    xm = x2;
                }
            }
            if (!after(dm, xm, dd, idx)) break;
            d[i * kBfLdsBlock] = dm;
            x[i * kBfLdsBlock] = xm;
            i = m;
        }
        d[i * kBfLdsBlock] = dd;
        x[i * kBfLdsBlock] = idx;
    }
    __device__ __forceinline__ void insert(T dd, int32_t idx) {
        if (!(dd < w)) return;
        sift_down(0, k, dd, idx);
        w = d[0];
    }
    // Heapsort: ascending (distance, index) in slots 0..k-1.
    __device__ __forceinline__ void sort() {
        for (int end = k - 1; end > 0; --end) {
            const T de = d[end * kBfLdsBlock];
            const int32_t xe = x[end * kBfLdsBlock];
            d[end * kBfLdsBlock] = d[0];
            x[end * kBfLdsBlock] = x[0];
            sift_down(0, end, de, xe);
        }
    }
};

template <typename T>
inline size_t bf_lds_list_bytes(int k) { return (size_t)k * kBfLdsBlock * (sizeof(T) + 4); }

template <typename T, int D>
__global__ __launch_bounds__(kBfLdsBlock) void bf_lds_kernel(const T* __restrict__ ref, int32_t n,
                                                             const T* __restrict__ queries, int64_t nq, int k,
                                                             int32_t chunk, T* __restrict__ out_d,
                                                             int32_t* __restrict__ out_i) {
    constexpr int S = bf_stride<D>();
    extern __shared__ unsigned char bf_smem[];
    __shared__ T tile[kBfTile * S];
    __shared__ T stage[kBfChunk * kBfLdsBlock];
    const int64_t qi = (int64_t)blockIdx.x * kBfLdsBlock + threadIdx.x;
    const bool active = qi < nq;
    T q[D];
#pragma unroll
    for (int i = 0; i < D; ++i) q[i] = active ? queries[qi * D + i] : T(0);
    BfLdsList<T> kb{reinterpret_cast<T*>(bf_smem) + threadIdx.x,
                    reinterpret_cast<int32_t*>(reinterpret_cast<T*>(bf_smem) + (size_t)k * kBfLdsBlock) + threadIdx.x,
                    k, T(0)};
    kb.init();
    T* st = stage + threadIdx.x;
    const int32_t cbeg = (int32_t)blockIdx.y * chunk;
    const int32_t cend = (int32_t)min((int64_t)n, (int64_t)cbeg + chunk);
    for (int32_t t0 = cbeg; t0 < cend; t0 += kBfTile) {
        const int32_t cnt = min(kBfTile, cend - t0);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * D; e += kBfLdsBlock) tile[(e / D) * S + e % D] = ref[(int64_t)t0 * D + e];
        __syncthreads();
        if (active) {
            for (int32_t j0 = 0; j0 < cnt; j0 += kBfChunk) {
                const T w = kb.w;
                uint32_t mask = 0;
#pragma unroll
                for (int c = 0; c < kBfChunk; ++c) {
                    const int32_t j = min(j0 + c, cnt - 1);
                    T r[D];
#pragma unroll
                    for (int i = 0; i < D; ++i) r[i] = tile[j * S + i];
                    const T dd = sq_dist<T, D>(q, r);
                    mask |= (dd < w && j0 + c < cnt) ? (1u << c) : 0u;
                    st[c * kBfLdsBlock] = dd;
                }
                while (mask) {
                    const int c = __builtin_ctz(mask);
                    mask &= mask - 1u;
                    kb.insert(st[c * kBfLdsBlock], t0 + j0 + c);
                }
            }
        }
    }
    if (!active) return;
    kb.sort();
    const int64_t row = ((int64_t)blockIdx.y * nq + qi) * k;
    for (int j = 0; j < k; ++j) {
        out_d[row + j] = kb.d[j * kBfLdsBlock];
        out_i[row + j] = kb.x[j * kBfLdsBlock];
    }
}

template <typename T>
__global__ __launch_bounds__(kBfLdsBlock) void bf_merge_lds_kernel(const T* __restrict__ part_d,
                                                                   const int32_t* __restrict__ part_i, int nchunks,
                                                                   int64_t nq, int k, T* __restrict__ out_d,
                                                                   int32_t* __restrict__ out_i) {
    extern __shared__ unsigned char bf_smem[];
    const int64_t qi = (int64_t)blockIdx.x * kBfLdsBlock + threadIdx.x;
    if (qi >= nq) return;
    BfLdsList<T> kb{reinterpret_cast<T*>(bf_smem) + threadIdx.x,
                    reinterpret_cast<int32_t*>(reinterpret_cast<T*>(bf_smem) + (size_t)k * kBfLdsBlock) + threadIdx.x,
                    k, T(0)};
    kb.init();
    for (int c = 0; c < nchunks; ++c) {
        const int64_t row = ((int64_t)c * nq + qi) * k;
        for (int j = 0; j < k; ++j) {
            const T dd = part_d[row + j];
            if (!(dd < kb.w)) break;  // the chunk list is sorted: nothing further enters
            kb.insert(dd, part_i[row + j]);
        }
    }
    kb.sort();
    for (int j = 0; j < k; ++j) {
        out_d[qi * k + j] = kb.d[j * kBfLdsBlock];
        out_i[qi * k + j] = kb.x[j * kBfLdsBlock];
    }
}

template <typename T, int D>
void bf_launch_lds(const T* ref, int32_t n, const T* q, int64_t nq, int k, T* dist, int32_t* idx, hipStream_t s,
                   unsigned char* ws, int nchunks, int32_t chunk, size_t part_bytes_d) {
    const size_t shmem = bf_lds_list_bytes<T>(k);
    const void* fk = reinterpret_cast<const void*>(&bf_lds_kernel<T, D>);
    const void* fm = reinterpret_cast<const void*>(&bf_merge_lds_kernel<T>);
    constexpr size_t kStatic = (size_t)kBfTile * bf_stride<D>() * sizeof(T) + (size_t)kBfChunk * kBfLdsBlock * sizeof(T);
    if (shmem + kStatic > 65536) TK_HIP_CHECK(hipFuncSetAttribute(fk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    if (shmem > 65536) TK_HIP_CHECK(hipFuncSetAttribute(fm, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    const dim3 grid((unsigned)((nq + kBfLdsBlock - 1) / kBfLdsBlock), (unsigned)nchunks);
    if (nchunks == 1) {
        hipLaunchKernelGGL((bf_lds_kernel<T, D>), grid, dim3(kBfLdsBlock), shmem, s, ref, n, q, nq, k, chunk, dist, idx);
        TK_HIP_CHECK(hipGetLastError());
    } else {
        T* pd = reinterpret_cast<T*>(ws);
        int32_t* pi = reinterpret_cast<int32_t*>(ws + part_bytes_d);
        hipLaunchKernelGGL((bf_lds_kernel<T, D>), grid, dim3(kBfLdsBlock), shmem, s, ref, n, q, nq, k, chunk, pd, pi);
        TK_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL((bf_merge_lds_kernel<T>), dim3((unsigned)((nq + kBfLdsBlock - 1) / kBfLdsBlock)),
                           dim3(kBfLdsBlock), shmem, s, pd, pi, nchunks, nq, k, dist, idx);
        TK_HIP_CHECK(hipGetLastError());
    }
}

struct BfPlan {
    int nchunks;
    int32_t chunk;
    size_t part_bytes_d, part_bytes_i, total;
};

BfPlan bf_plan(int32_t n, int64_t nq, int k, size_t tsize) {
    BfPlan p{};
    const int block = k > 32 ? kBfLdsBlock : kBfBlock;  // lists beyond 32 entries: bf_lds_kernel
    const int64_t qblocks = (nq + block - 1) / block;
    const int64_t max_chunks = std::max<int64_t>(1, (n + kBfTile - 1) / kBfTile);
    int64_t nc = std::max<int64_t>(1, (kBfTargetBlocks + qblocks - 1) / qblocks);
    nc = std::min<int64_t>(nc, max_chunks);
    nc = std::min<int64_t>(nc, 65535);
    int64_t chunk = (n + nc - 1) / nc;
    chunk = ((chunk + kBfTile - 1) / kBfTile) * kBfTile;
    nc = (n + chunk - 1) / chunk;
    p.nchunks = (int)nc;
    p.chunk = (int32_t)chunk;
    if (p.nchunks > 1) {
        p.part_bytes_d = ((size_t)nc * nq * k * tsize + 255) & ~size_t(255);
        p.part_bytes_i = ((size_t)nc * nq * k * 4 + 255) & ~size_t(255);
    }
    p.total = p.part_bytes_d + p.part_bytes_i;
    return p;
}

template <typename T, int D, int K>
void bf_launch(const T* ref, int32_t n, const T* q, int64_t nq, int k, T* dist, int32_t* idx, hipStream_t s,
               unsigned char* ws, const BfPlan& p) {
    const dim3 grid((unsigned)((nq + kBfBlock - 1) / kBfBlock), (unsigned)p.nchunks);
    if (p.nchunks == 1) {
        hipLaunchKernelGGL((bf_kernel<T, D, K>), grid, dim3(kBfBlock), 0, s, ref, n, q, nq, k, p.chunk, dist, idx);
        TK_HIP_CHECK(hipGetLastError());
    } else {
        T* pd = reinterpret_cast<T*>(ws);
        int32_t* pi = reinterpret_cast<int32_t*>(ws + p.part_bytes_d);
        hipLaunchKernelGGL((bf_kernel<T, D, K>), grid, dim3(kBfBlock), 0, s, ref, n, q, nq, k, p.chunk, pd, pi);
        TK_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL((bf_merge_kernel<T, K>), dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, pd, pi,
                           p.nchunks, nq, k, dist, idx);
        TK_HIP_CHECK(hipGetLastError());
    }
}

template <typename T, int D>
void bf_dispatch_k(const T* ref, int32_t n, const T* q, int64_t nq, int k, T* dist, int32_t* idx, hipStream_t s,
                   unsigned char* ws, const BfPlan& p) {
    int kb = 1;
    while (kb < k) kb <<= 1;
    switch (kb) {
        case 1: bf_launch<T, D, 1>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        case 2: bf_launch<T, D, 2>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        case 4: bf_launch<T, D, 4>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        case 8: bf_launch<T, D, 8>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        case 16: bf_launch<T, D, 16>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        case 32: bf_launch<T, D, 32>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        default: bf_launch_lds<T, D>(ref, n, q, nq, k, dist, idx, s, ws, p.nchunks, p.chunk, p.part_bytes_d); break;
    }
}

template <typename T>
void bf_dispatch(const T* ref, int32_t n, const T* q, int64_t nq, int dims, int k, T* dist, int32_t* idx,
                 hipStream_t s, unsigned char* ws, const BfPlan& p) {
    switch (dims) {
        case 1: bf_dispatch_k<T, 1>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        case 2: bf_dispatch_k<T, 2>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        case 3: bf_dispatch_k<T, 3>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        case 8: bf_dispatch_k<T, 8>(ref, n, q, nq, k, dist, idx, s, ws, p); break;
        default: throw std::runtime_error("unsupported dims");
    }
}

}  // namespace

size_t bruteforce_gpu_workspace(int32_t n, int64_t nq, int dims, int dtype, int k) {
    (void)dims;
    return bf_plan(n, nq, k, dtype_size(dtype)).total;
}

void bruteforce_gpu(const void* ref, int32_t n, const void* queries, int64_t nq, int dims, int dtype, int k, void* dist,
                    int32_t* idx, hipStream_t stream, void* workspace, size_t workspace_bytes) {
    if (!dims_supported(dims)) throw std::runtime_error("unsupported dims");
    if (k < 1) throw std::runtime_error("nr_nns_searches must be >= 1");
    if (n < 1 || k > n) throw std::runtime_error("You requested more nearest neighbors than there are reference points");
    if (k > kMaxGpuK)
        throw std::runtime_error("brute force: nr_nns_searches > " + std::to_string(kMaxGpuK) + " is not supported");
    if (nq == 0) return;
    const BfPlan p = bf_plan(n, nq, k, dtype_size(dtype));
    std::unique_ptr<StreamScratch> own;
    unsigned char* ws = static_cast<unsigned char*>(workspace);
    if (p.total) {
        if (!ws) {
            own.reset(new StreamScratch(p.total, stream));
            ws = static_cast<unsigned char*>(own->ptr);
        } else if (workspace_bytes < p.total) {
            throw std::runtime_error("brute-force workspace too small");
        }
    }
    ProfScope prof("bruteforce", stream);
    if (dtype == kF32)
        bf_dispatch<float>(static_cast<const float*>(ref), n, static_cast<const float*>(queries), nq, dims, k,
                           static_cast<float*>(dist), idx, stream, ws, p);
    else
        bf_dispatch<double>(static_cast<const double*>(ref), n, static_cast<const double*>(queries), nq, dims, k,
                            static_cast<double*>(dist), idx, stream, ws, p);
}

}  // namespace tk
