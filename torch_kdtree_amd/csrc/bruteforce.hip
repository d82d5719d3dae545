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

struct BfPlan {
    int nchunks;
    int32_t chunk;
    size_t part_bytes_d, part_bytes_i, total;
};

BfPlan bf_plan(int32_t n, int64_t nq, int k, size_t tsize) {
    BfPlan p{};
    const int64_t qblocks = (nq + kBfBlock - 1) / kBfBlock;
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
        default: throw std::runtime_error("brute force: nr_nns_searches > 32 is not supported");
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
