// Stable least-significant-digit radix sort of (uint32 key, uint32 value) pairs, 8 bits per pass.
//
// Each pass over n pairs:
//   1. radix_hist_kernel: every block counts the digits of its tile (kRadixItems items per
//      thread, 256 threads) in LDS and writes count[digit * n_blocks + block];
//   2. an exclusive scan of those counts gives, for each (digit, block), the output position of
//      the block's first pair with that digit (all smaller digits, then earlier blocks);
//   3. radix_scatter_kernel: every wavefront of a block ranks its quarter of the tile, in index
//      order, 64 pairs per round (lanes holding the same digit are found with one ballot per
//      digit bit; per-digit running counts live in the wavefront's own LDS row).  A pair's slot
//      in the tile is its digit's start in the tile, plus the pairs of its digit in lower
//      wavefronts, plus its rank within its wavefront.  The tile, rearranged so in LDS, is then
//      written out slot by slot: each digit's run goes to consecutive output positions from the
//      block's offset for that digit.  The order of equal digits is the input order: the sort is
//      stable, and the result does not depend on scheduling.
// Nothing is cleared with a memset and nothing waits on the host, so the sort can be captured in
// a HIP graph and replayed.
#include <algorithm>

#include "gpu_common.hpp"
#include "knn_gpu.hpp"

namespace tk {
namespace {

constexpr int kRadixBlock = 256;
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kRadixBlock * kRadixItems;
constexpr int kRadixBins = 256;

__global__ __launch_bounds__(kRadixBlock) void radix_hist_kernel(const uint32_t* __restrict__ keys, int64_t n,
                                                                  int shift, int32_t n_blocks,
                                                                  int32_t* __restrict__ hist) {
    __shared__ int32_t cnt[kRadixBins];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRadixTile;
#pragma unroll
    for (int s = 0; s < kRadixItems; ++s) {
        const int64_t i = base + (int64_t)s * kRadixBlock + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kRadixBins - 1)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * n_blocks + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(kRadixBlock) void radix_scatter_kernel(const uint32_t* __restrict__ keys_in,
                                                                     const uint32_t* __restrict__ vals_in, int64_t n,
                                                                     int shift, int32_t n_blocks,
                                                                     const int32_t* __restrict__ hist,
                                                                     const int32_t* __restrict__ offsets,
                                                                     uint32_t* __restrict__ keys_out,
                                                                     uint32_t* __restrict__ vals_out) {
    constexpr int kWaves = kRadixBlock / 64;
    constexpr int kRounds = kRadixTile / kRadixBlock;  // rounds of 64 pairs per wavefront
    __shared__ uint32_t tile_k[kRadixTile];           // the tile in digit order (stable)
    __shared__ uint32_t tile_v[kRadixTile];
    __shared__ int32_t tile_start[kRadixBins + 1];    // first tile slot of each digit
    __shared__ int32_t wave_cnt[kWaves][kRadixBins];  // per wavefront: pairs of each digit, then their first slot
    __shared__ int32_t scan_tmp[kWaves];
    __shared__ int32_t blk_off[kRadixBins];           // this block's output start of each digit
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int w = t >> 6;
    blk_off[t] = offsets[(int64_t)t * n_blocks + blockIdx.x];
    // tile_start: exclusive scan of this tile's digit counts
    {
        const int c = hist[(int64_t)t * n_blocks + blockIdx.x];
        int inc = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(inc, off, 64);
            if (lane >= off) inc += y;
        }
        if (lane == 63) scan_tmp[w] = inc;
        __syncthreads();
        int base = 0;
        for (int v = 0; v < w; ++v) base += scan_tmp[v];
        tile_start[t] = base + inc - c;
        if (t == kRadixBlock - 1) tile_start[kRadixBins] = base + inc;
    }
#pragma unroll
    for (int v = 0; v < kWaves; ++v) wave_cnt[v][t] = 0;
    __syncthreads();
    // Wavefront w ranks the pairs [w * kRounds * 64, (w + 1) * kRounds * 64) of the tile, 64 per
    // round: a pair's rank is the pairs of its digit in the wavefront's earlier rounds (its
    // running count in LDS) plus those in lower lanes of this round (lanes holding the same digit
    // are found with one ballot per digit bit).  One wavefront's LDS accesses take effect in
    // program order, so the running counts need no barrier.
    const int64_t base = (int64_t)blockIdx.x * kRadixTile + (int64_t)w * kRounds * 64;
    const uint64_t lower_lanes = (1ull << lane) - 1ull;
    uint32_t key[kRounds], val[kRounds];
    int32_t rank[kRounds];
#pragma unroll
    for (int s = 0; s < kRounds; ++s) {
        const int64_t i = base + (int64_t)s * 64 + lane;
        const bool valid = i < n;
        key[s] = valid ? keys_in[i] : 0u;
        val[s] = valid ? vals_in[i] : 0u;
    }
#pragma unroll
    for (int s = 0; s < kRounds; ++s) {
        const bool valid = base + (int64_t)s * 64 + lane < n;
        // invalid slots get digit 256, a group of their own
        const int dg = valid ? (int)((key[s] >> shift) & (kRadixBins - 1)) : kRadixBins;
        uint64_t same = ~0ull;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const bool bit = (dg >> b) & 1;
            const uint64_t bal = __ballot(bit);
            same &= bit ? bal : ~bal;
        }
        const int r = __popcll(same & lower_lanes);
        const int prev = valid ? wave_cnt[w][dg] : 0;
        rank[s] = prev + r;
        if (valid && (same >> lane) == 1ull) wave_cnt[w][dg] = prev + r + 1;
    }
    __syncthreads();
    // first slot of each digit's pairs from each wavefront: the wavefronts in order after tile_start
    {
        int run = tile_start[t];
#pragma unroll
        for (int v = 0; v < kWaves; ++v) {
            const int c = wave_cnt[v][t];
            wave_cnt[v][t] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kRounds; ++s) {
        if (base + (int64_t)s * 64 + lane < n) {
            const int dg = (int)((key[s] >> shift) & (kRadixBins - 1));
            const int pos = wave_cnt[w][dg] + rank[s];
            tile_k[pos] = key[s];
            tile_v[pos] = val[s];
        }
    }
    __syncthreads();
    // write out in tile order: each digit's slots go to consecutive output positions
    const int n_tile = tile_start[kRadixBins];
    for (int x = t; x < n_tile; x += kRadixBlock) {
        const uint32_t k = tile_k[x];
        const int dg = (int)((k >> shift) & (kRadixBins - 1));
        const int64_t pos = (int64_t)blk_off[dg] + (x - tile_start[dg]);
        keys_out[pos] = k;
        vals_out[pos] = tile_v[x];
    }
}

__global__ __launch_bounds__(256) void copy_u32_kernel(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                                       int64_t n, uint32_t* __restrict__ oa, uint32_t* __restrict__ ob) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        oa[i] = a[i];
        ob[i] = b[i];
    }
}

inline size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

struct RadixLayout {
    int32_t n_blocks, n_hist, n_sums;
    size_t keys_tmp, vals_tmp, hist, offs, sums, total;
};

RadixLayout radix_layout(int64_t n) {
    RadixLayout l{};
    l.n_blocks = (int32_t)std::max<int64_t>(1, (n + kRadixTile - 1) / kRadixTile);
    l.n_hist = l.n_blocks * kRadixBins;
    l.n_sums = scan_sums_count(l.n_hist);
    size_t off = 0;
    l.keys_tmp = off; off += a256(4 * (size_t)n);
    l.vals_tmp = off; off += a256(4 * (size_t)n);
    l.hist = off; off += a256(4 * (size_t)l.n_hist);
    l.offs = off; off += a256(4 * (size_t)l.n_hist);
    l.sums = off; off += a256(4 * (size_t)l.n_sums);
    l.total = off;
    return l;
}

}  // namespace

size_t radix_sort_workspace(int64_t n) { return radix_layout(n).total; }

void radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      int64_t n, int key_bits, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    if (n < 0 || n > (int64_t(1) << 28)) throw std::runtime_error("radix sort: n must be in [0, 2^28]");
    if (key_bits < 0 || key_bits > 32) throw std::runtime_error("radix sort: key_bits must be in [0, 32]");
    if (n == 0) return;
    const RadixLayout l = radix_layout(n);
    if (!workspace || workspace_bytes < l.total) throw std::runtime_error("radix sort: workspace too small");
    unsigned char* ws = static_cast<unsigned char*>(workspace);
    uint32_t* kt = reinterpret_cast<uint32_t*>(ws + l.keys_tmp);
    uint32_t* vt = reinterpret_cast<uint32_t*>(ws + l.vals_tmp);
    int32_t* hist = reinterpret_cast<int32_t*>(ws + l.hist);
    int32_t* offs = reinterpret_cast<int32_t*>(ws + l.offs);
    int32_t* sums = reinterpret_cast<int32_t*>(ws + l.sums);
    const int passes = (key_bits + 7) / 8;
    if (passes == 0) {
        hipLaunchKernelGGL(copy_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, keys_in, vals_in,
                           n, keys_out, vals_out);
        TK_HIP_CHECK(hipGetLastError());
        return;
    }
    // the last pass writes the output: alternate so that it does
    const uint32_t* ks = keys_in;
    const uint32_t* vs = vals_in;
    for (int p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* kd = to_out ? keys_out : kt;
        uint32_t* vd = to_out ? vals_out : vt;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(l.n_blocks), dim3(kRadixBlock), 0, stream, ks, n, 8 * p, l.n_blocks,
                           hist);
        TK_HIP_CHECK(hipGetLastError());
        exclusive_scan_i32(hist, l.n_hist, sums, offs, stream);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(l.n_blocks), dim3(kRadixBlock), 0, stream, ks, vs, n, 8 * p,
                           l.n_blocks, hist, offs, kd, vd);
        TK_HIP_CHECK(hipGetLastError());
        ks = kd;
        vs = vd;
    }
}

}  // namespace tk
