// GPU tree upload, the leaf-bucket presort (scan + scatter) and the query driver.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "knn_gpu.hpp"
#include "knn_tree.hpp"

namespace tk {

// ------------------------------------------------------------------------------------------
// Small helpers
// ------------------------------------------------------------------------------------------
StreamScratch::StreamScratch(size_t bytes, hipStream_t s) : stream(s) {
    if (bytes) TK_HIP_CHECK(hipMallocAsync(&ptr, bytes, s));
}
StreamScratch::~StreamScratch() {
    if (ptr) (void)hipFreeAsync(ptr, stream);
}

DeviceGuard::DeviceGuard(int device) {
    TK_HIP_CHECK(hipGetDevice(&prev));
    if (device >= 0 && device != prev) TK_HIP_CHECK(hipSetDevice(device));
}
DeviceGuard::~DeviceGuard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
}

namespace {
struct ProfEntry {
    std::string stage;
    hipEvent_t start, stop;
};
std::mutex g_prof_mu;
std::vector<ProfEntry> g_prof;
std::atomic<bool> g_prof_on{false};
}  // namespace

void prof_enable(bool on) { g_prof_on = on; }

void prof_reset() {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto& e : g_prof) {
        (void)hipEventDestroy(e.start);
        (void)hipEventDestroy(e.stop);
    }
    g_prof.clear();
}

bool prof_read(const char* stage, double* total_ms, int64_t* launches) {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    double ms = 0;
    int64_t n = 0;
    for (auto& e : g_prof) {
        if (e.stage != stage) continue;
        TK_HIP_CHECK(hipEventSynchronize(e.stop));
        float m = 0;
        TK_HIP_CHECK(hipEventElapsedTime(&m, e.start, e.stop));
        ms += m;
        ++n;
    }
    *total_ms = ms;
    *launches = n;
    return n > 0;
}

ProfScope::ProfScope(const char* st, hipStream_t s) {
    if (!g_prof_on) return;
    stage = st;
    stream = s;
    TK_HIP_CHECK(hipEventCreate(&start));
    TK_HIP_CHECK(hipEventRecord(start, s));
}

ProfScope::~ProfScope() {
    if (!start) return;
    hipEvent_t stop = nullptr;
    if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, stream) != hipSuccess) {
        (void)hipEventDestroy(start);
        if (stop) (void)hipEventDestroy(stop);
        return;
    }
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.push_back(ProfEntry{stage, start, stop});
}

int rec_bytes_for(int dtype, int dims) {
#define TK_RB(TT)                                  \
    switch (dims) {                                \
        case 1: return RecLayout<TT, 1>::kBytes;   \
        case 2: return RecLayout<TT, 2>::kBytes;   \
        case 3: return RecLayout<TT, 3>::kBytes;   \
        case 8: return RecLayout<TT, 8>::kBytes;   \
    }
    if (dtype == kF32) {
        TK_RB(float)
    } else {
        TK_RB(double)
    }
#undef TK_RB
    throw std::runtime_error("unsupported dims");
}

int box_bytes_for(int dtype, int dims) {
#define TK_BB(TT)                                  \
    switch (dims) {                                \
        case 1: return BoxLayout<TT, 1>::kBytes;   \
        case 2: return BoxLayout<TT, 2>::kBytes;   \
        case 3: return BoxLayout<TT, 3>::kBytes;   \
        case 8: return BoxLayout<TT, 8>::kBytes;   \
    }
    if (dtype == kF32) {
        TK_BB(float)
    } else {
        TK_BB(double)
    }
#undef TK_BB
    throw std::runtime_error("unsupported dims");
}

void free_device_tree(Tree& t) noexcept {
    DeviceTree* d = t.dev;
    if (!d) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (d->device >= 0) (void)hipSetDevice(d->device);
    (void)hipFree(d->split);
    (void)hipFree(d->recs);
    (void)hipFree(d->leaf_off);
    (void)hipFree(d->shuffled);
    (void)hipFree(d->inv_shuffled);
    (void)hipFree(d->boxes);
    (void)hipFree(d->group_boxes);
    (void)hipFree(d->group_base);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete d;
    t.dev = nullptr;
}

namespace {

// Whether exact distance ties look common among the cloud's neighbours, from up to 256 leaves
// spread over the tree and at most 64 points of each: more than one point in 256 with an exact
// duplicate among its leaf's sampled points, or, on some axis, fewer distinct values than half
// the sampled points (lattices, rounded coordinates).  Finite trees only (NaN has no order).
template <typename TT>
bool cloud_ties_often(const Tree& t) {
    const int D = t.dims;
    const TT* pts = reinterpret_cast<const TT*>(t.points.data());
    const int32_t nl = t.n_leaves;
    const int32_t S = std::min<int32_t>(nl, 256);
    std::vector<std::vector<TT>> vals((size_t)D);
    std::vector<int32_t> idx;
    int64_t sampled = 0, dups = 0;
    for (int32_t i = 0; i < S; ++i) {
        const int32_t l = (int32_t)((int64_t)i * nl / S);
        const int32_t b0 = t.leaf_off[(size_t)l];
        const int32_t b1 = std::min(t.leaf_off[(size_t)l + 1], b0 + 64);
        idx.resize((size_t)(b1 - b0));
        for (int32_t s = b0; s < b1; ++s) idx[(size_t)(s - b0)] = s;
        auto less = [&](int32_t a, int32_t b) {
            for (int c = 0; c < D; ++c) {
                const TT x = pts[(size_t)a * D + c], y = pts[(size_t)b * D + c];
                if (x != y) return x < y;
            }
            return false;
        };
        std::sort(idx.begin(), idx.end(), less);
        for (size_t j = 0; j + 1 < idx.size(); ++j) dups += !less(idx[j], idx[j + 1]) ? 1 : 0;
        for (int32_t s = b0; s < b1; ++s)
            for (int c = 0; c < D; ++c) vals[(size_t)c].push_back(pts[(size_t)s * D + c]);
        sampled += b1 - b0;
    }
    if (sampled < 2) return false;
    if (dups * 256 > sampled) return true;
    for (auto& v : vals) {
        std::sort(v.begin(), v.end());
        const int64_t distinct = std::unique(v.begin(), v.end()) - v.begin();
        if (distinct * 2 < sampled) return true;
    }
    return false;
}

}  // namespace

void upload_tree(Tree& t, int device) {
    free_device_tree(t);
    t.tie_prone = !t.finite || (t.dtype == kF32 ? cloud_ties_often<float>(t) : cloud_ties_often<double>(t));
    if (device < 0) TK_HIP_CHECK(hipGetDevice(&device));
    DeviceGuard g(device);
    std::unique_ptr<DeviceTree> d(new DeviceTree());
    d->device = device;
    d->rec_bytes = rec_bytes_for(t.dtype, t.dims);
    const size_t ts = dtype_size(t.dtype);
    const size_t coord_bytes = ts * (size_t)t.dims;
    const size_t n = (size_t)t.n_points;
    // kRecPad zero records after the last point: chunked leaf scans read them and ignore them
    std::vector<unsigned char> recs((n + kRecPad) * (size_t)d->rec_bytes, 0);
    std::vector<int32_t> inv(n);
    for (size_t s = 0; s < n; ++s) inv[(size_t)t.shuffled[s]] = (int32_t)s;
    const bool rec_idx = (size_t)d->rec_bytes >= coord_bytes + 4;  // RecLayout::kHasIdx
    for (size_t s = 0; s < n; ++s) {
        unsigned char* r = &recs[s * (size_t)d->rec_bytes];
        std::memcpy(r, &t.points[s * coord_bytes], coord_bytes);
        if (rec_idx) std::memcpy(r + coord_bytes, &t.shuffled[s], 4);
    }
    // boxes: lo then hi, each padded from dims to box_bytes / (2 * ts) elements
    d->box_bytes = box_bytes_for(t.dtype, t.dims);
    const size_t n_boxes = (size_t)t.n_nodes + (size_t)t.n_leaves;
    const size_t half = (size_t)d->box_bytes / 2;
    std::vector<unsigned char> boxes(n_boxes * (size_t)d->box_bytes, 0);
    for (size_t b = 0; b < n_boxes; ++b) {
        const unsigned char* src = &t.aabb[b * 2 * coord_bytes];
        unsigned char* dst = &boxes[b * (size_t)d->box_bytes];
        std::memcpy(dst, src, coord_bytes);
        std::memcpy(dst + half, src + coord_bytes, coord_bytes);
    }
    // Group boxes: min / max of each run of kGroupRecs records of a leaf, in the layout of `boxes`.
    // Only for leaves of several groups on average, where skipping a group pays for its test
    // (plain workload, 4M points: at depth 13, 488 points a leaf, the traversal takes 14.0 instead of
    // 20.1 ms; at depth 17, 30 a leaf, no gain), and only for finite trees, which prune with
    // boxes (a NaN coordinate would be dropped by min / max).
    const bool use_groups = t.finite && (int64_t)t.n_points >= (int64_t)kGroupMinLeaf * t.n_leaves;
    std::vector<int32_t> gbase;
    std::vector<unsigned char> gboxes;
    auto group_box = [&](auto tag) {
        using TT = decltype(tag);
        const TT* pts = reinterpret_cast<const TT*>(t.points.data());
        const int D = t.dims;
        for (int32_t l = 0; l < t.n_leaves; ++l) {
            const int32_t b0 = t.leaf_off[(size_t)l], b1 = t.leaf_off[(size_t)l + 1];
            for (int32_t g = 0; g < gbase[(size_t)l + 1] - gbase[(size_t)l]; ++g) {
                TT lo[8], hi[8];
                for (int a = 0; a < D; ++a) {
                    lo[a] = std::numeric_limits<TT>::infinity();
                    hi[a] = -std::numeric_limits<TT>::infinity();
                }
                for (int32_t s = b0 + g * kGroupRecs; s < std::min(b1, b0 + (g + 1) * kGroupRecs); ++s)
                    for (int a = 0; a < D; ++a) {
                        lo[a] = std::min(lo[a], pts[(size_t)s * D + a]);
                        hi[a] = std::max(hi[a], pts[(size_t)s * D + a]);
                    }
                unsigned char* dst = &gboxes[((size_t)gbase[(size_t)l] + g) * (size_t)d->box_bytes];
                std::memcpy(dst, lo, sizeof(TT) * D);
                std::memcpy(dst + half, hi, sizeof(TT) * D);
            }
        }
    };
    if (use_groups) {
        gbase.assign((size_t)t.n_leaves + 1, 0);
        for (int32_t l = 0; l < t.n_leaves; ++l) {
            const int32_t nl = t.leaf_off[(size_t)l + 1] - t.leaf_off[(size_t)l];
            gbase[(size_t)l + 1] = gbase[(size_t)l] + (nl + kGroupRecs - 1) / kGroupRecs;
        }
        gboxes.assign(std::max<size_t>((size_t)gbase[(size_t)t.n_leaves], 1) * (size_t)d->box_bytes, 0);
        if (t.dtype == kF32)
            group_box(float());
        else
            group_box(double());
    }
    try {
        TK_HIP_CHECK(hipMalloc(&d->split, (size_t)t.n_nodes * ts));
        TK_HIP_CHECK(hipMalloc((void**)&d->recs, recs.size()));
        TK_HIP_CHECK(hipMalloc((void**)&d->leaf_off, sizeof(int32_t) * t.leaf_off.size()));
        TK_HIP_CHECK(hipMalloc((void**)&d->shuffled, sizeof(int32_t) * n));
        TK_HIP_CHECK(hipMalloc((void**)&d->inv_shuffled, sizeof(int32_t) * n));
        TK_HIP_CHECK(hipMalloc((void**)&d->boxes, boxes.size()));
        TK_HIP_CHECK(hipMemcpy(d->split, t.split.data(), (size_t)t.n_nodes * ts, hipMemcpyHostToDevice));
        TK_HIP_CHECK(hipMemcpy(d->recs, recs.data(), recs.size(), hipMemcpyHostToDevice));
        TK_HIP_CHECK(hipMemcpy(d->leaf_off, t.leaf_off.data(), sizeof(int32_t) * t.leaf_off.size(),
                               hipMemcpyHostToDevice));
        TK_HIP_CHECK(hipMemcpy(d->shuffled, t.shuffled.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        TK_HIP_CHECK(hipMemcpy(d->inv_shuffled, inv.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        TK_HIP_CHECK(hipMemcpy(d->boxes, boxes.data(), boxes.size(), hipMemcpyHostToDevice));
        if (use_groups) {
            TK_HIP_CHECK(hipMalloc((void**)&d->group_boxes, gboxes.size()));
            TK_HIP_CHECK(hipMalloc((void**)&d->group_base, sizeof(int32_t) * gbase.size()));
            TK_HIP_CHECK(hipMemcpy(d->group_boxes, gboxes.data(), gboxes.size(), hipMemcpyHostToDevice));
            TK_HIP_CHECK(hipMemcpy(d->group_base, gbase.data(), sizeof(int32_t) * gbase.size(), hipMemcpyHostToDevice));
        }
    } catch (...) {
        (void)hipFree(d->split);
        (void)hipFree(d->recs);
        (void)hipFree(d->leaf_off);
        (void)hipFree(d->shuffled);
        (void)hipFree(d->inv_shuffled);
        (void)hipFree(d->boxes);
        (void)hipFree(d->group_boxes);
        (void)hipFree(d->group_base);
        throw;
    }
    t.dev = d.release();
}

// ------------------------------------------------------------------------------------------
// Streaming copy: 16-byte loads and stores, each thread striding over the array.  Its rate is the
// achievable HBM bandwidth the traversal's roofline is quoted against, and a copy of known size
// scales the FETCH_SIZE / WRITE_SIZE counters.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                          int64_t n16) {
    constexpr int kU = 4;  // loads in flight per thread
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += kU * stride) {
        uint4 v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i + u * stride < n16) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i + u * stride < n16) dst[i + u * stride] = v[u];
    }
}

void stream_copy_gpu(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    if (bytes % 16 != 0 || (((uintptr_t)dst | (uintptr_t)src) & 15u) != 0)
        throw std::runtime_error("stream copy: 16-byte aligned pointers and a multiple of 16 bytes needed");
    const int64_t n16 = (int64_t)(bytes / 16);
    if (n16 == 0) return;
    const unsigned blocks = (unsigned)std::min<int64_t>((n16 + 255) / 256, 256 * 16);
    hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, stream, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n16);
    TK_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------------------------------
// VALU issue probe: 8 independent chains per lane of 32-bit VALU operations (no FMA under
// -ffp-contract=off).  Mode 0: a v_mul_f32 and a v_add_f32 per chain, mode 1: two v_med3_f32 per
// chain (16 VALU instructions per loop iteration); mode 2: a compare, a select and an add per chain
// (v_cmp_gt_f32, v_cndmask_b32, v_add_f32: 24 per iteration), the mix of the searches' list and
// pruning code.  Launched over every CU, mode 0's rate is the issue ceiling of 4-cycle
// instructions (one wave instruction per SIMD every 4 cycles); compares and selects issue faster,
// which mode 2 shows.  The traversal's VALU instruction count is quoted against both.
// ------------------------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256) void valu_probe_kernel(float* __restrict__ out, int iters, float a, float b) {
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = (float)threadIdx.x * 0.001f + (float)i;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (MODE == 0) {
                x[i] = x[i] * a;
                x[i] = x[i] + b;
            } else if constexpr (MODE == 1) {
                x[i] = __builtin_amdgcn_fmed3f(x[i], a, b);
                x[i] = __builtin_amdgcn_fmed3f(x[i], b, a);
            } else {
                x[i] = x[i] + (x[i] < a ? b : -b);
            }
        }
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += x[i];
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

int64_t valu_probe_gpu(float* out, int64_t threads, int iters, int mode, hipStream_t stream) {
    if (threads <= 0 || threads % 256 != 0 || threads / 256 > INT32_MAX || iters <= 0)
        throw std::runtime_error("VALU probe: threads must be a positive multiple of 256, iters positive");
    if (mode < 0 || mode > 2) throw std::runtime_error("VALU probe: mode 0 (mul + add), 1 (med3) or 2 (compare + select)");
    const dim3 grid((unsigned)(threads / 256));
    if (mode == 0)
        hipLaunchKernelGGL(valu_probe_kernel<0>, grid, dim3(256), 0, stream, out, iters, 0.999f, 0.001f);
    else if (mode == 1)
        hipLaunchKernelGGL(valu_probe_kernel<1>, grid, dim3(256), 0, stream, out, iters, 0.999f, 0.001f);
    else
        hipLaunchKernelGGL(valu_probe_kernel<2>, grid, dim3(256), 0, stream, out, iters, 0.999f, 0.001f);
    TK_HIP_CHECK(hipGetLastError());
    return threads / 64 * (int64_t)iters * (mode == 2 ? 24 : 16);
}

// ------------------------------------------------------------------------------------------
// Scan / scatter kernels for the presort
// ------------------------------------------------------------------------------------------
__global__ void fill_i32_kernel(int32_t* __restrict__ p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(v, off, 64);
        if (lane >= off) v += y;
    }
    return v;
}

// Exclusive scan of n int32 counts (n <= kScanChunk * kScanChunk) in three kernels.
constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;
constexpr int kScanChunk = kScanBlock * kScanItems;

// Exclusive scan over a block of kScanBlock threads; `total` receives the block sum.
__device__ __forceinline__ int block_exclusive_scan(int v, int* wsum, int& total) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    constexpr int kWaves = kScanBlock / 64;
    const int inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        const int ws = lane < kWaves ? wsum[lane] : 0;
        const int wi = wave_inclusive_scan(ws);
        if (lane < kWaves) wsum[lane] = wi - ws;
        if (lane == kWaves - 1) wsum[kWaves] = wi;
    }
    __syncthreads();
    total = wsum[kWaves];
    const int r = inc - v + wsum[wid];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanBlock) void scan_block_sums_kernel(const int32_t* __restrict__ in, int32_t n,
                                                                      int32_t* __restrict__ sums) {
    __shared__ int wsum[kScanBlock / 64 + 1];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    int v = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < n) v += in[base + j];
    int total;
    (void)block_exclusive_scan(v, wsum, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void scan_sums_kernel(int32_t* __restrict__ sums, int32_t nsums) {
    __shared__ int wsum[kScanBlock / 64 + 1];
    const int base = threadIdx.x * kScanItems;
    int x[kScanItems];
    int v = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        x[j] = base + j < nsums ? sums[base + j] : 0;
        v += x[j];
    }
    int total;
    int run = block_exclusive_scan(v, wsum, total);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < nsums) sums[base + j] = run;
        run += x[j];
    }
}

__global__ __launch_bounds__(kScanBlock) void scan_apply_kernel(const int32_t* __restrict__ in, int32_t n,
                                                                 const int32_t* __restrict__ sums,
                                                                 int32_t* __restrict__ out) {
    __shared__ int wsum[kScanBlock / 64 + 1];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanItems;
    int x[kScanItems];
    int v = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        x[j] = base + j < n ? in[base + j] : 0;
        v += x[j];
    }
    int total;
    int run = block_exclusive_scan(v, wsum, total) + sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) out[base + j] = run;
        run += x[j];
    }
}


// ------------------------------------------------------------------------------------------
// Query driver
// ------------------------------------------------------------------------------------------
int32_t scan_sums_count(int32_t n) { return (n + kScanChunk - 1) / kScanChunk; }

void exclusive_scan_i32(const int32_t* in, int32_t n, int32_t* sums, int32_t* out, hipStream_t stream) {
    if (n > kScanChunk * kScanChunk) throw std::runtime_error("exclusive scan: too many values");
    const int32_t nsums = scan_sums_count(n);
    if (nsums == 0) return;
    hipLaunchKernelGGL(scan_block_sums_kernel, dim3(nsums), dim3(kScanBlock), 0, stream, in, n, sums);
    TK_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanBlock), 0, stream, sums, nsums);
    TK_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nsums), dim3(kScanBlock), 0, stream, in, n, sums, out);
    TK_HIP_CHECK(hipGetLastError());
}

namespace {

constexpr int64_t kPresortMinQueries = 4096;
constexpr int kMaxBucketBits = 22;

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

int presort_mode(const Tree& t, int64_t nq, const QueryOpts& o) {
    if (nq <= 0) return kPresortOff;
    // auto: the space-filling curve at d <= 3; the near-leaf order at d = 8, where the curve over a
    // 3-bit-per-axis grid groups queries more coarsely (d8_mixture, 2M queries, k = 32, depth 16:
    // 222 ms per query step against 231 ms, and 246 ms without a presort)
    if (o.presort == kPresortAuto)
        return (nq >= kPresortMinQueries && t.levels > 1) ? (t.dims == 8 ? kPresortLeaf : kPresortMorton) : kPresortOff;
    if (o.presort != kPresortOff && o.presort != kPresortMorton && o.presort != kPresortLeaf)
        throw std::runtime_error("unknown presort mode");
    return o.presort;
}

int bucket_shift(const Tree& t) { return std::max(0, t.levels - kMaxBucketBits); }

struct PresortLayout {
    size_t key, row, key_s, order, bbox, radix, radix_bytes, total;
    int key_bits;
};

int bits_for(int64_t n) {
    int b = 1;
    while (b < 32 && (int64_t(1) << b) < n) ++b;
    return b;
}

PresortLayout presort_layout(const Tree& t, int64_t nq, int mode) {
    PresortLayout l{};
    l.key_bits = mode == kPresortMorton ? bits_for(morton_bucket_count(t.dims)) : bits_for(t.n_leaves >> bucket_shift(t));
    size_t off = 0;
    l.key = off; off += align256(sizeof(int32_t) * (size_t)nq);
    l.row = off; off += align256(sizeof(int32_t) * (size_t)nq);
    l.key_s = off; off += align256(sizeof(int32_t) * (size_t)nq);
    l.order = off; off += align256(sizeof(int32_t) * (size_t)nq);
    l.bbox = off; off += align256(2 * 8 * (size_t)t.dims);
    l.radix = off;
    l.radix_bytes = radix_sort_workspace(nq);
    off += align256(l.radix_bytes);
    l.total = off;
    return l;
}

void check_launch() { TK_HIP_CHECK(hipGetLastError()); }

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, int n, unsigned long long lo_v,
                                unsigned long long hi_v, int half) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i < half ? lo_v : hi_v;
}

}  // namespace

void init_bbox(void* bbox, int dtype, int dims, hipStream_t stream) {
    // [min codes (all ones) | max codes (zero)] in the width of T
    if (dtype == kF32) {
        hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, stream, static_cast<int32_t*>(bbox), (int64_t)dims,
                           (int32_t)-1);
        hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, stream, static_cast<int32_t*>(bbox) + dims,
                           (int64_t)dims, 0);
    } else {
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, stream, static_cast<unsigned long long*>(bbox),
                           2 * dims, ~0ull, 0ull, dims);
    }
    TK_HIP_CHECK(hipGetLastError());
}

int canonical_morton_bits(int64_t nq, int dims) {
    int maxb = 0;
    switch (dims) {
        case 1: maxb = MortonBits<1>::kBits; break;
        case 2: maxb = MortonBits<2>::kBits; break;
        case 3: maxb = MortonBits<3>::kBits; break;
        case 8: maxb = MortonBits<8>::kBits; break;
        default: throw std::runtime_error("unsupported dims");
    }
    int b = 0;
    while (b < maxb && (int64_t(8) << ((b + 1) * dims)) <= nq) ++b;
    return b;
}

// Queries are presorted and searched in slices of at most this many (the presort's radix sort
// orders at most 2^28 pairs); queries are independent, so slicing changes no result.
constexpr int64_t kQuerySlice = int64_t(1) << 28;

size_t query_gpu_workspace(const Tree& t, int64_t nq, const QueryOpts& o) {
    nq = std::min(nq, kQuerySlice);
    const int mode = presort_mode(t, nq, o);
    if (mode == kPresortOff) return 0;
    return presort_layout(t, nq, mode).total;
}

void query_gpu(const Tree& t, const void* queries, int64_t nq, int k, void* dist, int32_t* idx, int64_t* orig,
               const QueryOpts& o, hipStream_t stream, void* workspace, size_t workspace_bytes) {
    if (!t.dev) throw std::runtime_error("The tree has no GPU copy (it was built for the CPU)");
    if (k < 1) throw std::runtime_error("nr_nns_searches must be >= 1");
    if (k > t.n_points) throw std::runtime_error("You requested more nearest neighbors than there are in the KD-Tree");
    if (k > kMaxGpuK)
        throw std::runtime_error("nr_nns_searches > " + std::to_string(kMaxGpuK) + " is not supported on the GPU");
    if (o.tie != kTieReference && o.tie != kTieLowerIndex) throw std::runtime_error("unknown tie policy");
    if (nq < 0) throw std::runtime_error("negative query count");
    if (nq == 0) return;
    if (nq >= (int64_t(1) << 31)) throw std::runtime_error("Only 32 bit signed query indexing implemented");
    if (nq > kQuerySlice) {
        const size_t ts = dtype_size(t.dtype);
        for (int64_t s0 = 0; s0 < nq; s0 += kQuerySlice) {
            const int64_t n1 = std::min(kQuerySlice, nq - s0);
            query_gpu(t, static_cast<const unsigned char*>(queries) + (size_t)s0 * t.dims * ts, n1, k,
                      static_cast<unsigned char*>(dist) + (size_t)s0 * k * ts, idx + s0 * k,
                      orig ? orig + s0 * k : nullptr, o, stream, workspace, workspace_bytes);
        }
        return;
    }
    DeviceGuard g(t.dev->device);
    const int32_t* order = nullptr;
    std::unique_ptr<StreamScratch> own;
    const int mode = presort_mode(t, nq, o);
    if (mode != kPresortOff) {
        ProfScope ps("presort", stream);
        const PresortLayout l = presort_layout(t, nq, mode);
        unsigned char* ws = static_cast<unsigned char*>(workspace);
        if (!ws) {
            own.reset(new StreamScratch(l.total, stream));
            ws = static_cast<unsigned char*>(own->ptr);
        } else if (workspace_bytes < l.total) {
            throw std::runtime_error("query workspace too small");
        }
        uint32_t* key = reinterpret_cast<uint32_t*>(ws + l.key);
        uint32_t* row = reinterpret_cast<uint32_t*>(ws + l.row);
        uint32_t* key_s = reinterpret_cast<uint32_t*>(ws + l.key_s);
        int32_t* ord = reinterpret_cast<int32_t*>(ws + l.order);
        void* bbox = ws + l.bbox;
        if (mode == kPresortMorton) {
            init_bbox(bbox, t.dtype, t.dims, stream);
            if (t.dtype == kF32)
                launch_morton_bucket_f32(t, static_cast<const float*>(queries), nq, bbox, key, row, stream);
            else
                launch_morton_bucket_f64(t, static_cast<const double*>(queries), nq, bbox, key, row, stream);
        } else if (t.dtype == kF32) {
            launch_leaf_bucket_f32(t, static_cast<const float*>(queries), nq, bucket_shift(t), key, row, stream);
        } else {
            launch_leaf_bucket_f64(t, static_cast<const double*>(queries), nq, bucket_shift(t), key, row, stream);
        }
        radix_sort_pairs(key, row, key_s, reinterpret_cast<uint32_t*>(ord), nq, l.key_bits, ws + l.radix,
                         l.radix_bytes, stream);
        order = ord;
    }
    ProfScope pt("traverse", stream);
    if (t.dtype == kF32)
        launch_tree_search_f32(t, static_cast<const float*>(queries), nq, order, k, static_cast<float*>(dist), idx, orig,
                               o.tie, stream);
    else
        launch_tree_search_f64(t, static_cast<const double*>(queries), nq, order, k, static_cast<double*>(dist), idx,
                               orig, o.tie, stream);
}

// ------------------------------------------------------------------------------------------
// The query that allocates nothing: its presort workspace is a static device array.  `scratch`
// below names the host-side stand-in of that array; the device address of each GPU's copy comes
// from hipGetSymbolAddress, taken once per device on the first call there (with that device
// current), and only that address is ever handed to kernels.
// ------------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device.
void* scratch_on_current_device() {
    int dev = -1;
    TK_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) throw std::runtime_error("no-allocation query: device out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("no-allocation query: hipGetSymbolAddress failed: ") +
                                     hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

void query_gpu_noalloc(const Tree& t, const void* queries, int64_t nq, int k, void* dist, int32_t* idx, int64_t* orig,
                       const QueryOpts& o, hipStream_t stream) {
    if (!t.dev) throw std::runtime_error("The tree has no GPU copy (it was built for the CPU)");
    if (nq > kNoAllocMaxQueries)
        throw std::runtime_error("no-allocation query: " + std::to_string(nq) + " queries exceed the limit of " +
                                 std::to_string(kNoAllocMaxQueries) + " per call (static scratch of " +
                                 std::to_string(kNoAllocScratchBytes >> 20) + " MiB per GPU)");
    const size_t need = query_gpu_workspace(t, nq, o);
    if (need > kNoAllocScratchBytes) throw std::runtime_error("no-allocation query: workspace exceeds the static scratch");
    DeviceGuard g(t.dev->device);
    void* ws = need ? scratch_on_current_device() : nullptr;
    query_gpu(t, queries, nq, k, dist, idx, orig, o, stream, ws, need ? kNoAllocScratchBytes : 0);
}

}  // namespace tk
