// Batched many-cloud mode: plan construction (arena layout and upload), forward, backward and
// HIP-graph capture.  Kernels live in batch_kernels.hpp.
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include "batch.hpp"
#include "gpu_common.hpp"
#include "knn_gpu.hpp"
#include "knn_tree.hpp"

namespace tk {
namespace {

// Cloud of each query row: the last cloud whose first row is <= the row (clouds with no queries
// share their first row with the next cloud and are never the last such cloud for a real row).
__global__ __launch_bounds__(256) void batch_iota_kernel(int32_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

__global__ __launch_bounds__(256) void batch_qcloud_kernel(const CloudDesc* __restrict__ clouds, int n_clouds,
                                                           int64_t total_q, int32_t* __restrict__ qcloud) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total_q) return;
    int lo = 0, hi = n_clouds - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (clouds[mid].q_base <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    qcloud[i] = lo;
}

int kbucket_of(int k) {
    int K = 1;
    while (K < k) K <<= 1;
    return K;
}

inline size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

template <typename P>
void dmalloc(P** p, size_t bytes) {
    TK_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(bytes, 256)));
}

template <typename P>
void dfree(P*& p) noexcept {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T, int D>
void stage_cloud(const Tree& t, const CloudDesc& cd, unsigned char* recs, T* split, int32_t* leaf_off,
                 int32_t* shuffled, int32_t* inv, unsigned char* boxes) {
    using R = RecLayout<T, D>;
    using B = BoxLayout<T, D>;
    const T* pts = t.points_ptr<T>();
    for (int32_t s = 0; s < t.n_points; ++s) {
        unsigned char* r = recs + (size_t)(cd.rec_base + s) * R::kBytes;
        std::memcpy(r, pts + (size_t)s * D, sizeof(T) * D);
        if constexpr (R::kHasIdx) std::memcpy(r + sizeof(T) * D, &t.shuffled[s], 4);
        shuffled[cd.rec_base + s] = t.shuffled[s];
        inv[cd.rec_base + t.shuffled[s]] = s;
    }
    std::memcpy(split + cd.node_base, t.split_ptr<T>(), sizeof(T) * (size_t)t.n_nodes);
    for (int32_t j = 0; j <= t.n_leaves; ++j) leaf_off[cd.leaf_base + j] = cd.rec_base + t.leaf_off[j];
    const T* box = t.aabb_ptr<T>();
    for (int32_t b = 0; b < t.n_nodes + t.n_leaves; ++b) {
        unsigned char* dst = boxes + (size_t)(cd.box_base + b) * B::kBytes;
        std::memcpy(dst, box + (size_t)b * 2 * D, sizeof(T) * D);
        std::memcpy(dst + sizeof(T) * B::kPad, box + (size_t)b * 2 * D + D, sizeof(T) * D);
    }
}

template <typename T>
void stage_cloud_t(const Tree& t, const CloudDesc& cd, unsigned char* recs, void* split, int32_t* leaf_off,
                   int32_t* shuffled, int32_t* inv, unsigned char* boxes) {
    T* sp = static_cast<T*>(split);
    switch (t.dims) {
        case 1: stage_cloud<T, 1>(t, cd, recs, sp, leaf_off, shuffled, inv, boxes); break;
        case 2: stage_cloud<T, 2>(t, cd, recs, sp, leaf_off, shuffled, inv, boxes); break;
        case 3: stage_cloud<T, 3>(t, cd, recs, sp, leaf_off, shuffled, inv, boxes); break;
        case 8: stage_cloud<T, 8>(t, cd, recs, sp, leaf_off, shuffled, inv, boxes); break;
        default: throw std::runtime_error("unsupported dims");
    }
}

int box_bytes_of(int dtype, int dims) { return box_bytes_for(dtype, dims); }

void enqueue_forward(BatchPlan& p, const void* q, void* dist, int32_t* idx, int64_t* orig, bool emit, const void* g,
                     void* dq, hipStream_t s) {
    if (p.total_q == 0) return;
    if (p.dtype == kF32) {
        batch_presort_f32(p, static_cast<const float*>(q), s);
        batch_traverse_f32(p, static_cast<const float*>(q), static_cast<float*>(dist), idx, orig, emit,
                           static_cast<const float*>(g), static_cast<float*>(dq), s);
    } else {
        batch_presort_f64(p, static_cast<const double*>(q), s);
        batch_traverse_f64(p, static_cast<const double*>(q), static_cast<double*>(dist), idx, orig, emit,
                           static_cast<const double*>(g), static_cast<double*>(dq), s);
    }
}

void enqueue_backward(BatchPlan& p, const void* q, const int32_t* idx, const void* g, void* dq, void* dr,
                      bool from_forward, hipStream_t s) {
    if (!from_forward && p.total_q > 0) {
        if (p.dtype == kF32)
            batch_presort_f32(p, static_cast<const float*>(q), s);
        else
            batch_presort_f64(p, static_cast<const double*>(q), s);
    }
    if (p.dtype == kF32)
        batch_backward_f32(p, static_cast<const float*>(q), idx, static_cast<const float*>(g), static_cast<float*>(dq),
                           static_cast<float*>(dr), from_forward, s);
    else
        batch_backward_f64(p, static_cast<const double*>(q), idx, static_cast<const double*>(g),
                           static_cast<double*>(dq), static_cast<double*>(dr), from_forward, s);
}

}  // namespace

void batch_init_qcloud(const BatchPlan& p, hipStream_t s) {
    if (p.total_q == 0) return;
    const dim3 grid((unsigned)((p.total_q + 255) / 256));
    hipLaunchKernelGGL(batch_qcloud_kernel, grid, dim3(256), 0, s, p.d_clouds, p.n_clouds, p.total_q, p.qcloud);
    TK_HIP_CHECK(hipGetLastError());
    // a valid permutation and its inverse before any forward has run
    hipLaunchKernelGGL(batch_iota_kernel, grid, dim3(256), 0, s, p.order, p.total_q);
    hipLaunchKernelGGL(batch_iota_kernel, grid, dim3(256), 0, s, reinterpret_cast<int32_t*>(p.rank), p.total_q);
    TK_HIP_CHECK(hipGetLastError());
}

BatchPlan* batch_create(const Tree* const* trees, int n_clouds, const int64_t* n_queries, const int32_t* ks, int tie,
                        int device) {
    if (n_clouds < 1) throw std::runtime_error("batch: need at least one tree");
    if (tie != kTieReference && tie != kTieLowerIndex) throw std::runtime_error("batch: unknown tie policy");
    const Tree& t0 = *trees[0];
    for (int c = 0; c < n_clouds; ++c) {
        const Tree& t = *trees[c];
        if (t.dims != t0.dims || t.dtype != t0.dtype) throw std::runtime_error("batch: trees differ in dims or dtype");
        if (n_queries[c] < 0) throw std::runtime_error("batch: negative query count");
        if (ks[c] < 1 || ks[c] > kBatchMaxK) throw std::runtime_error("batch: k must be in [1, 32]");
        if (tie == kTieReference && !t.finite)
            throw std::runtime_error("batch: cloud " + std::to_string(c) +
                                     " holds a non-finite coordinate or split value; the reference tie policy "
                                     "follows the reference's own steps for it, on its own tree (query it with "
                                     "build_kd_tree(...).query) or use the lower-index tie policy");
        if (ks[c] > t.n_points)
            throw std::runtime_error("You requested more nearest neighbors than there are in the KD-Tree");
    }
    if (device < 0) TK_HIP_CHECK(hipGetDevice(&device));
    DeviceGuard guard(device);
    std::unique_ptr<BatchPlan, void (*)(BatchPlan*)> p(new BatchPlan(), [](BatchPlan* x) { batch_free(x); });
    p->device = device;
    p->dtype = t0.dtype;
    for (int c = 0; c < n_clouds; ++c) p->all_finite = p->all_finite && trees[c]->finite;
    p->dims = t0.dims;
    p->tie = tie;
    p->n_clouds = n_clouds;
    p->clouds.resize(n_clouds);

    // arena layout
    int64_t rec = 0, node = 0, box = 0, leaf = 0, qb = 0, ob = 0;
    for (int c = 0; c < n_clouds; ++c) {
        const Tree& t = *trees[c];
        CloudDesc& cd = p->clouds[c];
        cd.rec_base = (int32_t)rec;
        cd.node_base = (int32_t)node;
        cd.box_base = (int32_t)box;
        cd.leaf_base = (int32_t)leaf;
        cd.levels = t.levels;
        cd.n_nodes = t.n_nodes;
        cd.n_points = t.n_points;
        cd.k = ks[c];
        cd.q_base = qb;
        cd.out_base = ob;
        cd.n_queries = (int32_t)n_queries[c];
        cd.kbucket = kbucket_of(ks[c]);
        rec += t.n_points;
        node += t.n_nodes;
        box += (int64_t)t.n_nodes + t.n_leaves;
        leaf += (int64_t)t.n_leaves + 1;
        qb += n_queries[c];
        ob += n_queries[c] * ks[c];
        if (n_queries[c] >= (int64_t(1) << 31)) throw std::runtime_error("batch: too many queries in one cloud");
    }
    if (rec + kRecPad >= (int64_t(1) << 31) || box >= (int64_t(1) << 31) || leaf >= (int64_t(1) << 31))
        throw std::runtime_error("batch: the trees hold too many points for 32-bit indexing");
    if (qb > (int64_t(1) << 27)) throw std::runtime_error("batch: at most 2^27 queries per batch");
    if (ob > (int64_t(1) << 28)) throw std::runtime_error("batch: at most 2^28 output neighbours per batch");
    p->total_ref = rec;
    p->total_q = qb;
    p->total_out = ob;

    // Morton grids (the canonical rule); cells, sorted positions and canonical slots ordered by
    // k bucket, then cloud
    for (CloudDesc& cd : p->clouds) cd.morton_bits = canonical_morton_bits(cd.n_queries, p->dims);
    int64_t nb = 0, pos = 0, slot = 0;
    for (int K = 1; K <= kBatchMaxK; K <<= 1) {
        BatchPlan::Range r{K, pos, 0};
        for (CloudDesc& cd : p->clouds) {
            if (cd.kbucket != K) continue;
            cd.bucket_base = (int32_t)nb;
            cd.sorted_q_base = pos + r.count;
            cd.sorted_out_base = slot;
            nb += int64_t(1) << (cd.morton_bits * p->dims);
            r.count += cd.n_queries;
            slot += (int64_t)cd.n_queries * cd.k;
        }
        if (r.count > 0) p->ranges.push_back(r);
        pos += r.count;
    }
    if (nb >= (int64_t(1) << 31)) throw std::runtime_error("batch: too many presort cells");
    int kb = 1;
    while (kb < 32 && (int64_t(1) << kb) < nb) ++kb;
    p->key_bits = kb;
    // pair keys: the record within its cloud when that saves a sort pass (8-bit digits)
    {
        int64_t max_n = 1;
        for (const CloudDesc& cd : p->clouds) max_n = std::max<int64_t>(max_n, cd.n_points);
        auto bits = [](int64_t n) {
            int b = 1;
            while (b < 32 && (int64_t(1) << b) < n) ++b;
            return b;
        };
        const int global_bits = bits(std::max<int64_t>(p->total_ref, 1));
        const int local_bits = bits(max_n);
        const int sort_bits = 8 * ((local_bits + 7) / 8);
        if (sort_bits < 32 && sort_bits / 8 < (global_bits + 7) / 8 &&
            (int64_t)n_clouds <= (int64_t(1) << (32 - sort_bits))) {
            p->pair_bits = sort_bits;
            p->pair_tag_shift = sort_bits;
        } else {
            p->pair_bits = global_bits;
            p->pair_tag_shift = 0;
        }
    }

    // host staging of the arena
    const size_t ts = dtype_size(p->dtype);
    const size_t rb = (size_t)rec_bytes_for(p->dtype, p->dims);
    const size_t bb = (size_t)box_bytes_of(p->dtype, p->dims);
    std::vector<unsigned char> h_recs((size_t)(rec + kRecPad) * rb, 0);
    std::vector<unsigned char> h_split((size_t)node * ts);
    std::vector<int32_t> h_leaf((size_t)leaf), h_shuf((size_t)rec), h_inv((size_t)rec), h_rcl((size_t)rec);
    std::vector<unsigned char> h_boxes((size_t)box * bb, 0);
    for (int c = 0; c < n_clouds; ++c) {
        const CloudDesc& cd = p->clouds[c];
        std::fill(h_rcl.begin() + cd.rec_base, h_rcl.begin() + cd.rec_base + cd.n_points, c);
        if (p->dtype == kF32)
            stage_cloud_t<float>(*trees[c], p->clouds[c], h_recs.data(), h_split.data(), h_leaf.data(), h_shuf.data(),
                                 h_inv.data(), h_boxes.data());
        else
            stage_cloud_t<double>(*trees[c], p->clouds[c], h_recs.data(), h_split.data(), h_leaf.data(),
                                  h_shuf.data(), h_inv.data(), h_boxes.data());
    }

    // device arena and scratch
    dmalloc(&p->recs, h_recs.size());
    dmalloc(&p->split, h_split.size());
    dmalloc(&p->leaf_off, 4 * h_leaf.size());
    dmalloc(&p->shuffled, 4 * h_shuf.size());
    dmalloc(&p->inv_shuffled, 4 * h_inv.size());
    dmalloc(&p->rec_cloud, 4 * h_rcl.size());
    dmalloc(&p->boxes, h_boxes.size());
    dmalloc(&p->d_clouds, sizeof(CloudDesc) * (size_t)n_clouds);
    TK_HIP_CHECK(hipMemcpy(p->recs, h_recs.data(), h_recs.size(), hipMemcpyHostToDevice));
    TK_HIP_CHECK(hipMemcpy(p->split, h_split.data(), h_split.size(), hipMemcpyHostToDevice));
    TK_HIP_CHECK(hipMemcpy(p->leaf_off, h_leaf.data(), 4 * h_leaf.size(), hipMemcpyHostToDevice));
    TK_HIP_CHECK(hipMemcpy(p->shuffled, h_shuf.data(), 4 * h_shuf.size(), hipMemcpyHostToDevice));
    TK_HIP_CHECK(hipMemcpy(p->inv_shuffled, h_inv.data(), 4 * h_inv.size(), hipMemcpyHostToDevice));
    TK_HIP_CHECK(hipMemcpy(p->rec_cloud, h_rcl.data(), 4 * h_rcl.size(), hipMemcpyHostToDevice));
    TK_HIP_CHECK(hipMemcpy(p->boxes, h_boxes.data(), h_boxes.size(), hipMemcpyHostToDevice));
    TK_HIP_CHECK(hipMemcpy(p->d_clouds, p->clouds.data(), sizeof(CloudDesc) * (size_t)n_clouds,
                           hipMemcpyHostToDevice));
    const size_t nq = (size_t)p->total_q, no = (size_t)p->total_out;
    dmalloc(&p->qcloud, 4 * nq);
    dmalloc(&p->key, 4 * nq);
    dmalloc(&p->row, 4 * nq);
    dmalloc(&p->key_s, 4 * nq);
    dmalloc(&p->order, 4 * nq);
    dmalloc(&p->rank, 4 * nq);
    dmalloc(&p->run_start, 4 * (size_t)p->total_ref);
    dmalloc(&p->run_end, 4 * (size_t)p->total_ref);
    dmalloc(reinterpret_cast<unsigned char**>(&p->q_sorted), ts * (size_t)p->dims * nq);
    dmalloc(&p->keys, 4 * no);
    dmalloc(&p->vals, 4 * no);
    dmalloc(&p->keys_s, 4 * no);
    dmalloc(&p->vals_s, 4 * no);
    p->radix_ws_bytes = std::max(radix_sort_workspace((int64_t)no), radix_sort_workspace((int64_t)nq));
    dmalloc(reinterpret_cast<unsigned char**>(&p->radix_ws), p->radix_ws_bytes);
    dmalloc(&p->work, 4 * 8);
    batch_init_qcloud(*p, nullptr);
    TK_HIP_CHECK(hipDeviceSynchronize());
    return p.release();
}

void batch_free(BatchPlan* p) noexcept {
    if (!p) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (p->device >= 0) (void)hipSetDevice(p->device);
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    dfree(p->split);
    dfree(p->recs);
    dfree(p->leaf_off);
    dfree(p->shuffled);
    dfree(p->inv_shuffled);
    dfree(p->rec_cloud);
    dfree(p->boxes);
    dfree(p->d_clouds);
    dfree(p->qcloud);
    dfree(p->key);
    dfree(p->row);
    dfree(p->key_s);
    dfree(p->order);
    dfree(p->rank);
    dfree(p->run_start);
    dfree(p->run_end);
    dfree(p->q_sorted);
    dfree(p->g_sorted);
    dfree(p->keys);
    dfree(p->vals);
    dfree(p->keys_s);
    dfree(p->vals_s);
    dfree(p->radix_ws);
    dfree(p->work);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete p;
}

void batch_forward(BatchPlan& p, const void* queries, void* dist, int32_t* idx, int64_t* orig, hipStream_t s) {
    DeviceGuard guard(p.device);
    ProfScope prof("batch_forward", s);
    enqueue_forward(p, queries, dist, idx, orig, false, nullptr, nullptr, s);
}

void batch_backward(BatchPlan& p, const void* queries, const int32_t* idx, const void* grad_dist, void* grad_q,
                    void* grad_ref, hipStream_t s) {
    DeviceGuard guard(p.device);
    ProfScope prof("batch_backward", s);
    enqueue_backward(p, queries, idx, grad_dist, grad_q, grad_ref, false, s);
}

void batch_capture(BatchPlan& p, const void* queries, void* dist, int32_t* idx, int64_t* orig, const void* grad_dist,
                   void* grad_q, void* grad_ref, hipStream_t s) {
    DeviceGuard guard(p.device);
    if (p.graph_exec) {
        TK_HIP_CHECK(hipGraphExecDestroy(p.graph_exec));
        p.graph_exec = nullptr;
    }
    if (p.graph) {
        TK_HIP_CHECK(hipGraphDestroy(p.graph));
        p.graph = nullptr;
    }
    // capture on a private stream (the legacy default stream cannot be captured), ordered after
    // the caller's stream
    hipStream_t cs = nullptr;
    TK_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t st;
        ~StreamGuard() { (void)hipStreamDestroy(st); }
    } sg{cs};
    TK_HIP_CHECK(hipStreamSynchronize(s));
    // a step with both gradients stores the upstream gradient by canonical slot in its forward
    // (allocated here, outside the capture)
    const bool both = grad_dist && grad_q && grad_ref && p.total_out > 0;
    if (both && !p.g_sorted)
        dmalloc(reinterpret_cast<unsigned char**>(&p.g_sorted), dtype_size(p.dtype) * (size_t)p.total_out);
    p.g_sorted_valid = both;
    TK_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
        // a captured training step: the upstream gradient is in place before each replay, so the
        // forward also writes dL/dq, and the backward only dL/dr
        const bool train = grad_dist && (grad_q || grad_ref);
        enqueue_forward(p, queries, dist, idx, orig, train && grad_ref, train ? grad_dist : nullptr,
                        train ? grad_q : nullptr, cs);
        if (train && grad_ref) enqueue_backward(p, queries, idx, grad_dist, nullptr, grad_ref, true, cs);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cs, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    TK_HIP_CHECK(hipStreamEndCapture(cs, &p.graph));
    TK_HIP_CHECK(hipGraphInstantiate(&p.graph_exec, p.graph, nullptr, nullptr, 0));
}

void batch_launch(BatchPlan& p, hipStream_t s) {
    if (!p.graph_exec) throw std::runtime_error("batch: nothing captured");
    DeviceGuard guard(p.device);
    ProfScope prof("batch_graph", s);
    TK_HIP_CHECK(hipGraphLaunch(p.graph_exec, s));
}

}  // namespace tk
