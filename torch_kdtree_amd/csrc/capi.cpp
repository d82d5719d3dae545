// extern "C" boundary (declared in include/torch_kdtree_amd.h).  Every entry point catches
// C++ exceptions and turns them into a nonzero return code plus a thread-local message.
#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>
#include <omp.h>

#include "../../include/torch_kdtree_amd.h"
#include "batch.hpp"
#include "knn_gpu.hpp"
#include "tree.hpp"

struct tk_tree {
    tk::Tree* t;
};

struct tk_batch {
    tk::BatchPlan* p;
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
    } catch (...) {
        g_last_error = "unknown error";
    }
    return 1;
}

const tk::Tree& tree_of(const tk_tree* h) {
    if (!h || !h->t) throw std::runtime_error("null tree handle");
    return *h->t;
}

void check_dtype(int dtype) {
    if (dtype != TK_F32 && dtype != TK_F64) throw std::runtime_error("unsupported dtype");
}

}  // namespace

extern "C" {

const char* tk_last_error(void) { return g_last_error.c_str(); }

const char* tk_version(void) { return "1.0"; }

int tk_gpu_available(void) { return 1; }

int tk_device_count(int* count) {
    return guarded([&] {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        *count = c;
    });
}

int tk_tree_build(const void* points, int64_t n_points, int dims, int dtype, int levels, int n_threads,
                  tk_tree** out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("null output handle");
        *out = nullptr;
        check_dtype(dtype);
        if (n_points > 0 && !points) throw std::runtime_error("null points");
        tk::Tree* t = tk::build_tree(points, n_points, dims, levels, dtype, n_threads);
        *out = new tk_tree{t};
    });
}

int tk_tree_build_many(const void* const* points, const int64_t* n_points, int n_trees, int dims, int dtype,
                       const int32_t* levels, int n_threads, tk_tree** out) {
    return guarded([&] {
        if (n_trees < 0 || (n_trees > 0 && (!points || !n_points || !levels || !out)))
            throw std::runtime_error("invalid arguments");
        check_dtype(dtype);
        for (int i = 0; i < n_trees; ++i) out[i] = nullptr;
        if (n_trees == 1) {
            out[0] = new tk_tree{tk::build_tree(points[0], n_points[0], dims, levels[0], dtype, n_threads)};
            return;
        }
        // largest first, one tree per thread (the build's own tasks stay on that thread)
        std::vector<int> order(n_trees);
        for (int i = 0; i < n_trees; ++i) order[i] = i;
        std::sort(order.begin(), order.end(), [&](int a, int b) { return n_points[a] > n_points[b]; });
        std::vector<std::string> errors(n_trees);
        const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
        for (int j = 0; j < n_trees; ++j) {
            const int i = order[j];
            try {
                out[i] = new tk_tree{tk::build_tree(points[i], n_points[i], dims, levels[i], dtype, 1)};
            } catch (const std::exception& e) {
                errors[i] = e.what();
            }
        }
        for (int i = 0; i < n_trees; ++i) {
            if (!errors[i].empty()) {
                for (int j = 0; j < n_trees; ++j) {
                    tk_tree_free(out[j]);
                    out[j] = nullptr;
                }
                throw std::runtime_error("tree " + std::to_string(i) + ": " + errors[i]);
            }
        }
    });
}

int tk_tree_build_gpu(const void* points, int64_t n_points, int dims, int dtype, int levels, int device, void* stream,
                      tk_tree** out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("invalid arguments");
        *out = nullptr;
        check_dtype(dtype);
        const int32_t lv = levels;
        std::vector<tk::Tree*> t = tk::build_trees_gpu(points, &n_points, &lv, 1, dims, dtype, device, stream);
        std::unique_ptr<tk::Tree> owned(t[0]);
        tk::upload_tree(*owned, device);
        *out = new tk_tree{owned.release()};
    });
}

int tk_tree_build_many_gpu(const void* points, const int64_t* n_points, int n_trees, int dims, int dtype,
                           const int32_t* levels, int device, void* stream, tk_tree** out) {
    return guarded([&] {
        if (n_trees < 1 || !n_points || !levels || !out) throw std::runtime_error("invalid arguments");
        check_dtype(dtype);
        for (int i = 0; i < n_trees; ++i) out[i] = nullptr;
        std::vector<tk::Tree*> t = tk::build_trees_gpu(points, n_points, levels, n_trees, dims, dtype, device, stream);
        for (int i = 0; i < n_trees; ++i) out[i] = new tk_tree{t[i]};
    });
}

int tk_tree_serialize(const tk_tree* tree, void* buf, size_t* bytes) {
    return guarded([&] {
        if (!tree || !tree->t || !bytes) throw std::runtime_error("invalid arguments");
        const std::vector<unsigned char> img = tk::serialize_tree(*tree->t);
        if (buf) {
            if (*bytes < img.size()) throw std::runtime_error("serialization buffer too small");
            std::memcpy(buf, img.data(), img.size());
        }
        *bytes = img.size();
    });
}

int tk_tree_deserialize(const void* buf, size_t bytes, tk_tree** out) {
    return guarded([&] {
        if (!out || (!buf && bytes)) throw std::runtime_error("invalid arguments");
        *out = nullptr;
        *out = new tk_tree{tk::deserialize_tree(static_cast<const unsigned char*>(buf), bytes)};
    });
}

int tk_tree_save(const tk_tree* tree, const char* path) {
    return guarded([&] {
        if (!tree || !tree->t || !path) throw std::runtime_error("invalid arguments");
        tk::save_tree(*tree->t, path);
    });
}

int tk_tree_load(const char* path, tk_tree** out) {
    return guarded([&] {
        if (!out || !path) throw std::runtime_error("invalid arguments");
        *out = nullptr;
        *out = new tk_tree{tk::load_tree(path)};
    });
}

int tk_tree_upload(tk_tree* tree, int device) {
    return guarded([&] {
        if (!tree || !tree->t) throw std::runtime_error("null tree handle");
        tk::upload_tree(*tree->t, device);
    });
}

void tk_tree_free(tk_tree* tree) {
    if (!tree) return;
    if (tree->t) {
        tk::free_device_tree(*tree->t);
        delete tree->t;
    }
    delete tree;
}

int tk_tree_info(const tk_tree* tree, int32_t out[6]) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        out[0] = t.dims;
        out[1] = t.dtype;
        out[2] = t.levels;
        out[3] = t.n_points;
        out[4] = t.dev ? t.dev->device : -1;
        out[5] = t.finite ? 1 : 0;
    });
}

int tk_tree_variant_info(int variant, int32_t out[3], int* count) {
    return guarded([&] {
        if (variant < 0 || variant >= tk::kNumTreeVariants) throw std::runtime_error("unknown launch variant");
        if (out) {
            out[0] = tk::kTreeVariants[variant].chunk;
            out[1] = tk::kTreeVariants[variant].waves;
            out[2] = tk::kTreeVariants[variant].block;
        }
        if (count) *count = tk::kNumTreeVariants;
    });
}

int tk_tree_set_variant(tk_tree* tree, int variant) {
    return guarded([&] {
        if (!tree || !tree->t) throw std::runtime_error("null tree handle");
        if (variant < 0 || variant >= tk::kNumTreeVariants) throw std::runtime_error("unknown launch variant");
        tree->t->variant = variant;
    });
}

int tk_tree_get_variant(const tk_tree* tree, int* variant) {
    return guarded([&] { *variant = tree_of(tree).variant; });
}

int tk_tree_set_lower_list(tk_tree* tree, int mode) {
    return guarded([&] {
        if (!tree || !tree->t) throw std::runtime_error("null tree handle");
        if (mode < 0 || mode > 2) throw std::runtime_error("lower-index list mode must be 0, 1 or 2");
        tree->t->lower_list = mode;
    });
}

int tk_tree_get_lower_list(const tk_tree* tree, int32_t out[2]) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        out[0] = t.lower_list;
        out[1] = t.tie_prone ? 1 : 0;
    });
}

int tk_tree_get_shuffled_inds(const tk_tree* tree, int32_t* out) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        std::memcpy(out, t.shuffled.data(), sizeof(int32_t) * (size_t)t.n_points);
    });
}

int tk_tree_get_structured_points(const tk_tree* tree, void* out) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        std::memcpy(out, t.points.data(), t.points.size());
    });
}

int tk_tree_get_nodes(const tk_tree* tree, int32_t* axis, void* split, int32_t* leaf_off) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        if (axis)
            for (int32_t i = 0; i < t.n_nodes; ++i) axis[i] = t.axis[i];
        if (split) std::memcpy(split, t.split.data(), t.split.size());
        if (leaf_off) std::memcpy(leaf_off, t.leaf_off.data(), sizeof(int32_t) * t.leaf_off.size());
    });
}

int tk_knn_query_cpu(const tk_tree* tree, const void* queries, int64_t n_queries, int k, void* dist, int32_t* idx,
                     int64_t* orig_idx, int n_threads) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        if (n_queries < 0) throw std::runtime_error("negative query count");
        tk::query_cpu(t, queries, n_queries, k, dist, idx, orig_idx, n_threads);
    });
}

int tk_knn_search_stats(const tk_tree* tree, const void* queries, int64_t n_queries, int k, int64_t out[2],
                        int n_threads) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        if (n_queries < 0 || !out) throw std::runtime_error("invalid arguments");
        const size_t kk = (size_t)std::max(k, 1);
        std::vector<unsigned char> dist((size_t)n_queries * kk * tk::dtype_size(t.dtype));
        std::vector<int32_t> idx((size_t)n_queries * kk);
        tk::query_cpu(t, queries, n_queries, k, dist.data(), idx.data(), nullptr, n_threads, out);
    });
}

int tk_knn_query_gpu(const tk_tree* tree, const void* queries, int64_t n_queries, int k, void* dist, int32_t* idx,
                     int64_t* orig_idx, int tie, int presort, void* workspace, size_t workspace_bytes, void* stream) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        tk::QueryOpts o;
        o.tie = tie;
        o.presort = presort;
        tk::query_gpu(t, queries, n_queries, k, dist, idx, orig_idx, o, static_cast<hipStream_t>(stream), workspace,
                      workspace_bytes);
    });
}

int tk_knn_query_gpu_noalloc(const tk_tree* tree, const void* queries, int64_t n_queries, int k, void* dist,
                             int32_t* idx, int64_t* orig_idx, int tie, int presort, void* stream) {
    static_assert(TK_NOALLOC_MAX_QUERIES == tk::kNoAllocMaxQueries, "header and library disagree on the limit");
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        tk::QueryOpts o;
        o.tie = tie;
        o.presort = presort;
        tk::query_gpu_noalloc(t, queries, n_queries, k, dist, idx, orig_idx, o, static_cast<hipStream_t>(stream));
    });
}

int tk_knn_query_gpu_workspace(const tk_tree* tree, int64_t n_queries, int k, int tie, int presort, size_t* bytes) {
    return guarded([&] {
        (void)k;
        const tk::Tree& t = tree_of(tree);
        tk::QueryOpts o;
        o.tie = tie;
        o.presort = presort;
        *bytes = tk::query_gpu_workspace(t, n_queries, o);
    });
}

int tk_knn_bruteforce_gpu(const void* ref, int32_t n_ref, const void* queries, int64_t n_queries, int dims, int dtype,
                          int k, void* dist, int32_t* idx, void* workspace, size_t workspace_bytes, void* stream) {
    return guarded([&] {
        check_dtype(dtype);
        tk::bruteforce_gpu(ref, n_ref, queries, n_queries, dims, dtype, k, dist, idx, static_cast<hipStream_t>(stream),
                           workspace, workspace_bytes);
    });
}

int tk_knn_bruteforce_gpu_workspace(int32_t n_ref, int64_t n_queries, int dims, int dtype, int k, size_t* bytes) {
    return guarded([&] {
        check_dtype(dtype);
        *bytes = tk::bruteforce_gpu_workspace(n_ref, n_queries, dims, dtype, k);
    });
}

int tk_knn_backward_gpu(const void* queries, const void* ref, int64_t n_queries, int32_t n_ref, int dims, int dtype,
                        int k, const int64_t* idx, const void* grad_dist, void* grad_queries, void* grad_ref,
                        void* workspace, size_t workspace_bytes, void* stream) {
    return guarded([&] {
        check_dtype(dtype);
        tk::knn_backward_gpu(queries, ref, n_queries, n_ref, dims, dtype, k, idx, grad_dist, grad_queries, grad_ref,
                             static_cast<hipStream_t>(stream), workspace, workspace_bytes);
    });
}

int tk_knn_backward_tree_gpu(const tk_tree* tree, const void* queries, const void* ref, int64_t n_queries, int k,
                             const int32_t* sidx, const void* grad_dist, void* grad_queries, void* grad_ref,
                             void* workspace, size_t workspace_bytes, void* stream) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        if (!t.dev) throw std::runtime_error("The tree has no GPU copy (it was built for the CPU)");
        tk::DeviceGuard g(t.dev->device);
        tk::knn_backward_structured_gpu(queries, ref, n_queries, t.n_points, t.dims, t.dtype, k, sidx, t.dev->shuffled,
                                        grad_dist, grad_queries, grad_ref, static_cast<hipStream_t>(stream), workspace,
                                        workspace_bytes);
    });
}

int tk_knn_backward_tree_gpu_workspace(const tk_tree* tree, int64_t n_queries, int k, size_t* bytes) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        *bytes = tk::knn_backward_structured_workspace(n_queries, t.n_points, k, t.dims, t.dtype);
    });
}

int tk_knn_backward_gpu_workspace(int64_t n_queries, int32_t n_ref, int dims, int dtype, int k, size_t* bytes) {
    return guarded([&] {
        check_dtype(dtype);
        *bytes = tk::knn_backward_workspace(n_queries, n_ref, k, dims, dtype);
    });
}

int tk_batch_create(const tk_tree* const* trees, int n_trees, const int64_t* n_queries, const int32_t* ks, int tie,
                    int device, tk_batch** out) {
    return guarded([&] {
        if (!out) throw std::runtime_error("null output handle");
        *out = nullptr;
        if (n_trees < 1 || !trees || !n_queries || !ks) throw std::runtime_error("invalid arguments");
        std::vector<const tk::Tree*> ts(n_trees);
        for (int i = 0; i < n_trees; ++i) {
            if (!trees[i] || !trees[i]->t) throw std::runtime_error("null tree handle");
            ts[i] = trees[i]->t;
        }
        *out = new tk_batch{tk::batch_create(ts.data(), n_trees, n_queries, ks, tie, device)};
    });
}

void tk_batch_free(tk_batch* batch) {
    if (!batch) return;
    tk::batch_free(batch->p);
    delete batch;
}

int tk_batch_sizes(const tk_batch* batch, int64_t out[4]) {
    return guarded([&] {
        if (!batch || !batch->p) throw std::runtime_error("null batch handle");
        out[0] = batch->p->total_q;
        out[1] = batch->p->total_out;
        out[2] = batch->p->total_ref;
        out[3] = batch->p->n_clouds;
    });
}

int tk_batch_offsets(const tk_batch* batch, int64_t* query_off, int64_t* out_off, int64_t* ref_off) {
    return guarded([&] {
        if (!batch || !batch->p) throw std::runtime_error("null batch handle");
        const tk::BatchPlan& p = *batch->p;
        for (int c = 0; c <= p.n_clouds; ++c) {
            const bool end = c == p.n_clouds;
            if (query_off) query_off[c] = end ? p.total_q : p.clouds[c].q_base;
            if (out_off) out_off[c] = end ? p.total_out : p.clouds[c].out_base;
            if (ref_off) ref_off[c] = end ? p.total_ref : p.clouds[c].rec_base;
        }
    });
}

int tk_batch_forward(tk_batch* batch, const void* queries, void* dist, int32_t* idx, int64_t* orig_idx, void* stream) {
    return guarded([&] {
        if (!batch || !batch->p) throw std::runtime_error("null batch handle");
        tk::batch_forward(*batch->p, queries, dist, idx, orig_idx, static_cast<hipStream_t>(stream));
    });
}

int tk_batch_backward(tk_batch* batch, const void* queries, const int32_t* idx, const void* grad_dist,
                      void* grad_queries, void* grad_ref, void* stream) {
    return guarded([&] {
        if (!batch || !batch->p) throw std::runtime_error("null batch handle");
        tk::batch_backward(*batch->p, queries, idx, grad_dist, grad_queries, grad_ref,
                           static_cast<hipStream_t>(stream));
    });
}

int tk_batch_capture(tk_batch* batch, const void* queries, void* dist, int32_t* idx, int64_t* orig_idx,
                     const void* grad_dist, void* grad_queries, void* grad_ref, void* stream) {
    return guarded([&] {
        if (!batch || !batch->p) throw std::runtime_error("null batch handle");
        tk::batch_capture(*batch->p, queries, dist, idx, orig_idx, grad_dist, grad_queries, grad_ref,
                          static_cast<hipStream_t>(stream));
    });
}

int tk_batch_set_persistent(tk_batch* batch, int on) {
    return guarded([&] {
        if (!batch || !batch->p) throw std::runtime_error("null batch handle");
        batch->p->persistent = on != 0;
    });
}

int tk_batch_launch(tk_batch* batch, void* stream) {
    return guarded([&] {
        if (!batch || !batch->p) throw std::runtime_error("null batch handle");
        tk::batch_launch(*batch->p, static_cast<hipStream_t>(stream));
    });
}

int tk_sort_pairs_gpu(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out,
                      int64_t n, int key_bits, void* workspace, size_t workspace_bytes, void* stream) {
    return guarded([&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        const size_t need = tk::radix_sort_workspace(n);
        std::unique_ptr<tk::StreamScratch> own;
        if (!workspace) {
            own.reset(new tk::StreamScratch(need, s));
            workspace = own->ptr;
            workspace_bytes = need;
        }
        tk::radix_sort_pairs(keys_in, vals_in, keys_out, vals_out, n, key_bits, workspace, workspace_bytes, s);
    });
}

int tk_sort_pairs_gpu_workspace(int64_t n, size_t* bytes) {
    return guarded([&] { *bytes = tk::radix_sort_workspace(n); });
}

int tk_copy_gpu(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        tk::ProfScope prof("copy", static_cast<hipStream_t>(stream));
        tk::stream_copy_gpu(dst, src, bytes, static_cast<hipStream_t>(stream));
    });
}

int tk_valu_probe_gpu(float* out, int64_t threads, int iters, int mode, void* stream, int64_t* wave_insts) {
    return guarded([&] {
        tk::ProfScope prof("valu_probe", static_cast<hipStream_t>(stream));
        const int64_t n = tk::valu_probe_gpu(out, threads, iters, mode, static_cast<hipStream_t>(stream));
        if (wave_insts) *wave_insts = n;
    });
}

int tk_profile_enable(int on) {
    return guarded([&] { tk::prof_enable(on != 0); });
}

int tk_profile_reset(void) {
    return guarded([&] { tk::prof_reset(); });
}

int tk_profile_read(const char* stage, double* total_ms, int64_t* calls) {
    return guarded([&] { tk::prof_read(stage, total_ms, calls); });
}

}  // extern "C"
