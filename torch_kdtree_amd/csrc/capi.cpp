// extern "C" boundary (declared in include/torch_kdtree_amd.h).  Every entry point catches
// C++ exceptions and turns them into a nonzero return code plus a thread-local message.
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include <hip/hip_runtime.h>

#include "../../include/torch_kdtree_amd.h"
#include "knn_gpu.hpp"
#include "tree.hpp"

struct tk_tree {
    tk::Tree* t;
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

int tk_tree_info(const tk_tree* tree, int32_t out[5]) {
    return guarded([&] {
        const tk::Tree& t = tree_of(tree);
        out[0] = t.dims;
        out[1] = t.dtype;
        out[2] = t.levels;
        out[3] = t.n_points;
        out[4] = t.dev ? t.dev->device : -1;
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

int tk_knn_backward_gpu_workspace(int64_t n_queries, int32_t n_ref, int k, size_t* bytes) {
    return guarded([&] { *bytes = tk::knn_backward_workspace(n_queries, n_ref, k); });
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
