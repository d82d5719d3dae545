// pybind11 module `torch_knn`: the reference's binding surface (src/interface.cpp:13-150)
// rebuilt over this library's C ABI (include/torch_kdtree_amd.h).
//
// Classes KDTree{CPU,GPU}{1,2,3,8}D (float64) and KDTree{CPU,GPU}{1,2,3,8}DF (float32), with
// the reference's methods:
//   __init__(points_ref, levels)            src/interface.cpp:25-43
//   get_shuffled_inds() -> int32[N]          src/interface.cpp:45-50
//   get_structured_points() -> T[N, d]       src/interface.cpp:52-58
//   query(points_query_ptr, nr_query_points, nr_nns_searches, dist_arr_ptr, knn_idx_ptr)
//                                            src/interface.cpp:79-86
// plus keyword-only extensions that default to the reference's behaviour: the constructor's
// `device`, and query's `orig_idx_ptr`, `tie`, `presort`, `stream`, `n_threads`.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/torch_kdtree_amd.h"

namespace py = pybind11;

namespace {

void check(int rc) {
    if (rc != 0) throw std::runtime_error(tk_last_error());
}

template <typename T>
constexpr int dtype_code() {
    return std::is_same<T, float>::value ? TK_F32 : TK_F64;
}

template <typename T, int D, bool kGpu>
struct KDTree {
    tk_tree* tree = nullptr;
    int levels = 0;

    KDTree(const py::array_t<T, py::array::c_style | py::array::forcecast>& points_ref, int levels_, int device) {
        if (points_ref.ndim() != 2 || points_ref.shape(1) != D)
            throw std::runtime_error("points_ref must have shape (N, " + std::to_string(D) + ")");
        const T* p = points_ref.data();
        const int64_t n = points_ref.shape(0);
        levels = levels_;
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_tree_build(p, n, D, dtype_code<T>(), levels_, 0, &tree);
            if (rc == 0 && kGpu) rc = tk_tree_upload(tree, device);
        }
        if (rc != 0) {
            const std::string msg = tk_last_error();
            tk_tree_free(tree);
            tree = nullptr;
            throw std::runtime_error(msg);
        }
    }
    // Adopts a built or loaded tree (uploading it for GPU classes unless it is on the GPU already).
    KDTree(tk_tree* loaded, int device, bool uploaded = false) : tree(loaded) {
        int32_t info[6];
        check(tk_tree_info(tree, info));
        levels = info[2];
        if (kGpu && !uploaded) {
            int rc;
            {
                py::gil_scoped_release release;
                rc = tk_tree_upload(tree, device);
            }
            if (rc != 0) {
                const std::string msg = tk_last_error();
                tk_tree_free(tree);
                tree = nullptr;
                throw std::runtime_error(msg);
            }
        }
    }
    ~KDTree() { tk_tree_free(tree); }

    void save(const std::string& path) const {
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_tree_save(tree, path.c_str());
        }
        check(rc);
    }

    // Every coordinate and split value finite (otherwise GPU queries under the reference tie
    // policy run the reference's own steps, without point-box pruning).
    void set_variant(int v) { check(tk_tree_set_variant(tree, v)); }
    void set_lower_list(int mode) { check(tk_tree_set_lower_list(tree, mode)); }
    py::tuple get_lower_list() const {
        int32_t out[2] = {0, 0};
        check(tk_tree_get_lower_list(tree, out));
        return py::make_tuple(out[0], out[1] != 0);
    }
    int get_variant() const {
        int v = 0;
        check(tk_tree_get_variant(tree, &v));
        return v;
    }

    // Leaves and points scanned by the search of host queries (summed), see tk_knn_search_stats.
    py::tuple search_stats(const py::array_t<T, py::array::c_style | py::array::forcecast>& q, int k, int n_threads) {
        int64_t out[2] = {0, 0};
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_knn_search_stats(tree, q.data(), q.shape(0), k, out, n_threads);
        }
        check(rc);
        return py::make_tuple(out[0], out[1]);
    }

    // The backward for this GPU tree's query, neighbours by structured index (tk_knn_backward_tree_gpu).
    void backward(size_t q_ptr, size_t ref_ptr, int64_t n_q, int k, size_t sidx_ptr, size_t grad_ptr,
                  size_t grad_q_ptr, size_t grad_ref_ptr, size_t stream) const {
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_knn_backward_tree_gpu(tree, reinterpret_cast<const void*>(q_ptr),
                                          reinterpret_cast<const void*>(ref_ptr), n_q, k,
                                          reinterpret_cast<const int32_t*>(sidx_ptr),
                                          reinterpret_cast<const void*>(grad_ptr), reinterpret_cast<void*>(grad_q_ptr),
                                          reinterpret_cast<void*>(grad_ref_ptr), nullptr, 0,
                                          reinterpret_cast<void*>(stream));
        }
        check(rc);
    }

    bool is_finite() const {
        int32_t info[6];
        check(tk_tree_info(tree, info));
        return info[5] != 0;
    }

    int32_t n_points() const {
        int32_t info[6];
        check(tk_tree_info(tree, info));
        return info[3];
    }

    py::array_t<int32_t> get_shuffled_inds() const {
        py::array_t<int32_t> out(n_points());
        check(tk_tree_get_shuffled_inds(tree, out.mutable_data()));
        return out;
    }

    py::array_t<T> get_structured_points() const {
        py::array_t<T> out(std::vector<py::ssize_t>{n_points(), D});
        check(tk_tree_get_structured_points(tree, out.mutable_data()));
        return out;
    }

    py::tuple get_nodes() const {
        int32_t info[6];
        check(tk_tree_info(tree, info));
        const int L = info[2];
        const py::ssize_t nodes = (py::ssize_t(1) << L) - 1;
        py::array_t<int32_t> axis(nodes);
        py::array_t<T> split(nodes);
        py::array_t<int32_t> leaf_off((py::ssize_t(1) << L) + 1);
        check(tk_tree_get_nodes(tree, axis.mutable_data(), split.mutable_data(), leaf_off.mutable_data()));
        return py::make_tuple(axis, split, leaf_off);
    }

    void query(size_t points_query_ptr, size_t nr_query_points, int32_t nr_nns_searches, size_t dist_arr_ptr,
               size_t knn_idx_ptr, size_t orig_idx_ptr, int tie, int presort, size_t stream, int n_threads) {
        int rc;
        {
            py::gil_scoped_release release;
            if (kGpu)
                rc = tk_knn_query_gpu(tree, reinterpret_cast<const void*>(points_query_ptr), (int64_t)nr_query_points,
                                      nr_nns_searches, reinterpret_cast<void*>(dist_arr_ptr),
                                      reinterpret_cast<int32_t*>(knn_idx_ptr), reinterpret_cast<int64_t*>(orig_idx_ptr),
                                      tie, presort, nullptr, 0, reinterpret_cast<void*>(stream));
            else
                rc = tk_knn_query_cpu(tree, reinterpret_cast<const void*>(points_query_ptr), (int64_t)nr_query_points,
                                      nr_nns_searches, reinterpret_cast<void*>(dist_arr_ptr),
                                      reinterpret_cast<int32_t*>(knn_idx_ptr), reinterpret_cast<int64_t*>(orig_idx_ptr),
                                      n_threads);
        }
        check(rc);
    }
};

// Host trees of many clouds, built in parallel (for the batched mode).
struct TreeSet {
    std::vector<tk_tree*> trees;
    int dims = 0;
    int dtype = TK_F32;

    TreeSet() = default;
    TreeSet(const TreeSet&) = delete;
    TreeSet& operator=(const TreeSet&) = delete;
    TreeSet(const py::list& clouds, const std::vector<int32_t>& levels, int n_threads) {
        const size_t n = clouds.size();
        if (n == 0) throw std::runtime_error("no clouds");
        if (levels.size() != n) throw std::runtime_error("one levels value per cloud is required");
        std::vector<py::array> arrs;
        std::vector<const void*> ptrs(n);
        std::vector<int64_t> counts(n);
        for (size_t i = 0; i < n; ++i) {
            py::array a = py::array::ensure(clouds[i], py::array::c_style);
            if (!a) throw std::runtime_error("clouds must be arrays");
            const bool f32 = a.dtype().is(py::dtype::of<float>());
            const bool f64 = a.dtype().is(py::dtype::of<double>());
            if (!f32 && !f64) throw std::runtime_error("clouds must be float32 or float64");
            if (a.ndim() != 2) throw std::runtime_error("clouds must have shape (N, d)");
            if (i == 0) {
                dims = (int)a.shape(1);
                dtype = f32 ? TK_F32 : TK_F64;
            } else if ((int)a.shape(1) != dims || (f32 ? TK_F32 : TK_F64) != dtype) {
                throw std::runtime_error("all clouds must have the same dims and dtype");
            }
            ptrs[i] = a.data();
            counts[i] = a.shape(0);
            arrs.push_back(a);
        }
        trees.assign(n, nullptr);
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_tree_build_many(ptrs.data(), counts.data(), (int)n, dims, dtype, levels.data(), n_threads,
                                    trees.data());
        }
        check(rc);
    }
    ~TreeSet() {
        for (tk_tree* t : trees) tk_tree_free(t);
    }
    // Trees built in one GPU build from a concatenated device array (tk_tree_build_many_gpu).
    static std::shared_ptr<TreeSet> from_device(size_t points_ptr, const std::vector<int64_t>& n_points,
                                                const std::vector<int32_t>& levels, int dims_, int dtype_, int device,
                                                size_t stream) {
        const size_t n = n_points.size();
        if (n == 0) throw std::runtime_error("no clouds");
        if (levels.size() != n) throw std::runtime_error("one levels value per cloud is required");
        auto ts = std::make_shared<TreeSet>();
        ts->dims = dims_;
        ts->dtype = dtype_;
        ts->trees.assign(n, nullptr);
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_tree_build_many_gpu(reinterpret_cast<const void*>(points_ptr), n_points.data(), (int)n, dims_,
                                        dtype_, levels.data(), device, reinterpret_cast<void*>(stream),
                                        ts->trees.data());
        }
        check(rc);
        return ts;
    }
    std::vector<int> info(int field) const {
        std::vector<int> out;
        for (tk_tree* t : trees) {
            int32_t v[6];
            check(tk_tree_info(t, v));
            out.push_back(v[field]);
        }
        return out;
    }
};

// Batched query plan over a TreeSet (tk_batch_*).
struct BatchPlan {
    tk_batch* b = nullptr;
    int n_clouds = 0;

    BatchPlan(const TreeSet& ts, const std::vector<int64_t>& n_queries, const std::vector<int32_t>& ks, int tie,
              int device) {
        n_clouds = (int)ts.trees.size();
        if ((int)n_queries.size() != n_clouds || (int)ks.size() != n_clouds)
            throw std::runtime_error("one query count and one k per cloud are required");
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_batch_create(const_cast<const tk_tree* const*>(ts.trees.data()), n_clouds, n_queries.data(),
                                 ks.data(), tie, device, &b);
        }
        check(rc);
    }
    ~BatchPlan() { tk_batch_free(b); }
    py::tuple sizes() const {
        int64_t v[4];
        check(tk_batch_sizes(b, v));
        return py::make_tuple(v[0], v[1], v[2], v[3]);
    }
    py::tuple offsets() const {
        py::array_t<int64_t> q(n_clouds + 1), o(n_clouds + 1), r(n_clouds + 1);
        check(tk_batch_offsets(b, q.mutable_data(), o.mutable_data(), r.mutable_data()));
        return py::make_tuple(q, o, r);
    }
    void forward(size_t q, size_t dist, size_t idx, size_t orig, size_t stream) {
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_batch_forward(b, reinterpret_cast<const void*>(q), reinterpret_cast<void*>(dist),
                                  reinterpret_cast<int32_t*>(idx), reinterpret_cast<int64_t*>(orig),
                                  reinterpret_cast<void*>(stream));
        }
        check(rc);
    }
    void backward(size_t q, size_t idx, size_t g, size_t dq, size_t dr, size_t stream) {
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_batch_backward(b, reinterpret_cast<const void*>(q), reinterpret_cast<const int32_t*>(idx),
                                   reinterpret_cast<const void*>(g), reinterpret_cast<void*>(dq),
                                   reinterpret_cast<void*>(dr), reinterpret_cast<void*>(stream));
        }
        check(rc);
    }
    void capture(size_t q, size_t dist, size_t idx, size_t orig, size_t g, size_t dq, size_t dr, size_t stream) {
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_batch_capture(b, reinterpret_cast<const void*>(q), reinterpret_cast<void*>(dist),
                                  reinterpret_cast<int32_t*>(idx), reinterpret_cast<int64_t*>(orig),
                                  reinterpret_cast<const void*>(g), reinterpret_cast<void*>(dq),
                                  reinterpret_cast<void*>(dr), reinterpret_cast<void*>(stream));
        }
        check(rc);
    }
    void launch(size_t stream) {
        int rc;
        {
            py::gil_scoped_release release;
            rc = tk_batch_launch(b, reinterpret_cast<void*>(stream));
        }
        check(rc);
    }
    void set_persistent(bool on) { check(tk_batch_set_persistent(b, on ? 1 : 0)); }
};

template <typename T, int D, bool kGpu>
void bind_tree(py::module_& m, const char* name) {
    using C = KDTree<T, D, kGpu>;
    py::class_<C, std::shared_ptr<C>>(m, name)
        .def(py::init<const py::array_t<T, py::array::c_style | py::array::forcecast>&, int, int>(),
             py::arg("points_ref"), py::arg("levels"), py::kw_only(), py::arg("device") = -1)
        .def("save", &C::save, py::arg("path"), "Writes the tree to a file (load it with torch_knn.load_tree)")
        .def("get_shuffled_inds", &C::get_shuffled_inds,
             "Returns the shuffled indices to translate from local to global indices")
        .def("get_structured_points", &C::get_structured_points,
             "Returns the ordered points how they are used in the KD-Tree")
        .def("get_nodes", &C::get_nodes, "Returns (axis[int32], split[T], leaf_offsets[int32]) of the tree")
        .def("is_finite", &C::is_finite, "True when every coordinate and split value of the tree is finite")
        .def("backward", &C::backward, py::arg("queries_ptr"), py::arg("ref_ptr"), py::arg("n_queries"), py::arg("k"),
             py::arg("sidx_ptr"), py::arg("grad_dist_ptr"), py::arg("grad_queries_ptr"), py::arg("grad_ref_ptr"),
             py::arg("stream") = 0,
             "Backward of this GPU tree's squared distances, neighbours by structured index (tk_knn_backward_tree_gpu)")
        .def("set_variant", &C::set_variant, py::arg("variant"),
             "Launch variant of later GPU queries (see torch_knn.tree_variants())")
        .def("get_variant", &C::get_variant)
        .def("set_lower_list", &C::set_lower_list, py::arg("mode"),
             "List of later lower-index GPU queries: 0 = chosen from the cloud, 1 = tie-flagged, 2 = (distance, index)")
        .def("get_lower_list", &C::get_lower_list, "(mode, whether mode 0 takes the (distance, index) list)")
        .def("search_stats", &C::search_stats, py::arg("queries"), py::arg("k"), py::arg("n_threads") = 0,
             "(leaves, points) scanned by the search of these host queries, summed")
        .def("query", &C::query, py::arg("points_query_ptr"), py::arg("nr_query_points"), py::arg("nr_nns_searches"),
             py::arg("dist_arr_ptr"), py::arg("knn_idx_ptr"), py::kw_only(), py::arg("orig_idx_ptr") = 0,
             py::arg("tie") = 0, py::arg("presort") = -1, py::arg("stream") = 0, py::arg("n_threads") = 0,
             "Queries the KNN from the KD-Tree and puts the results in the array pointed to by dist_arr_ptr and "
             "knn_idx_ptr")
        .def_readonly("levels", &C::levels);
}

int dtype_from_str(const std::string& s) {
    if (s == "float32") return TK_F32;
    if (s == "float64") return TK_F64;
    throw std::runtime_error("unsupported dtype " + s);
}

}  // namespace

PYBIND11_MODULE(torch_knn, m) {
    m.doc() = "KD-tree KNN for AMD Instinct MI355X (gfx950): pybind11 layer over the torch_kdtree_amd C ABI";

    bind_tree<float, 3, false>(m, "KDTreeCPU3DF");
    bind_tree<double, 3, false>(m, "KDTreeCPU3D");
    bind_tree<float, 3, true>(m, "KDTreeGPU3DF");
    bind_tree<double, 3, true>(m, "KDTreeGPU3D");
    bind_tree<float, 2, false>(m, "KDTreeCPU2DF");
    bind_tree<double, 2, false>(m, "KDTreeCPU2D");
    bind_tree<float, 2, true>(m, "KDTreeGPU2DF");
    bind_tree<double, 2, true>(m, "KDTreeGPU2D");
    bind_tree<float, 1, false>(m, "KDTreeCPU1DF");
    bind_tree<double, 1, false>(m, "KDTreeCPU1D");
    bind_tree<float, 1, true>(m, "KDTreeGPU1DF");
    bind_tree<double, 1, true>(m, "KDTreeGPU1D");
    bind_tree<float, 8, false>(m, "KDTreeCPU8DF");
    bind_tree<double, 8, false>(m, "KDTreeCPU8D");
    bind_tree<float, 8, true>(m, "KDTreeGPU8DF");
    bind_tree<double, 8, true>(m, "KDTreeGPU8D");

    py::class_<TreeSet, std::shared_ptr<TreeSet>>(m, "TreeSet", "Host trees of many clouds, built in parallel")
        .def(py::init<const py::list&, const std::vector<int32_t>&, int>(), py::arg("clouds"), py::arg("levels"),
             py::arg("n_threads") = 0)
        .def_property_readonly("levels", [](const TreeSet& t) { return t.info(2); })
        .def_property_readonly("n_points", [](const TreeSet& t) { return t.info(3); })
        .def_property_readonly("dims", [](const TreeSet& t) { return t.dims; })
        .def_static("from_device", &TreeSet::from_device, py::arg("points_ptr"), py::arg("n_points"),
                    py::arg("levels"), py::arg("dims"), py::arg("dtype"), py::arg("device") = -1,
                    py::arg("stream") = 0, "Builds the trees of a concatenated device array on the GPU")
        .def("__len__", [](const TreeSet& t) { return t.trees.size(); });
    py::class_<BatchPlan, std::shared_ptr<BatchPlan>>(m, "BatchPlan", "Batched KNN plan over a TreeSet")
        .def(py::init<const TreeSet&, const std::vector<int64_t>&, const std::vector<int32_t>&, int, int>(),
             py::arg("trees"), py::arg("n_queries"), py::arg("ks"), py::arg("tie") = 0, py::arg("device") = -1)
        .def("sizes", &BatchPlan::sizes, "(total queries, total outputs, total reference points, clouds)")
        .def("offsets", &BatchPlan::offsets, "(query row offsets, output offsets, reference row offsets)")
        .def("forward", &BatchPlan::forward, py::arg("queries_ptr"), py::arg("dist_ptr"), py::arg("idx_ptr"),
             py::arg("orig_idx_ptr") = 0, py::arg("stream") = 0)
        .def("backward", &BatchPlan::backward, py::arg("queries_ptr"), py::arg("idx_ptr"), py::arg("grad_dist_ptr"),
             py::arg("grad_queries_ptr") = 0, py::arg("grad_ref_ptr") = 0, py::arg("stream") = 0)
        .def("capture", &BatchPlan::capture, py::arg("queries_ptr"), py::arg("dist_ptr"), py::arg("idx_ptr"),
             py::arg("orig_idx_ptr") = 0, py::arg("grad_dist_ptr") = 0, py::arg("grad_queries_ptr") = 0,
             py::arg("grad_ref_ptr") = 0, py::arg("stream") = 0)
        .def("launch", &BatchPlan::launch, py::arg("stream") = 0)
        .def("set_persistent", &BatchPlan::set_persistent, py::arg("on"),
             "One persistent traversal launch for every k bucket (work queues) instead of one launch per bucket");

    m.def(
        "load_tree",
        [](const std::string& path, bool gpu, int device) -> py::object {
            tk_tree* t = nullptr;
            int rc;
            {
                py::gil_scoped_release release;
                rc = tk_tree_load(path.c_str(), &t);
            }
            check(rc);
            int32_t info[6];
            if (tk_tree_info(t, info) != 0) {
                tk_tree_free(t);
                throw std::runtime_error(tk_last_error());
            }
            const int dims = info[0], dtype = info[1];
#define TK_MAKE(TT, DD)                                                                              \
    if (dims == DD && dtype == dtype_code<TT>()) {                                                   \
        if (gpu) return py::cast(std::make_shared<KDTree<TT, DD, true>>(t, device));                 \
        return py::cast(std::make_shared<KDTree<TT, DD, false>>(t, device));                         \
    }
            TK_MAKE(float, 1)
            TK_MAKE(float, 2)
            TK_MAKE(float, 3)
            TK_MAKE(float, 8)
            TK_MAKE(double, 1)
            TK_MAKE(double, 2)
            TK_MAKE(double, 3)
            TK_MAKE(double, 8)
#undef TK_MAKE
            tk_tree_free(t);
            throw std::runtime_error("tree file: unsupported dims or dtype");
        },
        py::arg("path"), py::arg("gpu") = false, py::arg("device") = -1,
        "Loads a tree written by save(); returns the KDTree{CPU,GPU}{d}D[F] object of its dims and dtype");

    m.def(
        "build_tree_gpu",
        [](size_t points_ptr, int64_t n_points, int dims, const std::string& dtype, int levels, int device,
           size_t stream) -> py::object {
            const int dt = dtype_from_str(dtype);
            tk_tree* t = nullptr;
            int rc;
            {
                py::gil_scoped_release release;
                rc = tk_tree_build_gpu(reinterpret_cast<const void*>(points_ptr), n_points, dims, dt, levels, device,
                                       reinterpret_cast<void*>(stream), &t);
            }
            check(rc);
#define TK_MAKE(TT, DD) \
    if (dims == DD && dt == dtype_code<TT>()) return py::cast(std::make_shared<KDTree<TT, DD, true>>(t, device, true));
            TK_MAKE(float, 1)
            TK_MAKE(float, 2)
            TK_MAKE(float, 3)
            TK_MAKE(float, 8)
            TK_MAKE(double, 1)
            TK_MAKE(double, 2)
            TK_MAKE(double, 3)
            TK_MAKE(double, 8)
#undef TK_MAKE
            tk_tree_free(t);
            throw std::runtime_error("unsupported dims or dtype");
        },
        py::arg("points_ptr"), py::arg("n_points"), py::arg("dims"), py::arg("dtype"), py::arg("levels"),
        py::arg("device") = -1, py::arg("stream") = 0,
        "Builds a tree on the GPU from a device array; returns the KDTreeGPU{d}D[F] object");

    m.def("check_for_gpu", [] { return tk_gpu_available() != 0; }, "Check if the library was compiled with GPU support");
    m.def("device_count", [] {
        int c = 0;
        check(tk_device_count(&c));
        return c;
    });

    m.def(
        "knn_bruteforce",
        [](size_t ref_ptr, int32_t n_ref, size_t q_ptr, int64_t n_q, int dims, const std::string& dtype, int k,
           size_t dist_ptr, size_t idx_ptr, size_t stream) {
            int rc;
            {
                py::gil_scoped_release release;
                rc = tk_knn_bruteforce_gpu(reinterpret_cast<const void*>(ref_ptr), n_ref,
                                           reinterpret_cast<const void*>(q_ptr), n_q, dims, dtype_from_str(dtype), k,
                                           reinterpret_cast<void*>(dist_ptr), reinterpret_cast<int32_t*>(idx_ptr),
                                           nullptr, 0, reinterpret_cast<void*>(stream));
            }
            check(rc);
        },
        py::arg("ref_ptr"), py::arg("n_ref"), py::arg("queries_ptr"), py::arg("n_queries"), py::arg("dims"),
        py::arg("dtype"), py::arg("k"), py::arg("dist_ptr"), py::arg("idx_ptr"), py::arg("stream") = 0);

    m.def(
        "knn_backward",
        [](size_t q_ptr, size_t ref_ptr, int64_t n_q, int32_t n_ref, int dims, const std::string& dtype, int k,
           size_t idx_ptr, size_t grad_ptr, size_t grad_q_ptr, size_t grad_ref_ptr, size_t stream) {
            int rc;
            {
                py::gil_scoped_release release;
                rc = tk_knn_backward_gpu(reinterpret_cast<const void*>(q_ptr), reinterpret_cast<const void*>(ref_ptr),
                                         n_q, n_ref, dims, dtype_from_str(dtype), k,
                                         reinterpret_cast<const int64_t*>(idx_ptr),
                                         reinterpret_cast<const void*>(grad_ptr), reinterpret_cast<void*>(grad_q_ptr),
                                         reinterpret_cast<void*>(grad_ref_ptr), nullptr, 0,
                                         reinterpret_cast<void*>(stream));
            }
            check(rc);
        },
        py::arg("queries_ptr"), py::arg("ref_ptr"), py::arg("n_queries"), py::arg("n_ref"), py::arg("dims"),
        py::arg("dtype"), py::arg("k"), py::arg("idx_ptr"), py::arg("grad_dist_ptr"), py::arg("grad_queries_ptr"),
        py::arg("grad_ref_ptr"), py::arg("stream") = 0);

    m.def(
        "sort_pairs",
        [](size_t keys_in, size_t vals_in, size_t keys_out, size_t vals_out, int64_t n, int key_bits, size_t stream) {
            int rc;
            {
                py::gil_scoped_release release;
                rc = tk_sort_pairs_gpu(reinterpret_cast<const uint32_t*>(keys_in),
                                       reinterpret_cast<const uint32_t*>(vals_in), reinterpret_cast<uint32_t*>(keys_out),
                                       reinterpret_cast<uint32_t*>(vals_out), n, key_bits, nullptr, 0,
                                       reinterpret_cast<void*>(stream));
            }
            check(rc);
        },
        py::arg("keys_in_ptr"), py::arg("vals_in_ptr"), py::arg("keys_out_ptr"), py::arg("vals_out_ptr"), py::arg("n"),
        py::arg("key_bits"), py::arg("stream") = 0,
        "Stable GPU radix sort of uint32 (key, value) pairs by the low key_bits bits of the keys");
    m.def("profile_enable", [](bool on) { check(tk_profile_enable(on ? 1 : 0)); }, py::arg("on"));
    m.def("profile_reset", [] { check(tk_profile_reset()); });
    m.def(
        "tree_variants",
        [] {
            int n = 0;
            int32_t v[3];
            check(tk_tree_variant_info(0, v, &n));
            py::list out;
            for (int i = 0; i < n; ++i) {
                check(tk_tree_variant_info(i, v, nullptr));
                py::dict d;
                d["chunk"] = v[0];
                d["waves"] = v[1];
                d["block"] = v[2];
                out.append(d);
            }
            return out;
        },
        "Launch variants of the GPU traversal: [{chunk, waves, block}] (index 0 = the built-in choice)");
    m.def(
        "copy_gpu",
        [](size_t dst_ptr, size_t src_ptr, size_t bytes, size_t stream) {
            int rc;
            {
                py::gil_scoped_release release;
                rc = tk_copy_gpu(reinterpret_cast<void*>(dst_ptr), reinterpret_cast<const void*>(src_ptr), bytes,
                                 reinterpret_cast<void*>(stream));
            }
            check(rc);
        },
        py::arg("dst_ptr"), py::arg("src_ptr"), py::arg("bytes"), py::arg("stream") = 0,
        "Streaming device copy (16-byte loads and stores), timed as stage 'copy'");
    m.def(
        "valu_probe",
        [](size_t out_ptr, int64_t threads, int iters, int mode, size_t stream) {
            int rc;
            int64_t n = 0;
            {
                py::gil_scoped_release release;
                rc = tk_valu_probe_gpu(reinterpret_cast<float*>(out_ptr), threads, iters, mode,
                                       reinterpret_cast<void*>(stream), &n);
            }
            check(rc);
            return n;
        },
        py::arg("out_ptr"), py::arg("threads"), py::arg("iters"), py::arg("mode") = 0, py::arg("stream") = 0,
        "VALU issue probe (16 independent 32-bit VALU instructions per iteration and lane), timed as stage "
        "'valu_probe'; returns the VALU wave instructions of its loop");
    m.def(
        "profile_read",
        [](const std::string& stage) {
            double ms = 0;
            int64_t calls = 0;
            check(tk_profile_read(stage.c_str(), &ms, &calls));
            return py::make_tuple(ms, calls);
        },
        py::arg("stage"), "Summed kernel milliseconds and call count of a stage (waits for its events)");

    m.attr("__version__") = tk_version();
}
