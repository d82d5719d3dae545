// Host self-test of the product's CPU code (tree build, CPU search, serialization) built with
// AddressSanitizer and UndefinedBehaviorSanitizer and with the debug cross-checks (TK_DEBUG_CHECKS:
// the reference's iterative-vs-recursive search check, src/kdtree.cpp:182-189, and its leaf-offset
// checks, src/kdtree.hpp:517-529).  Every result is compared with the oracle (the CPU restatement of
// the reference) linked into the same binary.  Exit status 0 and a last line "host selftest OK" when
// every check passes; a sanitizer report or a mismatch ends it with a nonzero status.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../torch_kdtree_amd/csrc/tree.hpp"

extern "C" {
void* oracle_build(const void* points, int32_t n, int dims, int levels, int dtype);
void oracle_free(void* h);
int oracle_search(const void* h, const void* q, int64_t nq, int k, void* dist, int32_t* idx, int nthreads,
                  int recursive);
const char* oracle_last_error();
}

namespace {

int g_fail = 0;

template <typename T>
void check_case(const char* name, int dims, std::vector<T> pts, std::vector<T> qry, int levels, int k) {
    const int dtype = sizeof(T) == 4 ? tk::kF32 : tk::kF64;
    const int64_t n = (int64_t)pts.size() / dims, nq = (int64_t)qry.size() / dims;
    tk::Tree* t = tk::build_tree(pts.data(), n, dims, levels, dtype, 0);
    std::vector<T> d(nq * k), od(nq * k);
    std::vector<int32_t> i(nq * k), oi(nq * k);
    std::vector<int64_t> o(nq * k);
    tk::query_cpu(*t, qry.data(), nq, k, d.data(), i.data(), o.data(), 0);
    void* h = oracle_build(pts.data(), (int32_t)n, dims, levels, dtype);
    if (!h || oracle_search(h, qry.data(), nq, k, od.data(), oi.data(), 0, 0) != 0) {
        std::printf("%s: oracle failed: %s\n", name, oracle_last_error());
        ++g_fail;
        return;
    }
    oracle_free(h);
    const bool same = std::memcmp(d.data(), od.data(), d.size() * sizeof(T)) == 0 && i == oi;
    // serialization round trip, then the same search on the reloaded tree
    const std::vector<unsigned char> img = tk::serialize_tree(*t);
    tk::Tree* t2 = tk::deserialize_tree(img.data(), img.size());
    std::vector<T> d2(nq * k);
    std::vector<int32_t> i2(nq * k);
    tk::query_cpu(*t2, qry.data(), nq, k, d2.data(), i2.data(), nullptr, 0);
    const bool same2 = std::memcmp(d.data(), d2.data(), d.size() * sizeof(T)) == 0 && i == i2;
    std::printf("%-28s d=%d %s levels=%2d k=%3d finite=%d: %s\n", name, dims, sizeof(T) == 4 ? "f32" : "f64", levels,
                k, (int)t->finite, same && same2 ? "equal to the oracle" : "MISMATCH");
    if (!(same && same2)) ++g_fail;
    delete t;
    delete t2;
}

template <typename T>
std::vector<T> normal(std::mt19937_64& g, int64_t n, double scale) {
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<T> v((size_t)n);
    for (auto& x : v) x = (T)(nd(g) * scale);
    return v;
}

template <typename T>
void run_dtype() {
    std::mt19937_64 g(7);
    for (int dims : {1, 2, 3, 8}) {
        check_case<T>("normal", dims, normal<T>(g, 3000 * dims, 1e3), normal<T>(g, 500 * dims, 1e3), 6, 7);
        std::vector<T> lat((size_t)2000 * dims), ql((size_t)300 * dims);
        std::uniform_int_distribution<int> u(0, 5);
        for (auto& x : lat) x = (T)u(g);
        for (auto& x : ql) x = (T)(u(g) + 0.5 * (u(g) & 1));
        check_case<T>("integer lattice (ties)", dims, lat, ql, 5, 16);
        std::vector<T> withnan = normal<T>(g, 2000 * dims, 1e3);
        for (int r = 0; r < 20; ++r) withnan[(size_t)(r * 97 % 2000) * dims + r % dims] = std::numeric_limits<T>::quiet_NaN();
        withnan[5 * dims] = std::numeric_limits<T>::infinity();
        std::vector<T> qn = normal<T>(g, 300 * dims, 1e3);
        qn[0] = std::numeric_limits<T>::quiet_NaN();
        qn[dims] = -std::numeric_limits<T>::infinity();
        check_case<T>("NaN and inf coordinates", dims, withnan, qn, 5, 5);
        check_case<T>("one leaf per point pair", dims, normal<T>(g, 64 * dims, 1.0), normal<T>(g, 50 * dims, 1.0), 6, 40);
    }
    // the "Ran out of points" path
    try {
        std::vector<T> few((size_t)10 * 3, T(1));
        delete tk::build_tree(few.data(), 10, 3, 6, sizeof(T) == 4 ? tk::kF32 : tk::kF64, 0);
        std::printf("ran-out-of-points: no error\n");
        ++g_fail;
    } catch (const std::runtime_error&) {
    }
}

}  // namespace

int main() {
    try {
        run_dtype<float>();
        run_dtype<double>();
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
    if (g_fail) {
        std::printf("%d case(s) failed\n", g_fail);
        return 1;
    }
    std::printf("host selftest OK\n");
    return 0;
}
