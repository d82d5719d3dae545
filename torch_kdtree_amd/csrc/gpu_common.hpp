// Device-side building blocks shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace tk {

#define TK_HIP_CHECK(expr)                                                                          \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess)                                                                       \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " + \
                                     __FILE__ + ":" + std::to_string(__LINE__));                    \
    } while (0)

// ---------------------------------------------------------------------------------------------
// XCD-aware tile order.  The dispatcher hands workgroups to the 8 XCDs round-robin (workgroup b
// runs on XCD b % 8), and each XCD has its own L2.  xcd_tile maps workgroup b to a tile so that
// each XCD takes runs of TK_XCD_GROUP consecutive tiles, the runs dealt to the XCDs in turn:
// presorted neighbouring queries, which read the same leaves, then share one L2 instead of
// fetching those leaves into eight, while every XCD still sees all regions of the query set (one
// contiguous range per XCD, TK_XCD_GROUP 0, leaves the XCDs unevenly loaded when regions differ in
// cost).  The map is a bijection of [0, n_tiles) for any n_tiles.
// ---------------------------------------------------------------------------------------------
#ifndef TK_XCDS
#define TK_XCDS 8
#endif
#ifndef TK_XCD_GROUP
#define TK_XCD_GROUP 256
#endif

__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t n_tiles) {
    if (TK_XCDS <= 1) return b;
    const int64_t x = b % TK_XCDS, j = b / TK_XCDS;
    if (TK_XCD_GROUP > 0) {
        // runs of TK_XCD_GROUP tiles, dealt to the XCDs in turn; the last partial round in order
        constexpr int64_t G = TK_XCD_GROUP;
        const int64_t full = n_tiles / (G * TK_XCDS) * (G * TK_XCDS);
        if (b >= full) return b;
        return ((j / G) * TK_XCDS + x) * G + j % G;
    }
    const int64_t per = n_tiles / TK_XCDS, extra = n_tiles % TK_XCDS;
    return x * per + (x < extra ? x : extra) + j;
}

// ---------------------------------------------------------------------------------------------
// Point records in HBM.  A leaf is a contiguous run of records in structured order; each record
// holds the point's coordinates, padded so that one record is one (or a few) 16-byte loads, and
// the point's original index where the coordinates leave room for it in the padding: fp32 d=3 is
// exactly one float4 (x, y, z, index bits).  Coordinates that fill whole 16-byte words (fp32 d=8,
// fp64 d=2 and d=8) are stored alone, 32 instead of 48 bytes at fp32 d=8; a search that needs the
// original index of such a record reads it from the tree's `shuffled` map (rec_orig).
// ---------------------------------------------------------------------------------------------
template <typename T, int D>
struct RecLayout {
    static constexpr int kCoord = D * (int)sizeof(T);
    static constexpr bool kHasIdx = kCoord <= 4 || kCoord % 16 != 0;
    static constexpr int kRaw = kCoord + (kHasIdx ? 4 : 0);
    static constexpr int kBytes = kRaw <= 8 ? 8 : ((kRaw + 15) / 16) * 16;
    static constexpr int kWords = kBytes / 4;
    static constexpr int kIdxWord = kCoord / 4;  // when kHasIdx
};

template <typename T, int D>
struct Rec {
    T c[D];
    int32_t o;
};

template <typename T, int D>
__device__ __forceinline__ Rec<T, D> load_rec(const unsigned char* __restrict__ base, int32_t s) {
    using L = RecLayout<T, D>;
    uint32_t w[L::kWords];
    if constexpr (L::kBytes % 16 == 0) {
        const uint4* p = reinterpret_cast<const uint4*>(base + (size_t)s * L::kBytes);
#pragma unroll
        for (int i = 0; i < L::kBytes / 16; ++i) {
            const uint4 v = p[i];
            w[4 * i + 0] = v.x;
            w[4 * i + 1] = v.y;
            w[4 * i + 2] = v.z;
            w[4 * i + 3] = v.w;
        }
    } else {
        const uint2 v = reinterpret_cast<const uint2*>(base + (size_t)s * L::kBytes)[0];
        w[0] = v.x;
        w[1] = v.y;
    }
    Rec<T, D> r;
    __builtin_memcpy(r.c, w, sizeof(T) * D);
    if constexpr (L::kHasIdx)
        r.o = (int32_t)w[L::kIdxWord];
    else
        r.o = 0;  // not stored: rec_orig reads it from the tree's map
    return r;
}

// ---------------------------------------------------------------------------------------------
// Arithmetic.  The reference computes every distance and bound in T with separate rounding of
// each subtraction, product and sum (src/kdtree.hpp:546-614).  All kernels are compiled with
// -ffp-contract=off, so a*b+c below is a rounded product followed by a rounded sum.
// ---------------------------------------------------------------------------------------------
// Per-lane point/offset vectors live in clang vector types: element access with a run-time
// axis then lowers to register selects instead of forcing the values into memory.
template <typename T, int D>
using Pt = T __attribute__((ext_vector_type(D)));

template <typename T, int D, typename A, typename B>
__device__ __forceinline__ T sq_dist(const A& q, const B& r) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const T v = q[i] - r[i];
        s = s + v * v;
    }
    return s;
}

// Reference pruning bound, src/kdtree.hpp:606-614: v = pp - q, v[a] += m - pp[a], sum of squares.
template <typename T, int D, typename V>
__device__ __forceinline__ T proj_bound(const V& q, const V& pp, int a, T m) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        T v = pp[i] - q[i];
        if (i == a) v = v + (m - pp[i]);
        s = s + v * v;
    }
    return s;
}

// Cell bound used by the lower-index tie policy: v[a] = m - q[a] rounded once, so that by the
// monotonicity of rounding it never exceeds the computed distance of any point beyond the
// split (the reference's two-rounding form above can exceed it by an ulp).  pp[i] is q[i] or a
// split the path crossed on axis i, so |p[i] - q[i]| >= |pp[i] - q[i]| for every point p there.
template <typename T, int D, typename V>
__device__ __forceinline__ T clean_bound(const V& q, const V& pp, int a, T m) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const T v = (i == a) ? (m - q[i]) : (pp[i] - q[i]);
        s = s + v * v;
    }
    return s;
}

// Point bounding box of a tree node, lo[D] then hi[D], each padded to kPad elements so that a
// box is a whole number of aligned 16-byte words (8 bytes for a 1-d f32 box).
template <typename T, int D>
struct BoxLayout {
    static constexpr int kPad = D == 3 ? 4 : D;
    static constexpr int kBytes = 2 * kPad * (int)sizeof(T);
    static constexpr int kWords = kBytes / 4;
};

// Squared distance from q to the box of heap node `node`: per axis the gap to the box, one
// rounded difference (max(lo - q, q - hi, 0), the two candidates cannot both be positive),
// squared and summed in axis order like a distance.  By the monotonicity of rounding it never
// exceeds the computed distance of a point in the box.  An empty box (lo = +inf, hi = -inf)
// gives +inf.
// The coordinates of one box, loaded (lo then hi).
template <typename T, int D>
struct BoxVals {
    T lo[D], hi[D];
};

template <typename T, int D>
__device__ __forceinline__ BoxVals<T, D> load_box(const unsigned char* __restrict__ boxes, int32_t node) {
    using L = BoxLayout<T, D>;
    uint32_t w[L::kWords];
    if constexpr (L::kBytes % 16 == 0) {
        const uint4* p = reinterpret_cast<const uint4*>(boxes + (size_t)node * L::kBytes);
#pragma unroll
        for (int i = 0; i < L::kBytes / 16; ++i) {
            const uint4 v = p[i];
            w[4 * i + 0] = v.x;
            w[4 * i + 1] = v.y;
            w[4 * i + 2] = v.z;
            w[4 * i + 3] = v.w;
        }
    } else {
        const uint2 v = reinterpret_cast<const uint2*>(boxes + (size_t)node * L::kBytes)[0];
        w[0] = v.x;
        w[1] = v.y;
    }
    BoxVals<T, D> b;
    __builtin_memcpy(b.lo, w, sizeof(T) * D);
    __builtin_memcpy(b.hi, reinterpret_cast<const unsigned char*>(w) + sizeof(T) * L::kPad, sizeof(T) * D);
    return b;
}

template <typename T, int D>
__device__ __forceinline__ T box_dist(const BoxVals<T, D>& b, const Pt<T, D>& q) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const T v = fmax(fmax(b.lo[i] - q[i], q[i] - b.hi[i]), T(0));
        s = s + v * v;
    }
    return s;
}

template <typename T, int D>
__device__ __forceinline__ T aabb_bound(const unsigned char* __restrict__ boxes, int32_t node, const Pt<T, D>& q) {
    return box_dist<T, D>(load_box<T, D>(boxes, node), q);
}

template <typename T>
__device__ __forceinline__ T pos_inf();
template <>
__device__ __forceinline__ float pos_inf<float>() { return __builtin_huge_valf(); }
template <>
__device__ __forceinline__ double pos_inf<double>() { return __builtin_huge_val(); }

// ---------------------------------------------------------------------------------------------
// k-best list in VGPRs.  K is the compile-time capacity (a power of two >= k); the first K-k
// slots hold -inf and never move, so the live list is slots K-k .. K-1 and the worst entry is
// always slot K-1.  Fully unrolled, branch-free compare-and-shift.
//
// Each entry is (distance, key).
// kLower == false (reference policy, src/nndistance.hpp:41-85): key = structured index; insert
//   iff d < worst, at the upper-bound position (after every equal distance already present).
// kLower == true (lower-index policy): key = original index; the list is ordered by (d, key).
//   Empty slots start as (+inf, 0), which no (+inf, key) or NaN entry orders before: like the
//   reference policy, an infinite or NaN distance is never inserted, and a list left unfilled
//   (a query with a NaN coordinate) reports (+inf, index 0).
// ---------------------------------------------------------------------------------------------
template <typename T, int K, bool kLower>
struct KBest {
    T d[K];
    int32_t x[K];
    static constexpr bool kOrigKeys = kLower;  // keys are original (else structured) indices

    __device__ __forceinline__ void init(int k) {
        const int pad = K - k;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            d[j] = j < pad ? -pos_inf<T>() : pos_inf<T>();
            x[j] = kLower ? (j < pad ? -1 : 0) : 0;
        }
    }

    __device__ __forceinline__ T worst() const { return d[K - 1]; }
    __device__ __forceinline__ T dist(int j) const { return d[j]; }
    __device__ __forceinline__ int32_t key(int j) const { return x[j]; }
    // slot j set directly (a result known without a search)
    __device__ __forceinline__ void set(int j, T dist, int32_t key) {
        d[j] = dist;
        x[j] = key;
    }

    __device__ __forceinline__ bool accepts(T dist, int32_t key) const {
        if constexpr (kLower)
            return dist < d[K - 1] || (dist == d[K - 1] && key < x[K - 1]);
        else
            return dist < d[K - 1];
    }
    // The second acceptance test of a staged leaf scan (against the current list).
    __device__ __forceinline__ bool admit(T dist, int32_t key) { return accepts(dist, key); }

    // The worst entry as of now, for acceptance tests against a stale list.
    struct Stale {
        T d;
        int32_t x;
    };
    __device__ __forceinline__ Stale stale() const { return Stale{d[K - 1], kLower ? x[K - 1] : 0}; }
    static __device__ __forceinline__ bool accepts_stale(const Stale& w, T dist, int32_t key) {
        if constexpr (kLower)
            return dist < w.d || (dist == w.d && key < w.x);
        else
            return dist < w.d;
    }

    __device__ __forceinline__ bool after(int j, T dist, int32_t key) const {
        if constexpr (kLower)
            return d[j] > dist || (d[j] == dist && x[j] > key);
        else
            return d[j] > dist;
    }

    // Sorted insert of an accepted candidate.  With p the insert position, slot j becomes d[j-1]
    // (j > p), dist (j == p) or d[j] (j < p); for a sorted list that value is the median of
    // (d[j-1], d[j], dist) whatever p is, so distances need one med3 per slot and only the keys
    // need the position masks.
    __device__ __forceinline__ void insert(T dist, int32_t key) {
        bool here_next = after(K - 1, dist, key);
#pragma unroll
        for (int j = K - 1; j > 0; --j) {
            const bool shift = after(j - 1, dist, key);
            const bool here = here_next;
            here_next = shift;
            d[j] = med3(d[j - 1], d[j], dist);
            x[j] = shift ? x[j - 1] : (here ? key : x[j]);
        }
        d[0] = here_next ? dist : d[0];
        x[0] = here_next ? key : x[0];
    }

    // The median of (a, b, c) for a <= b (neighbouring slots of the sorted list): max(a, min(b, c)),
    // two instructions for double (there is no v_med3_f64) instead of four for any order.
    static __device__ __forceinline__ T med3(T a, T b, T c) {
        if constexpr (sizeof(T) == 4)
            return __builtin_amdgcn_fmed3f(a, b, c);
        else
            return fmax(a, fmin(b, c));
    }
};

// The lower-index list for float: an entry is one 64-bit key, the distance's bits above the
// original index.  Distances are never negative, so their bits order like the distances (+inf
// included; a NaN orders after everything and is never accepted), and one unsigned compare orders
// (distance, index) pairs lexicographically: an insertion takes one compare and two 64-bit selects
// per slot.  Padding slots hold 0, which nothing orders before; empty slots hold (+inf, 0).
template <int K>
struct KBest<float, K, true> {
    uint64_t k[K];
    static constexpr bool kOrigKeys = true;

    static __device__ __forceinline__ uint64_t pack(float d, int32_t key) {
        return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)key;
    }

    __device__ __forceinline__ void init(int kk) {
        const int pad = K - kk;
#pragma unroll
        for (int j = 0; j < K; ++j) k[j] = j < pad ? uint64_t(0) : pack(__builtin_huge_valf(), 0);
    }

    __device__ __forceinline__ float worst() const { return __uint_as_float((uint32_t)(k[K - 1] >> 32)); }
    __device__ __forceinline__ float dist(int j) const { return __uint_as_float((uint32_t)(k[j] >> 32)); }
    __device__ __forceinline__ int32_t key(int j) const { return (int32_t)(uint32_t)k[j]; }
    __device__ __forceinline__ void set(int j, float dist, int32_t key) { k[j] = pack(dist, key); }

    __device__ __forceinline__ bool accepts(float dist, int32_t key) const { return pack(dist, key) < k[K - 1]; }
    __device__ __forceinline__ bool admit(float dist, int32_t key) { return accepts(dist, key); }

    using Stale = uint64_t;
    __device__ __forceinline__ Stale stale() const { return k[K - 1]; }
    static __device__ __forceinline__ bool accepts_stale(Stale w, float dist, int32_t key) { return pack(dist, key) < w; }

    __device__ __forceinline__ void insert(float dist, int32_t key) {
        const uint64_t c = pack(dist, key);
        bool here_next = k[K - 1] > c;
#pragma unroll
        for (int j = K - 1; j > 0; --j) {
            const bool shift = k[j - 1] > c;
            const bool here = here_next;
            here_next = shift;
            k[j] = shift ? k[j - 1] : (here ? c : k[j]);
        }
        k[0] = here_next ? c : k[0];
    }
};

}  // namespace tk
