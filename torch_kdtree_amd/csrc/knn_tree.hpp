// KD-tree KNN traversal kernels for gfx950.
//
// One query per lane.  Each lane runs the reference's near-first depth-first search
// (src/kdtree.cpp:16-110) without a stack or a tag array: the state is the node index, its
// depth and the projected point pp (D registers).  Coming back up from a child, the lane
// recomputes which side was the near one (q[a] < m) to tell whether the far side is still
// to be tested; that is exactly the information the reference's per-node tags hold for the
// nodes on the active path, and nodes off the path are never revisited.  The pruning bound,
// the pp updates and the candidate insertion repeat the reference's arithmetic, so with the
// reference tie policy every lane produces what the reference CPU search produces.
//
// Queries arrive in an order given by `order` (a presort below: Morton buckets of the query
// positions by default, or buckets by near leaf), so the lanes of a wavefront walk similar paths
// and their record loads coalesce.
#pragma once

#include <type_traits>

#include "gpu_common.hpp"

namespace tk {

// Device view of a tree.  kBatched: the arrays belong to an arena of many trees and this lane's
// tree starts at the per-lane offsets below (set per query); otherwise the offsets are zero and
// fold away.  Leaf offsets hold arena-wide record positions in both cases.
template <typename T, bool kBatched = false>
struct TreeView {
    const T* __restrict__ split;            // [n_nodes] heap order
    const unsigned char* __restrict__ recs; // [n_points] records, structured order
    const int32_t* __restrict__ leaf_off;   // [n_leaves + 1]
    const int32_t* __restrict__ shuffled;   // [n_points] original index of a structured position
    const int32_t* __restrict__ inv_shuffled; // [n_points] structured position of an original index
    const unsigned char* __restrict__ boxes;  // [n_nodes + n_leaves] point boxes, heap order
    int levels;
    int32_t n_nodes;
    // batched only: where this tree's nodes, boxes, leaf offsets and records start in the arena
    int32_t node_base = 0, box_base = 0, leaf_base = 0, rec_base = 0;
    // point boxes of each leaf's record groups (kGroupRecs records from the leaf's first), and the
    // index of each leaf's first group box; null: no group tests
    const unsigned char* __restrict__ group_boxes = nullptr;
    const int32_t* __restrict__ group_base = nullptr;
    // Every split value separates its node's points (true when all coordinates and split values
    // are finite).  A NaN coordinate leaves std::sort's order unspecified and an overflowing median
    // is +-inf, so the lower-index policy then prunes with point boxes only (they hold the actual
    // extents of every subtree whatever the order).
    bool split_bounds = true;
    // Lower-index searches launch the 64-bit keyed list's kernel (kTieOK false) rather than the
    // KBestTie one: a cloud whose neighbours tie often, where nearly every lane would be searched
    // twice.  Read by the launcher.
    bool lower_direct = false;

    __device__ __forceinline__ T split_at(int32_t node) const { return split[kBatched ? node_base + node : node]; }
    __device__ __forceinline__ int32_t leaf_begin(int32_t leaf) const {
        return leaf_off[kBatched ? leaf_base + leaf : leaf];
    }
    __device__ __forceinline__ int32_t leaf_end(int32_t leaf) const {
        return leaf_off[kBatched ? leaf_base + leaf + 1 : leaf + 1];
    }
    __device__ __forceinline__ int32_t box_of(int32_t heap_node) const {
        return kBatched ? box_base + heap_node : heap_node;
    }
};

// base[i] for 0 <= i < 2^31, addressed with a 32-bit byte offset (keeps i a 32-bit value; a
// sign-extended 64-bit index would tie each k-list key to an aligned register pair).
__device__ __forceinline__ int32_t load_i32_at(const int32_t* __restrict__ base, int32_t i) {
    return *reinterpret_cast<const int32_t*>(reinterpret_cast<const char*>(base) + ((uint32_t)i << 2));
}

// Records per group box (the host writes them in runs of the same length), and the mean leaf
// size from which a tree gets group boxes.
constexpr int kGroupRecs = 16;
constexpr int kGroupMinLeaf = 4 * kGroupRecs;

// Threads per block of the traversal kernel; per-lane LDS staging slots are strided by it.
#ifndef TK_TREE_BLOCK
#define TK_TREE_BLOCK 128
#endif
#ifndef TK_STAGE_CHUNK_F32
#define TK_STAGE_CHUNK_F32 8
#endif
#ifndef TK_STAGE_CHUNK_F32_LOWER
#define TK_STAGE_CHUNK_F32_LOWER 4
#endif
#ifndef TK_STAGE_CHUNK_SMALLK
#define TK_STAGE_CHUNK_SMALLK 8
#endif
#ifndef TK_STAGE_CHUNK_F32_K16
#define TK_STAGE_CHUNK_F32_K16 16
#endif
#ifndef TK_SCAN_LOADS_F32
#define TK_SCAN_LOADS_F32 4
#endif
#ifndef TK_STAGE_CHUNK_F64
#define TK_STAGE_CHUNK_F64 16
#endif
#ifndef TK_SCAN_LOADS_F64
#define TK_SCAN_LOADS_F64 4
#endif
// Whether the batched kernels' climb also loads the next level's box ahead (on: 24.9 -> 24.5 ms
// per 256-cloud replay; with the batch occupancy targets the k bucket 16 kernel stays at 127
// VGPRs without spills).
#ifndef TK_BATCH_CLIMB_PREFETCH
#define TK_BATCH_CLIMB_PREFETCH 1
#endif
// The same for d > 3, where a box is 2d values (on: d = 8, k = 32, 500K queries 95.9 ms; off:
// 99.0 ms, with the same registers).
#ifndef TK_WIDE_CLIMB_PREFETCH
#define TK_WIDE_CLIMB_PREFETCH 1
#endif
// Largest list held in registers (larger k use the LDS-list kernel).
#ifndef TK_REG_LIST_MAX_K
#define TK_REG_LIST_MAX_K 32
#endif
#ifndef TK_STAGE_CHUNK_WIDE
#define TK_STAGE_CHUNK_WIDE 4
#endif
// Occupancy target (waves per SIMD) of the register-list traversal kernels, for float searches
// with d <= 3 and lists of at most 16 (measured on MI355X, plain workload, k = 16):
//   reference policy, TK_F32_WAVES = 7: 72 VGPRs, no spills, 8.48 ms (5 waves: the same 72 VGPRs
//     scheduled differently, 8.77 ms; no target: 91 VGPRs, 9.14 ms; 8 waves: spills, 9.29 ms);
//   lower-index policy: lists of 8 and 16 (KBestTie, keyed like the reference policy's) take the
//     reference policy's shape (k = 16: 72 VGPRs, 8.29-8.63 ms); TK_F32_LOWER_WAVES = 7 for its
//     64-bit keyed lists (lists of at most 4; with that list at k = 16, chunks of 4 at 7 waves had
//     been best: 72 VGPRs and 16 bytes of spills, 9.75 ms; with chunks of 8, 92 VGPRs at 5 waves,
//     10.62 ms).
// Other kernels: see TK_F64_WAVES and TK_WIDE_WAVES below, no target for the rest.
// TK_WAVES_PER_EU overrides every target (0 = no target anywhere).
#ifndef TK_WAVES_PER_EU
#define TK_WAVES_PER_EU -1
#endif
#ifndef TK_F32_WAVES
#define TK_F32_WAVES 7
#endif
#ifndef TK_F32_LOWER_WAVES
#define TK_F32_LOWER_WAVES 7
#endif
// Lower-index float searches with lists of 8, 16 or 32 keep the reference policy's list plus tie
// flags (KBestTie) instead of the 64-bit keyed list (0).
#ifndef TK_LOWER_TIE_LIST
#define TK_LOWER_TIE_LIST 1
#endif
// Double, d = 3, k = 16: no target (119 VGPRs, 4 waves; a 5-wave target spills 48 bytes, and a
// 4-wave one measured neutral).  Float, d = 8, k = 32: 4 waves (128 VGPRs, 12 bytes of spills,
// 95.9 ms per 500K queries against 99.1 ms at 3 waves and 133 VGPRs).
#ifndef TK_F64_WAVES
#define TK_F64_WAVES 0
#endif
#ifndef TK_WIDE_WAVES
#define TK_WIDE_WAVES 4
#endif
// kTie: the lower-index search keeps a KBestTie list (TreeList).
template <typename T, int D, int K, bool kLower, bool kTie = false>
struct TreeOccupancy {
    static constexpr int kWaves =
        TK_WAVES_PER_EU >= 0 ? TK_WAVES_PER_EU
        : (sizeof(T) == 4 && D <= 3 && K <= 16) ? (kLower && !kTie ? TK_F32_LOWER_WAVES : TK_F32_WAVES)
        : (sizeof(T) == 8 && D <= 3 && K == 16 && !kLower) ? TK_F64_WAVES
        : (sizeof(T) == 4 && D == 8 && K == 32 && (!kLower || kTie)) ? TK_WIDE_WAVES
                                                                     : 0;
};
#define TK_TREE_OCCUPANCY(T, D, K, L, TIE) __attribute__((amdgpu_waves_per_eu(TreeOccupancy<T, D, K, L, TIE>::kWaves > 0 ? TreeOccupancy<T, D, K, L, TIE>::kWaves : 1)))
// The same for the batched search kernels (more live state per lane: the cloud's offsets).
#ifndef TK_BATCH_WAVES_K16
#define TK_BATCH_WAVES_K16 4
#endif
#ifndef TK_BATCH_WAVES_K8
#define TK_BATCH_WAVES_K8 5
#endif
// k buckets 4 and 1 (67-75 VGPRs, 6 waves): 7- and 8-wave targets measured within noise.
#ifndef TK_BATCH_WAVES_K4
#define TK_BATCH_WAVES_K4 0
#endif
#ifndef TK_BATCH_WAVES_K1
#define TK_BATCH_WAVES_K1 0
#endif
template <typename T, int D, int K, bool kLower>
struct BatchOccupancy {
    static constexpr int kWaves =
        (sizeof(T) == 4 && D <= 3 && !kLower)
            ? (K == 16 ? TK_BATCH_WAVES_K16
                       : K == 8 ? TK_BATCH_WAVES_K8 : K == 4 ? TK_BATCH_WAVES_K4 : K == 1 ? TK_BATCH_WAVES_K1 : 0)
            : 0;
};
#define TK_BATCH_OCCUPANCY(T, D, K, L) __attribute__((amdgpu_waves_per_eu(BatchOccupancy<T, D, K, L>::kWaves > 0 ? BatchOccupancy<T, D, K, L>::kWaves : 1)))
constexpr int kTreeBlock = TK_TREE_BLOCK;
constexpr int kTreeBlockStride = kTreeBlock;

// k-best list held in LDS for k beyond the register buckets: slot j of lane t lives at
// base[j * blockDim + t] (consecutive lanes hit consecutive banks).
template <typename T, bool kLower>
struct KBestLds {
    T* d;
    int32_t* s;
    int32_t* o;
    int k;
    int stride;
    T w;        // cached worst distance
    int32_t wo; // cached worst original index

    // empty slots: (+inf, structured index s0, original index 0); see KBest for why 0
    __device__ __forceinline__ void init(int32_t s0) {
        for (int j = 0; j < k; ++j) {
            d[j * stride] = pos_inf<T>();
            s[j * stride] = s0;
            if constexpr (kLower) o[j * stride] = 0;
        }
        w = pos_inf<T>();
        wo = 0;
    }
    __device__ __forceinline__ T worst() const { return w; }
    __device__ __forceinline__ bool accepts(T dist, int32_t orig) const {
        if constexpr (kLower)
            return dist < w || (dist == w && orig < wo);
        else
            return dist < w;
    }
    __device__ __forceinline__ bool after(int j, T dist, int32_t orig) const {
        const T dj = d[j * stride];
        if constexpr (kLower)
            return dj > dist || (dj == dist && o[j * stride] > orig);
        else
            return dj > dist;
    }
    __device__ __forceinline__ void insert(T dist, int32_t sidx, int32_t orig) {
        int j = k - 1;
        while (j > 0 && after(j - 1, dist, orig)) {
            d[j * stride] = d[(j - 1) * stride];
            s[j * stride] = s[(j - 1) * stride];
            if constexpr (kLower) o[j * stride] = o[(j - 1) * stride];
            --j;
        }
        d[j * stride] = dist;
        s[j * stride] = sidx;
        if constexpr (kLower) o[j * stride] = orig;
        w = d[(k - 1) * stride];
        if constexpr (kLower) wo = o[(k - 1) * stride];
    }
};

// The reference's own list, for trees with non-finite values under the reference tie policy.  Held
// in LDS like KBestLds.  A candidate is rejected only when d >= worst (src/nndistance.hpp:77) and
// is placed by the reference's binary search (upperBound, src/nndistance.hpp:41-60), so a NaN
// distance is taken and goes to the front, and the search's path through a list holding NaN
// entries decides where each later value lands; a NaN reaching the last slot makes worst NaN,
// after which every candidate is taken and every bound test fails, as in the reference.
template <typename T>
struct KBestExact {
    T* d;
    int32_t* s;
    int k;
    int stride;

    __device__ __forceinline__ void init() {
        for (int j = 0; j < k; ++j) {
            d[j * stride] = pos_inf<T>();
            s[j * stride] = 0;
        }
    }
    __device__ __forceinline__ T worst() const { return d[(k - 1) * stride]; }
    __device__ __forceinline__ bool accepts(T dist, int32_t) const { return !(dist >= worst()); }
    __device__ __forceinline__ void insert(T dist, int32_t sidx, int32_t) {
        int first = 0, count = k;
        while (count > 0) {
            const int step = count / 2;
            if (dist >= d[(first + step) * stride]) {
                first += step + 1;
                count -= step + 1;
            } else {
                count = step;
            }
        }
        if (first >= k) return;  // not reached after accepts(); kept as the reference keeps it
        for (int j = k - 1; j > first; --j) {
            d[j * stride] = d[(j - 1) * stride];
            s[j * stride] = s[(j - 1) * stride];
        }
        d[first * stride] = dist;
        s[first * stride] = sidx;
    }
};

// Lower-index policy on the reference policy's register list (one med3 per slot; the keys are
// structured indices, as under the reference policy, since the key order no longer decides
// anything: ties are settled by the flag below).  The two policies' lists can differ only through
// exact distance ties: when no point the search scans outside the list has a distance equal to the
// final worst W and the list's k distances are distinct, both hold the k points of smallest
// distance in ascending order.  `tie` records every event after which that may not hold:
// * a candidate equal to the current worst in the second pass (the first pass passes every
//   distance <= the stale worst, so a point at W is always a candidate, and the current worst it
//   then meets is >= W);
// * an insertion whose evicted entry equals the new worst (an evicted distance is >= W, and
//   equal to W only if equal to that worst, which never falls below W; empty slots aside);
// and unsure() adds equal distances among the k live entries (their order must follow the
// original index).  The search prunes with the lower-index policy's tests, so every point with
// distance <= W is scanned.  A lane that becomes unsure stops after that leaf pair and searches
// again, node by node, with the lower-index list (research_lower): rare on clouds without exact
// fp32 distance ties.  Trees whose clouds are full of them (lattices, rounded coordinates,
// duplicated points; Tree::tie_prone) launch the lower-index list's kernel instead.
template <typename T, int K>
struct KBestTie : KBest<T, K, false> {
    using Base = KBest<T, K, false>;
    bool tie;

    __device__ __forceinline__ void init(int k) {
        Base::init(k);
        tie = false;
    }
    static __device__ __forceinline__ bool accepts_stale(const typename Base::Stale& w, T dist, int32_t) {
        return dist <= w.d;
    }
    __device__ __forceinline__ bool admit(T dist, int32_t) {
        tie |= dist == Base::worst();
        return dist < Base::worst();
    }
    // (an entry at +inf is an empty slot: a +inf distance is never inserted, so evicting one is no tie)
    __device__ __forceinline__ void insert(T dist, int32_t key) {
        const T old = Base::worst();
        Base::insert(dist, key);
        tie |= dist < old && Base::worst() == old && old < pos_inf<T>();
    }
    __device__ __forceinline__ bool unsure(int k) const {
        bool t = tie;
#pragma unroll
        for (int j = 0; j + 1 < K; ++j) t |= j >= K - k && Base::d[j] == Base::d[j + 1];
        return t;
    }
    // unsure() during the search: equal entries are looked for only in a list without padding slots
    // (k == K, a wave-uniform test), where every adjacent pair is live
    __device__ __forceinline__ bool unsure_early(int k) const {
        bool t = tie;
        if (k == K) {
#pragma unroll
            for (int j = 0; j + 1 < K; ++j) t |= Base::d[j] == Base::d[j + 1];
        }
        return t;
    }
};

// The list of a single-tree register search.  Lower-index float searches with lists of 8, 16 or
// 32 take KBestTie (TK_LOWER_TIE_LIST 1): the 64-bit keyed list spends more issue slots per
// insertion than the med3 list plus the tie flags, and needs the original indices staged with the
// distances (plain workload, k = 16: 9.96-9.99 ms against 8.29-8.63; k = 32: 22.1-22.2 against
// 16.9-17.1; d8_mixture, k = 32, 500K queries: 87.1-90.4 against 78.1-79.1).  Lists of at most 4
// keep it: their insertions are short.
// kTieOK false: the 64-bit keyed list in any case (trees whose clouds tie often, Tree::tie_prone).
template <typename T, int K, bool kLower, bool kTieOK = true>
struct TreeList {
    using type = KBest<T, K, kLower>;
    static constexpr bool kTie = false;
};
template <int K, bool kTieOK>
struct TreeList<float, K, true, kTieOK> {
    static constexpr bool kTie = kTieOK && TK_LOWER_TIE_LIST && (K == 8 || K == 16 || K == 32);
    using type = typename std::conditional<kTie, KBestTie<float, K>, KBest<float, K, true>>::type;
};

// The original index of the record at structured position s: from the record where its layout
// has room for it, else from the tree's `shuffled` map (indexed, like the records, by position).
template <typename T, int D, bool kB>
__device__ __forceinline__ int32_t rec_orig(const TreeView<T, kB>& tv, int32_t s, const Rec<T, D>& r) {
    if constexpr (RecLayout<T, D>::kHasIdx)
        return r.o;
    else
        return load_i32_at(tv.shuffled, s);
}

template <typename T, int D, typename List, bool kB>
__device__ __forceinline__ void scan_leaf(const TreeView<T, kB>& tv, int32_t leaf, const Pt<T, D>& q, List& kb) {
    const int32_t beg = tv.leaf_begin(leaf);
    const int32_t end = tv.leaf_end(leaf);
    for (int32_t s = beg; s < end; ++s) {
        const Rec<T, D> r = load_rec<T, D>(tv.recs, s);
        const T dd = sq_dist<T, D>(q, r.c);
        const int32_t o = rec_orig(tv, s, r);
        if (kb.accepts(dd, o)) kb.insert(dd, s, o);
    }
}

// Bits l < 32 with l mod D == a.
template <int D>
__device__ __forceinline__ uint32_t axis_levels(int a) {
    uint32_t m = 0;
#pragma unroll
    for (int l = 0; l < 32; ++l)
        if (l % D == a) m |= 1u << l;
    return m;
}

// Pruning test of the lower-index policy: a side is skipped only when its bound exceeds worst.  A
// NaN bound (a NaN split value of a tree with non-finite coordinates) never prunes.
template <typename T>
__device__ __forceinline__ bool lower_passes(T bound, T worst) {
    return !(bound > worst);
}

// The lower-index policy's split-plane test: passes on trees whose splits do not bound their points.
template <typename T, int D, bool kB>
__device__ __forceinline__ bool lower_split_passes(const TreeView<T, kB>& tv, const Pt<T, D>& q, const Pt<T, D>& pp,
                                                   int a, T m, T worst) {
    return !tv.split_bounds || lower_passes(clean_bound<T, D>(q, pp, a, m), worst);
}

// Far side test for the child `far_child` of split (a, m).  Reference policy: the reference's own
// test (proj_bound < worst) decides whether the reference would look at the far side; the
// distance to the box of the points there then tells whether that look could change anything:
// when it is >= worst every point there has d >= worst and would be rejected.  Lower-index
// policy: the cell bound with one rounding per term and the box distance, both non-strict, so
// that a point at exactly the worst distance with a lower index is still found.
// Returns 0: skip (the reference skips too), 1: enter, 2: skip a side the reference enters.
template <typename T, int D, bool kLower, bool kB>
__device__ __forceinline__ int far_side(const TreeView<T, kB>& tv, const Pt<T, D>& q, const Pt<T, D>& pp, int a, T m,
                                       int32_t far_child, T worst) {
    if constexpr (kLower) {
        if (!lower_split_passes<T, D>(tv, q, pp, a, m, worst)) return 0;
        return lower_passes(aabb_bound<T, D>(tv.boxes, tv.box_of(far_child), q), worst) ? 1 : 0;
    } else {
        if (!(proj_bound<T, D>(q, pp, a, m) < worst)) return 0;
        return aabb_bound<T, D>(tv.boxes, tv.box_of(far_child), q) < worst ? 1 : 2;
    }
}

// Whether a leaf can hold a point the list accepts (the reference scans it anyway; when the
// distance to its box is >= worst it would reject every point).
template <typename T, int D, bool kLower, bool kB>
__device__ __forceinline__ bool leaf_useful(const TreeView<T, kB>& tv, const Pt<T, D>& q, int32_t leaf, T worst) {
    const T b = aabb_bound<T, D>(tv.boxes, tv.box_of(tv.n_nodes + leaf), q);
    return kLower ? lower_passes(b, worst) : b < worst;
}

// Run-time axis access.  The element is taken from / put into a copy of the whole vector, so
// the lane state itself is only ever read and written whole and stays in registers (an
// element address computed from a run-time index would pin the surrounding object in memory).
template <typename T, int D>
__device__ __forceinline__ T ax_get(const Pt<T, D>& v, int a) {
    if constexpr (D == 1) {
        return v[0];
    } else {
        const Pt<T, D> c = v;
        return c[a];
    }
}

template <typename T, int D>
__device__ __forceinline__ void ax_set(Pt<T, D>& v, int a, T x) {
    if constexpr (D == 1) {
        v[0] = x;
    } else {
        Pt<T, D> c = v;
        c[a] = x;
        v = c;
    }
}

template <int D>
__device__ __forceinline__ int ax_next(int a) { return a + 1 == D ? 0 : a + 1; }
template <int D>
__device__ __forceinline__ int ax_prev(int a) { return a == 0 ? D - 1 : a - 1; }

// Records past the last point that a chunked leaf scan may read (and ignore): the record buffer is
// allocated with this many extra records, so the CH loads of a chunk need no per-point guard.
constexpr int kRecPad = 16;

// Leaf scan in two passes, for register lists.  Pass 1 computes the distances of CH consecutive
// records (loads issued back to back; those past the leaf are ignored), tests them against the
// worst distance at the start of the chunk and stages them in LDS; pass 2 pops each lane's
// accepted points in scan order and inserts those still accepted by the current list.  A
// wavefront then runs the insertion network once per accepted point of its busiest lane instead
// of once per point any lane accepts.  The outcome is that of a one-pass scan: the list after
// pass 2 holds exactly the points a one-pass scan would have inserted, because a point rejected
// by the stale worst is rejected by every later (smaller or equal) worst as well.
// In pass 2 a candidate that the current list rejects is inserted as +inf, which leaves the list
// unchanged: the insertion is then unconditional and updates the list in place.
template <typename T, int D, int K, bool kLower, int CH, int LD, int SB = kTreeBlockStride, bool kGroups = false,
          bool kB, typename List>
__device__ __forceinline__ void scan_leaf_staged(const TreeView<T, kB>& tv, int32_t leaf, const Pt<T, D>& q,
                                                 List& kb, T* __restrict__ st_d,
                                                 int32_t* __restrict__ st_o) {
    static_assert(CH % LD == 0 && CH <= 32, "a chunk is a whole number of load groups, at most 32 points");
    static_assert(!kGroups || kGroupRecs % CH == 0, "chunks nest in record groups");
    const int32_t beg = tv.leaf_begin(leaf);
    const int32_t end = tv.leaf_end(leaf);
    const int32_t gb = kGroups ? tv.group_base[leaf] : 0;
    for (int32_t c0 = beg; c0 < end; c0 += CH) {
        // With kGroups (the tree has group boxes): a group of records whose box lies at least the
        // worst distance away holds no point the list would take (every distance is at least the box
        // distance, in rounded arithmetic too), so skipping it changes nothing but the work.
        if (kGroups && (c0 - beg) % kGroupRecs == 0) {
            const T b = aabb_bound<T, D>(tv.group_boxes, gb + (c0 - beg) / kGroupRecs, q);
            if (kLower ? !lower_passes(b, kb.worst()) : !(b < kb.worst())) {
                c0 += kGroupRecs - CH;
                continue;
            }
        }
        const int n = min(CH, end - c0);
        const auto w = kb.stale();
        uint32_t mask = 0;
        // pass 1, LD records in flight at a time; every group is read (records past the leaf's
        // end are padding or the next leaf's, and are masked out), so the loads of a chunk are
        // issued without a data-dependent exit
#pragma unroll
        for (int g = 0; g < CH; g += LD) {
            Rec<T, D> rg[LD];
#pragma unroll
            for (int jj = 0; jj < LD; ++jj) rg[jj] = load_rec<T, D>(tv.recs, c0 + g + jj);
            // keeps the group's loads issued together ahead of their use
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int jj = 0; jj < LD; ++jj) {
                const int j = g + jj;
                const Rec<T, D>& r = rg[jj];
                const T dd = sq_dist<T, D>(q, r.c);
                const int32_t o = List::kOrigKeys ? rec_orig(tv, c0 + j, r) : 0;
                const bool acc = kb.accepts_stale(w, dd, o);
                mask |= acc ? (1u << j) : 0u;
                st_d[j * SB] = dd;
                if constexpr (List::kOrigKeys) st_o[j * SB] = o;
            }
        }
        mask &= n >= 32 ? ~0u : ((1u << n) - 1u);  // only the leaf's points
        // pass 2
        while (mask) {
            const int j = __builtin_ctz(mask);
            mask &= mask - 1u;
            const T dd = st_d[j * SB];
            const int32_t o = List::kOrigKeys ? st_o[j * SB] : 0;
            const bool ok = kb.admit(dd, o);
            kb.insert(ok ? dd : pos_inf<T>(), ok ? (List::kOrigKeys ? o : c0 + j) : 0x7fffffff);
        }
    }
}

template <int D>
__device__ __forceinline__ int axis_of_level(int l) { return l % D; }

// The search of one query; `kb` must be initialised.
//
// Besides pp (the reference's projected point, kept exactly as the reference keeps it) each lane
// carries far_mask, the levels at which the current path took a far child.  Leaves and far sides
// are also tested against the distance from q to the bounding box of their points.  When the
// reference would enter a far side whose box distance is >= worst, the lane skips it and applies
// the only lasting effect the reference's visit would have had: pp[a] = q[a] for the axes of the
// levels below (the reference always walks the near path down to the leaf-parent level and
// resets each axis it leaves).
//
// Climbing: coming back up, a level whose far child is on the path (bit set in far_mask) is
// simply left (its axis is reset in pp); only levels entered through their near child need a
// far-side test.  So the climb jumps from one such level to the next with bit scans and
// handles the far levels in between in one step.
// Lane state of one search: q, pp, the k-list, far_mask, node, level.  They are kept as
// separate variables of the kernel (not one aggregate), so that each can live in registers.

template <typename T, int D>
__device__ __forceinline__ bool pp_is_q(const Pt<T, D>& pp, const Pt<T, D>& q) {
    bool e = true;
#pragma unroll
    for (int i = 0; i < D; ++i) e = e && (pp[i] == q[i]);
    return e;
}

// Walks the near path from `node` down to the leaf-parent level.
template <typename T, int D, bool kB>
__device__ __forceinline__ void lane_descend(const TreeView<T, kB>& tv, const Pt<T, D>& q, int32_t& node, int& level) {
    int a = axis_of_level<D>(level);
    while (level < tv.levels - 1) {
        node = 2 * node + (ax_get(q, a) < tv.split_at(node) ? 1 : 2);
        ++level;
        a = ax_next<D>(a);
    }
}

// A query with a NaN coordinate forms only NaN distances and bounds.  The reference's insertion
// (src/nndistance.hpp:77-85) rejects a candidate only when d >= worst and places it at the upper
// bound found with `value >= *it`, so a NaN distance is always taken and goes to the front; every
// bound test against worst fails, so its search (src/kdtree.cpp:16-110) scans exactly the leaf
// its descent reaches (a NaN coordinate never compares below a split) and prunes the rest.  Under
// the reference policy the list is that leaf's last k records in reverse order, with their
// distances, then the untouched (+inf, 0) entries when the leaf holds fewer than k points.  The
// lower-index policy gives what the brute force gives (src/nndistance_g.cu:57 inserts only when
// d < worst, which a NaN never is): the empty list, (+inf, original index 0) in every slot.
template <typename T, int D>
__device__ __forceinline__ bool has_nan(const Pt<T, D>& q) {
    bool r = false;
#pragma unroll
    for (int i = 0; i < D; ++i) r |= __builtin_isnan(q[i]);
    return r;
}

// Record range of the leaf the descent of q reaches.
template <typename T, int D, bool kB>
__device__ __forceinline__ void descent_leaf(const TreeView<T, kB>& tv, const Pt<T, D>& q, int32_t& beg,
                                             int32_t& end) {
    int32_t node = 0;
    int a = 0;
    for (int level = 0; level < tv.levels; ++level) {
        node = 2 * node + (ax_get(q, a) < tv.split_at(node) ? 1 : 2);
        a = ax_next<D>(a);
    }
    const int32_t leaf = node - tv.n_nodes;
    beg = tv.leaf_begin(leaf);
    end = tv.leaf_end(leaf);
}

// The list of a NaN query under the reference policy, from an initialised list: slot c (0 =
// nearest) holds record end-1-c of the descent's leaf while that is in the leaf.  Under the
// lower-index policy the initialised list (+inf, original index 0) is already the result; a list
// keyed by structured index then takes the position of original index 0 as its key.
template <typename T, int D, int K, bool kLower, bool kB, typename List>
__device__ __forceinline__ void nan_fill(const TreeView<T, kB>& tv, const Pt<T, D>& q, int k, List& kb) {
    if constexpr (!kLower) {
        int32_t beg = 0, end = 0;
        descent_leaf<T, D, kB>(tv, q, beg, end);
        const int pad = K - k;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int32_t rec = end - 1 - (j - pad);
            if (j >= pad && rec >= beg) kb.set(j, sq_dist<T, D>(q, load_rec<T, D>(tv.recs, rec).c), rec);
        }
    } else if constexpr (!List::kOrigKeys) {
        const int32_t s0 = load_i32_at(tv.inv_shuffled, 0);
#pragma unroll
        for (int j = 0; j < K; ++j) kb.set(j, kb.dist(j), s0);
    }
}

// Starts the search of the query in q.
template <typename T, int D, int K, bool kLower, bool kB, typename List>
__device__ __forceinline__ void lane_start(const TreeView<T, kB>& tv, int k, const Pt<T, D>& q, Pt<T, D>& pp,
                                           List& kb, uint32_t& far_mask, int32_t& node,
                                           int& level) {
    pp = q;
    kb.init(k);
    far_mask = 0;
    node = 0;
    level = 0;
    lane_descend<T, D>(tv, q, node, level);
}

// Scans the leaf pair of the current leaf-parent, then climbs to the next far side to enter and
// walks down to its leaf-parent.  Returns true when the search is complete.
template <typename T, int D, int K, bool kLower, int CH, int LD, int SB = kTreeBlockStride, bool kGroups = false,
          bool kB, typename List>
__device__ __forceinline__ bool lane_step(const TreeView<T, kB>& tv, const Pt<T, D>& q, Pt<T, D>& pp,
                                          List& kb, uint32_t& far_mask, int32_t& node, int& level,
                                          T* __restrict__ st_d, int32_t* __restrict__ st_o) {
    const int L = tv.levels;
    const int a_lp = axis_of_level<D>(L - 1);  // split axis of every leaf-parent
    {
        const T m = tv.split_at(node);
        const T qa = ax_get(q, a_lp);
        const bool lower = qa < m;
        const int32_t left_leaf = 2 * node + 1 - tv.n_nodes;
        const int32_t near_leaf = lower ? left_leaf : left_leaf + 1;
        const int32_t far_leaf = lower ? left_leaf + 1 : left_leaf;
        // near leaf, then the far leaf if the reference looks at it (one scan site for both)
#pragma nounroll
        for (int side = 0; side < 2; ++side) {
            const int32_t leaf = side ? far_leaf : near_leaf;
            const bool looks = side == 0 || (kLower ? lower_split_passes<T, D>(tv, q, pp, a_lp, m, kb.worst())
                                                    : proj_bound<T, D>(q, pp, a_lp, m) < kb.worst());
            if (looks && leaf_useful<T, D, kLower>(tv, q, leaf, kb.worst()))
                scan_leaf_staged<T, D, K, kLower, CH, LD, SB, kGroups>(tv, leaf, q, kb, st_d, st_o);
        }
        if (level == 0) return true;
        ax_set(pp, a_lp, qa);
    }
    int l = level;
    for (;;) {
        if (pp_is_q<T, D>(pp, q)) {
            // With pp equal to q the reference's bound for a far side is fl(m - q[a])^2, which the
            // distance to the far child's box never undercuts (the box lies beyond m on axis a), so
            // the far side is entered iff that box distance passes, and a skip leaves pp as it is.
            // The far child of a level entered through its near child is the sibling of the path's
            // child one level down, so these tests need neither the split nor the axis.
            uint32_t near_lv = ((1u << l) - 1u) & ~far_mask;
            if constexpr ((kB && !TK_BATCH_CLIMB_PREFETCH) || (D > 3 && !TK_WIDE_CLIMB_PREFETCH)) {
                while (near_lv) {
                    const int j = 31 - __builtin_clz(near_lv);
                    const int32_t c = ((node + 1) >> (level - 1 - j)) - 1;
                    const int32_t fc = ((c - 1) ^ 1) + 1;
                    const T b = aabb_bound<T, D>(tv.boxes, tv.box_of(fc), q);
                    if (kLower ? lower_passes(b, kb.worst()) : b < kb.worst()) {
                        ax_set(pp, axis_of_level<D>(j), tv.split_at((c - 1) >> 1));
                        far_mask = (far_mask & ((1u << j) - 1u)) | (1u << j);
                        node = fc;
                        level = j + 1;
                        lane_descend<T, D>(tv, q, node, level);
                        return false;
                    }
                    near_lv ^= 1u << j;
                }
                return true;
            }
            // The box of the next candidate level is loaded while the current one is tested.
            if (!near_lv) return true;
            int j = 31 - __builtin_clz(near_lv);
            int32_t c = ((node + 1) >> (level - 1 - j)) - 1;
            BoxVals<T, D> bx = load_box<T, D>(tv.boxes, tv.box_of(((c - 1) ^ 1) + 1));
            for (;;) {
                const uint32_t rest = near_lv ^ (1u << j);
                const int j2 = rest ? 31 - __builtin_clz(rest) : j;
                const int32_t c2 = ((node + 1) >> (level - 1 - j2)) - 1;
                const BoxVals<T, D> bx2 = load_box<T, D>(tv.boxes, tv.box_of(((c2 - 1) ^ 1) + 1));
                const T b = box_dist<T, D>(bx, q);
                if (kLower ? lower_passes(b, kb.worst()) : b < kb.worst()) {
                    ax_set(pp, axis_of_level<D>(j), tv.split_at((c - 1) >> 1));
                    far_mask = (far_mask & ((1u << j) - 1u)) | (1u << j);
                    node = ((c - 1) ^ 1) + 1;
                    level = j + 1;
                    lane_descend<T, D>(tv, q, node, level);
                    return false;
                }
                if (!rest) return true;
                near_lv = rest;
                j = j2;
                c = c2;
                bx = bx2;
            }
        }
        const uint32_t below = (1u << l) - 1u;
        const uint32_t near_lv = below & ~far_mask;
        const int j = near_lv ? 31 - __builtin_clz(near_lv) : -1;
        const uint32_t popped = far_mask & below & ~(j >= 0 ? ((2u << j) - 1u) : 0u);
        if (popped) {
            far_mask &= ~popped;
#pragma unroll
            for (int ax = 0; ax < D; ++ax)
                if (popped & axis_levels<D>(ax)) pp[ax] = q[ax];
        }
        if (j < 0) return true;
        const int32_t nj = ((node + 1) >> (level - j)) - 1;
        const int aj = axis_of_level<D>(j);
        const T mj = tv.split_at(nj);
        const T qj = ax_get(q, aj);
        const int32_t far_child = 2 * nj + (qj < mj ? 2 : 1);
        ax_set(pp, aj, mj);
        const int f = far_side<T, D, kLower>(tv, q, pp, aj, mj, far_child, kb.worst());
        if (f == 1) {
            far_mask |= 1u << j;
            node = far_child;
            level = j + 1;
            lane_descend<T, D>(tv, q, node, level);
            return false;
        }
        if (f == 2) {
            // the reference's pass over this side resets pp on the axes of the next
            // min(D, L-1-j) levels
            const int nlev = min(D, L - 1 - j);
#pragma unroll
            for (int i = 0; i < D; ++i) {
                const int off = (i - aj - 1 + 2 * D) % D;
                if (off < nlev) pp[i] = q[i];
            }
        }
        ax_set(pp, aj, qj);
        if (j == 0) return true;
        l = j;
    }
}

// Reference-order search with a generic list (used with the LDS list for large k): the same walk
// as lane_start/lane_step, one node at a time.  kBoxes false: no point-box tests, so the walk is
// the reference's own (src/kdtree.cpp:16-110), step for step (trees with non-finite values).
template <typename T, int D, bool kLower, bool kBoxes = true, typename List, bool kB>
__device__ __forceinline__ void knn_search_one(const TreeView<T, kB>& tv, const Pt<T, D>& q, List& kb) {
    static_assert(kBoxes || !kLower, "the lower-index policy always prunes with boxes");
    Pt<T, D> pp = q;
    const int L = tv.levels;
    const int32_t nn = tv.n_nodes;
    int32_t node = 0;
    int level = 0;
    int a = 0;  // split axis of `node`: level mod D
    for (;;) {
        while (level < L - 1) {
            node = 2 * node + (ax_get(q, a) < tv.split_at(node) ? 1 : 2);
            ++level;
            a = ax_next<D>(a);
        }
        {
            const T m = tv.split_at(node);
            const T qa = ax_get(q, a);
            const bool lower = qa < m;
            const int32_t near_leaf = 2 * node + (lower ? 1 : 2) - nn;
            const int32_t far_leaf = 2 * node + (lower ? 2 : 1) - nn;
            if (!kBoxes || leaf_useful<T, D, kLower>(tv, q, near_leaf, kb.worst()))
                scan_leaf<T, D>(tv, near_leaf, q, kb);
            if ((kLower ? lower_split_passes<T, D>(tv, q, pp, a, m, kb.worst())
                        : proj_bound<T, D>(q, pp, a, m) < kb.worst()) &&
                (!kBoxes || leaf_useful<T, D, kLower>(tv, q, far_leaf, kb.worst())))
                scan_leaf<T, D>(tv, far_leaf, q, kb);
            if (level == 0) return;
            ax_set(pp, a, qa);
        }
        bool entered = false;
        while (level > 0) {
            const int32_t child = node;
            node = (node - 1) >> 1;
            --level;
            a = ax_prev<D>(a);
            const T m2 = tv.split_at(node);
            const T q2 = ax_get(q, a);
            const bool lower2 = q2 < m2;
            if (child == 2 * node + (lower2 ? 1 : 2)) {
                const int32_t far_child = 2 * node + (lower2 ? 2 : 1);
                ax_set(pp, a, m2);
                const int f = kBoxes ? far_side<T, D, kLower>(tv, q, pp, a, m2, far_child, kb.worst())
                                     : (proj_bound<T, D>(q, pp, a, m2) < kb.worst() ? 1 : 0);
                if (f == 1) {
                    node = far_child;
                    ++level;
                    a = ax_next<D>(a);
                    entered = true;
                    break;
                }
                if (f == 2) {
                    const int nlev = min(D, L - 1 - level);
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        const int off = (i - a - 1 + 2 * D) % D;
                        if (off < nlev) pp[i] = q[i];
                    }
                }
            }
            ax_set(pp, a, q2);
        }
        if (!entered) return;
    }
}

// Stores val(j) for the slots j in [J0, J0 + N) that are live (j >= pad) to p[j - pad]; with vec
// (k == K and p 16-byte aligned) in 16-byte stores.
template <typename V, int J0, int N, typename F>
__device__ __forceinline__ void store_slots(V* __restrict__ p, bool vec, int pad, F&& val) {
    constexpr int kPer = 16 / (int)sizeof(V);
    if constexpr (N % kPer == 0) {
        if (vec) {
#pragma unroll
            for (int j = J0; j < J0 + N; j += kPer) {
                V tmp[kPer];
#pragma unroll
                for (int u = 0; u < kPer; ++u) tmp[u] = val(j + u);
                uint4 w;
                __builtin_memcpy(&w, tmp, 16);
                *reinterpret_cast<uint4*>(p + j) = w;
            }
            return;
        }
    }
#pragma unroll
    for (int j = J0; j < J0 + N; ++j)
        if (j >= pad) p[j - pad] = val(j);
}

// Writes one query's list to row qi of the outputs: distances, structured indices and (with
// out_o) original indices.  A list keyed by structured index (kOrig false) gathers the original
// ones through `shuffled`; one keyed by original index (the lower-index policy's 64-bit keyed
// list) gathers the structured ones through `inv_shuffled`.  The gathers of up to 16 slots are
// issued before their stores, so their latencies overlap instead of following one another (a
// 32-entry list goes out in two halves, which keeps the live registers below those of its
// search); when k == K and the three arrays are 16-byte aligned, the rows go out in 16-byte
// stores.
template <typename T, int K, bool kOrig, int J0 = 0, typename List>
__device__ __forceinline__ void store_list(const List& kb, int k, const TreeView<T>& tv, int64_t qi,
                                           T* __restrict__ out_d, int32_t* __restrict__ out_i,
                                           int64_t* __restrict__ out_o) {
    constexpr int N = K - J0 < 16 ? K - J0 : 16;
    const int pad = K - k;
    const bool vec = pad == 0 && ((((uintptr_t)out_d | (uintptr_t)out_i | (uintptr_t)out_o) & 15u) == 0);
    int32_t m[N] = {};
    if (kOrig || out_o) {
        const int32_t* map = kOrig ? tv.inv_shuffled : tv.shuffled;
#pragma unroll
        for (int u = 0; u < N; ++u) m[u] = load_i32_at(map, J0 + u >= pad ? kb.key(J0 + u) : 0);
    }
    T* od = out_d + qi * k;
    int32_t* oi = out_i + qi * k;
    store_slots<T, J0, N>(od, vec, pad, [&](int j) { return kb.dist(j); });
    store_slots<int32_t, J0, N>(oi, vec, pad, [&](int j) { return kOrig ? m[j - J0] : kb.key(j); });
    if (out_o)
        store_slots<int64_t, J0, N>(out_o + qi * k, vec, pad,
                                    [&](int j) { return (int64_t)(kOrig ? kb.key(j) : m[j - J0]); });
    if constexpr (J0 + N < K) store_list<T, K, kOrig, J0 + N>(kb, k, tv, qi, out_d, out_i, out_o);
}

// ---------------------------------------------------------------------------------------------
// Row stores through LDS for a full wavefront.  Lane l holds output row qi(l); stored lane by lane,
// every 16-byte store instruction touches 64 different rows (64 partial lines).  Passing a slab of
// SW bytes of each row through the wave's part of the leaf-scan staging buffer lets one
// instruction write 64 consecutive 16-byte chunks of the slab matrix instead: SW/16 lanes per row
// and 1024/SW whole row slabs per instruction.
// ---------------------------------------------------------------------------------------------

// The wave's part of a block's staging buffer T[CH * BLK] (lane i of the block owns elements i,
// i + BLK, i + 2*BLK, ...): CH segments of 64 elements, BLK elements apart.  Chunk c (16 bytes) of
// the wave's logical buffer lies in one segment, since a segment is a multiple of 16 bytes.
template <typename T, int BLK>
struct WaveStage {
    unsigned char* base;  // the wave's part of segment 0

    __device__ __forceinline__ uint4* chunk(int c) const {
        constexpr int kSeg = 64 * (int)sizeof(T);
        const int byte = c * 16;
        return reinterpret_cast<uint4*>(base + (byte / kSeg) * (BLK * (int)sizeof(T)) + byte % kSeg);
    }
};

// Orders the wave's LDS writes before its reads and the reads before the next writes.  A
// wavefront's LDS instructions execute in issue order, so only the compiler must keep them there.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Writes bytes [col, col + 16*NC) of every lane's row: v holds this lane's, row_bytes is the row
// stride of `base`, qr[p] the row of the lanes whose chunk pass p stores (row p*64/C + lane/C of
// the wave, C = SW/16 chunks per lane and slab).
template <typename T, int BLK, int SW, int NC>
__device__ __forceinline__ void wave_store_part(const WaveStage<T, BLK>& st, int lane, const int32_t (&qr)[SW / 16],
                                                unsigned char* __restrict__ base, int64_t row_bytes, int col,
                                                const uint4 (&v)[NC]) {
    constexpr int C = SW / 16;
    static_assert(NC % C == 0, "a row part is a whole number of slabs");
#pragma unroll
    for (int s = 0; s < NC / C; ++s) {
#pragma unroll
        for (int u = 0; u < C; ++u) *st.chunk(lane * C + u) = v[s * C + u];
        wave_lds_sync();
#pragma unroll
        for (int p = 0; p < C; ++p) {
            const uint4 w = *st.chunk(p * 64 + lane);
            *reinterpret_cast<uint4*>(base + qr[p] * row_bytes + col + s * SW + (lane % C) * 16) = w;
        }
        wave_lds_sync();
    }
}

// Slab width of the wave stores: at most 64 bytes, what the wave's staging part holds per lane
// (CH elements), and a divisor of every row part (up to 16 entries of 4, 8 or sizeof(T) bytes).
template <typename T, int K, int CH>
constexpr int wave_slab_bytes() {
    const int a = CH * (int)sizeof(T) < 64 ? CH * (int)sizeof(T) : 64;
    const int n4 = (K < 16 ? K : 16) * 4;
    return a < n4 ? a : n4;
}

// store_list for a full wavefront whose rows are k == K entries long and 16-byte aligned: the same
// values, through wave_store_part.  The distances go out before the index gathers are issued, so
// the two are not live at once.
template <typename T, int K, bool kOrig, int BLK, int SW, int J0 = 0, typename List>
__device__ __forceinline__ void store_list_wave(const List& kb, const TreeView<T>& tv,
                                                const WaveStage<T, BLK>& st, int lane,
                                                const int32_t (&qr)[SW / 16], T* __restrict__ out_d,
                                                int32_t* __restrict__ out_i, int64_t* __restrict__ out_o) {
    constexpr int N = K - J0 < 16 ? K - J0 : 16;
    {
        T tmp[N];
#pragma unroll
        for (int u = 0; u < N; ++u) tmp[u] = kb.dist(J0 + u);
        uint4 v[N * sizeof(T) / 16];
        __builtin_memcpy(v, tmp, sizeof(tmp));
        wave_store_part<T, BLK, SW>(st, lane, qr, reinterpret_cast<unsigned char*>(out_d), (int64_t)K * sizeof(T),
                                    J0 * (int)sizeof(T), v);
    }
    int32_t m[N] = {};
    if (kOrig || out_o) {
        const int32_t* map = kOrig ? tv.inv_shuffled : tv.shuffled;
#pragma unroll
        for (int u = 0; u < N; ++u) m[u] = load_i32_at(map, kb.key(J0 + u));
    }
    {
        int32_t tmp[N];
#pragma unroll
        for (int u = 0; u < N; ++u) tmp[u] = kOrig ? m[u] : kb.key(J0 + u);
        uint4 v[N * 4 / 16];
        __builtin_memcpy(v, tmp, sizeof(tmp));
        wave_store_part<T, BLK, SW>(st, lane, qr, reinterpret_cast<unsigned char*>(out_i), (int64_t)K * 4, J0 * 4, v);
    }
    if (out_o) {
        int64_t tmp[N];
#pragma unroll
        for (int u = 0; u < N; ++u) tmp[u] = (int64_t)(kOrig ? kb.key(J0 + u) : m[u]);
        uint4 v[N * 8 / 16];
        __builtin_memcpy(v, tmp, sizeof(tmp));
        wave_store_part<T, BLK, SW>(st, lane, qr, reinterpret_cast<unsigned char*>(out_o), (int64_t)K * 8, J0 * 8, v);
    }
    if constexpr (J0 + N < K)
        store_list_wave<T, K, kOrig, BLK, SW, J0 + N>(kb, tv, st, lane, qr, out_d, out_i, out_o);
}

template <typename T, int D>
__device__ __forceinline__ void load_query(const T* __restrict__ queries, int64_t qi, Pt<T, D>& q) {
#pragma unroll
    for (int i = 0; i < D; ++i) q[i] = queries[qi * D + i];
}

// Output rows of full wavefronts written through LDS (1) or lane by lane (0).
#ifndef TK_WAVE_STORES
#define TK_WAVE_STORES 1
#endif

// Points per staged chunk of the leaf scan (measured on MI355X), for float searches with d <= 3
// and lists of 8 or more: TK_STAGE_CHUNK_F32_K16 (16) for the single-tree reference-policy
// search with a 16-entry list (plain workload 8.32 ms against 8.44 at 8), TK_STAGE_CHUNK_F32 (8)
// for the other reference-policy ones (k = 8: 6.14 ms against 6.20 at 16; batched replay 21.76 ms
// against 21.93); the lower-index policy's KBestTie lists take the same, its 64-bit keyed lists
// TK_STAGE_CHUNK_F32_LOWER (4; 10.14 ms against 10.56 at 8).  For lists of at most 4 at d <= 3, TK_STAGE_CHUNK_SMALLK (8) in the single-tree search
// (k = 4: 4.78 ms against 4.91 at 4; k = 1: 3.75 against 3.80) and TK_STAGE_CHUNK_WIDE (4) in the
// batched one (replay 20.43 ms against 20.74 at 8).  TK_STAGE_CHUNK_WIDE (4) also for d = 8, where a
// larger chunk's extra stale-worst pops cost more than its fewer exits; TK_STAGE_CHUNK_F64 for double.
// kB: the batched search.
template <typename T, int D, int K, bool kLower, bool kB = false, bool kTieOK = true>
struct StageBytes {
    static constexpr bool kTieList = !kB && TreeList<T, K, kLower, kTieOK>::kTie;
    static constexpr int kChunk =
        sizeof(T) == 4
            ? ((D <= 3 && K >= 8) ? (kLower && !kTieList ? TK_STAGE_CHUNK_F32_LOWER
                                                         : (K == 16 && !kB ? TK_STAGE_CHUNK_F32_K16 : TK_STAGE_CHUNK_F32))
                                  : (D <= 3 && !kB ? TK_STAGE_CHUNK_SMALLK : TK_STAGE_CHUNK_WIDE))
            : TK_STAGE_CHUNK_F64;
    static constexpr int kLoads = sizeof(T) == 4 ? (TK_SCAN_LOADS_F32 < kChunk ? TK_SCAN_LOADS_F32 : kChunk)
                                                 : (TK_SCAN_LOADS_F64 < kChunk ? TK_SCAN_LOADS_F64 : kChunk);
    static constexpr int kBytes = kChunk * kTreeBlock * ((int)sizeof(T) + (kLower ? 4 : 0));
};

// The list interface of scan_leaf / knn_search_one over a register list.
template <typename T, typename L>
struct RegListRef {
    L& l;
    __device__ __forceinline__ T worst() const { return l.worst(); }
    __device__ __forceinline__ bool accepts(T dist, int32_t orig) const { return l.accepts(dist, orig); }
    __device__ __forceinline__ void insert(T dist, int32_t, int32_t orig) { l.insert(dist, orig); }
};

// The lower-index search of q again, node by node with the 64-bit keyed list (little code beside
// the staged search, for the rare unsure lane), into kb, with the original indices turned into
// structured ones.
template <typename T, int D, int K, bool kB>
__device__ __forceinline__ void research_lower(const TreeView<T, kB>& tv, const Pt<T, D>& q, int k,
                                               KBestTie<T, K>& kb) {
    KBest<T, K, true> ex;
    ex.init(k);
    RegListRef<T, KBest<T, K, true>> ref{ex};
    knn_search_one<T, D, true>(tv, q, ref);
#pragma unroll
    for (int j = 0; j < K; ++j) kb.set(j, ex.dist(j), load_i32_at(tv.inv_shuffled, ex.key(j)));
}

// One lane per query.  The lanes of a wavefront take 64 consecutive presorted queries and run
// their searches to completion together: neighbouring queries walk the tree nearly in step, which
// keeps the wavefront's control flow and loads coherent.  (Handing a finished lane the next query
// was measured to be slower: it breaks that lockstep.)  BLK threads per block, leaf chunks of CH
// records with LD loads in flight.
template <typename T, int D, int K, bool kLower, int CH, int LD, int BLK, bool kGroups = false, bool kTieOK = true>
__device__ __forceinline__ void tree_search_lanes(const TreeView<T>& tv, const T* __restrict__ queries, int64_t nq,
                                                  const int32_t* __restrict__ order, int k, T* __restrict__ out_d,
                                                  int32_t* __restrict__ out_i, int64_t* __restrict__ out_o) {
    static_assert(CH <= kRecPad, "a chunk may read at most kRecPad records past a leaf's end");
    static_assert(BLK % 64 == 0, "whole wavefronts per block");
    using List = typename TreeList<T, K, kLower, kTieOK>::type;
    constexpr bool kO = List::kOrigKeys;  // list keys are original indices (staged with the distances)
    constexpr bool kTie = TreeList<T, K, kLower, kTieOK>::kTie;
    __shared__ __align__(16) T stage_d[CH * BLK];
    __shared__ int32_t stage_o[kO ? CH * BLK : 1];
    // t = wave0 + lane with wave0 in scalars; the lane index and the row are found again after the
    // search, so that no per-lane index has to stay live through it
    const int wave_off = __builtin_amdgcn_readfirstlane((int)(threadIdx.x & ~63u));
    const int64_t wave0 = xcd_tile(blockIdx.x, gridDim.x) * BLK + wave_off;
    if (wave0 + (int)(threadIdx.x & 63u) >= nq) return;
    Pt<T, D> q, pp;
    List kb;
    uint32_t far_mask;
    int32_t node;
    int level;
    {
        const int64_t t = wave0 + (int)(threadIdx.x & 63u);
        load_query<T, D>(queries, order ? (int64_t)order[t] : t, q);
    }
    lane_start<T, D, K, kLower>(tv, k, q, pp, kb, far_mask, node, level);
    if (has_nan<T, D>(q)) {
        nan_fill<T, D, K, kLower>(tv, q, k, kb);
    } else {
        // (a KBestTie lane stops after the leaf pair where it becomes unsure: it is searched again
        // in any case)
        while (!lane_step<T, D, K, kLower, CH, LD, BLK, kGroups>(tv, q, pp, kb, far_mask, node, level,
                                                              stage_d + threadIdx.x, stage_o + (kO ? threadIdx.x : 0))) {
            if constexpr (kTie)
                if (kb.unsure_early(k)) break;
        }
        if constexpr (kTie)
            if (kb.unsure(k)) research_lower<T, D, K>(tv, q, k, kb);
    }
    // (the lane index from mbcnt in a volatile asm, and the row by an atomic load, so that the
    // compiler recomputes both here rather than carrying the first values through the search)
    int lane;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane));
    const int32_t qi = order ? __hip_atomic_load(order + wave0 + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT)
                             : (int32_t)(wave0 + lane);  // nq < 2^31
    constexpr int SW = wave_slab_bytes<T, K, CH>();
    // (float only: the double kernels already run at 2 to 5 waves per SIMD, where the stores' extra
    // registers cost a wave for some lists)
    if constexpr (TK_WAVE_STORES && sizeof(T) == 4 && K >= 8 && SW >= 32) {
        // When the wave's 64 queries all exist, every lane of the wave is here (the search loop has
        // ended for all of them) and the wave's rows go out through its part of stage_d.  The wave's
        // first position comes from scalars, so that t need not stay live through the search.
        const bool full = wave0 + 64 <= nq;
        const bool vec = k == K && ((((uintptr_t)out_d | (uintptr_t)out_i | (uintptr_t)out_o) & 15u) == 0);
        if (full && vec) {
            constexpr int C = SW / 16;
            int32_t qr[C];
#pragma unroll
            for (int p = 0; p < C; ++p) qr[p] = __shfl(qi, p * (64 / C) + lane / C, 64);
            const WaveStage<T, BLK> st{reinterpret_cast<unsigned char*>(stage_d + wave_off)};
            store_list_wave<T, K, kO, BLK, SW>(kb, tv, st, lane, qr, out_d, out_i, out_o);
            return;
        }
    }
    store_list<T, K, kO>(kb, k, tv, qi, out_d, out_i, out_o);
}

// kGroups: the tree has record-group boxes (tv.group_boxes), and leaf scans test them.  kTieOK
// false: a lower-index search keeps the 64-bit keyed list (a tree whose cloud ties often).
template <typename T, int D, int K, bool kLower, bool kGroups = false, bool kTieOK = true>
__global__ __launch_bounds__(kTreeBlock) TK_TREE_OCCUPANCY(T, D, K, kLower, (TreeList<T, K, kLower, kTieOK>::kTie))
void knn_tree_kernel(TreeView<T> tv, const T* __restrict__ queries, int64_t nq, const int32_t* __restrict__ order,
                     int k, T* __restrict__ out_d, int32_t* __restrict__ out_i, int64_t* __restrict__ out_o) {
    tree_search_lanes<T, D, K, kLower, StageBytes<T, D, K, kLower, false, kTieOK>::kChunk,
                      StageBytes<T, D, K, kLower, false, kTieOK>::kLoads, kTreeBlock, kGroups, kTieOK>(
        tv, queries, nq, order, k, out_d, out_i, out_o);
}

// The same search with its launch shape as template parameters: the prebuilt variants an autotuner
// picks among (leaf chunk CH with LD loads in flight, BLK threads per block, an occupancy target of
// WAVES waves per SIMD, 0 = none).  Results do not depend on them.
template <typename T, int D, int K, bool kLower, int CH, int LD, int BLK, int WAVES>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1))) void knn_tree_kernel_v(
    TreeView<T> tv, const T* __restrict__ queries, int64_t nq, const int32_t* __restrict__ order, int k,
    T* __restrict__ out_d, int32_t* __restrict__ out_i, int64_t* __restrict__ out_o) {
    tree_search_lanes<T, D, K, kLower, CH, LD, BLK>(tv, queries, nq, order, k, out_d, out_i, out_o);
}

constexpr int kLdsBlock = 64;

template <typename T, int D, bool kLower>
__global__ __launch_bounds__(kLdsBlock) void knn_tree_lds_kernel(TreeView<T> tv, const T* __restrict__ queries,
                                                                  int64_t nq, const int32_t* __restrict__ order,
                                                                  int k, T* __restrict__ out_d,
                                                                  int32_t* __restrict__ out_i,
                                                                  int64_t* __restrict__ out_o) {
    extern __shared__ unsigned char smem[];
    const int64_t t = xcd_tile(blockIdx.x, gridDim.x) * kLdsBlock + threadIdx.x;
    if (t >= nq) return;
    const int64_t qi = order ? (int64_t)order[t] : t;
    T* sd = reinterpret_cast<T*>(smem);
    int32_t* ss = reinterpret_cast<int32_t*>(sd + (size_t)k * kLdsBlock);
    int32_t* so = ss + (size_t)k * kLdsBlock;
    KBestLds<T, kLower> kb{sd + threadIdx.x, ss + threadIdx.x, so + threadIdx.x, k, kLdsBlock, T(0), 0};
    kb.init(kLower ? tv.inv_shuffled[0] : 0);
    Pt<T, D> q;
    load_query<T, D>(queries, qi, q);
    if (has_nan<T, D>(q)) {
        // the reference policy's NaN result (nan_fill); the lower-index list stays empty
        if constexpr (!kLower) {
            int32_t beg = 0, end = 0;
            descent_leaf<T, D>(tv, q, beg, end);
            for (int c = 0; c < k && end - 1 - c >= beg; ++c) {
                kb.d[c * kLdsBlock] = sq_dist<T, D>(q, load_rec<T, D>(tv.recs, end - 1 - c).c);
                kb.s[c * kLdsBlock] = end - 1 - c;
            }
        }
    } else {
        knn_search_one<T, D, kLower>(tv, q, kb);
    }
    for (int j = 0; j < k; ++j) {
        out_d[qi * k + j] = kb.d[j * kLdsBlock];
        out_i[qi * k + j] = kb.s[j * kLdsBlock];
        if (out_o) out_o[qi * k + j] = kLower ? kb.o[j * kLdsBlock] : tv.shuffled[kb.s[j * kLdsBlock]];
    }
}

// Reference tie policy on a tree with a non-finite coordinate or split value: the reference's walk
// step for step (no point boxes) with the reference's list (KBestExact), one lane per query.
// Queries with a NaN coordinate need no special case here.
template <typename T, int D>
__global__ __launch_bounds__(kLdsBlock) void knn_tree_exact_kernel(TreeView<T> tv, const T* __restrict__ queries,
                                                                    int64_t nq, const int32_t* __restrict__ order,
                                                                    int k, T* __restrict__ out_d,
                                                                    int32_t* __restrict__ out_i,
                                                                    int64_t* __restrict__ out_o) {
    extern __shared__ unsigned char smem[];
    const int64_t t = xcd_tile(blockIdx.x, gridDim.x) * kLdsBlock + threadIdx.x;
    if (t >= nq) return;
    const int64_t qi = order ? (int64_t)order[t] : t;
    T* sd = reinterpret_cast<T*>(smem);
    int32_t* ss = reinterpret_cast<int32_t*>(sd + (size_t)k * kLdsBlock);
    KBestExact<T> kb{sd + threadIdx.x, ss + threadIdx.x, k, kLdsBlock};
    kb.init();
    Pt<T, D> q;
    load_query<T, D>(queries, qi, q);
    knn_search_one<T, D, false, false>(tv, q, kb);
    for (int j = 0; j < k; ++j) {
        const int32_t sj = kb.s[j * kLdsBlock];
        out_d[qi * k + j] = kb.d[j * kLdsBlock];
        out_i[qi * k + j] = sj;
        if (out_o) out_o[qi * k + j] = tv.shuffled[sj];
    }
}

inline size_t exact_list_bytes(int k, size_t tsize) { return (size_t)k * kLdsBlock * (tsize + 4); }

inline size_t lds_list_bytes(int k, size_t tsize, bool lower) {
    return (size_t)k * kLdsBlock * (tsize + 4 + (lower ? 4 : 0));
}

// ---------------------------------------------------------------------------------------------
// Query presorts.  Each presort kernel writes one integer key per query and the query's row; a
// stable radix sort of (key, row) (radix_sort.hip) then gives the order, so the order is
// deterministic: equal keys keep increasing rows.
//
// Leaf keys: each query descends to the leaf on its near side (no pruning); leaves are numbered
// left to right, so neighbouring keys are neighbouring cells of space.
// ---------------------------------------------------------------------------------------------
constexpr int kPresortBlock = 256;

template <typename T, int D>
__global__ __launch_bounds__(kPresortBlock) void leaf_bucket_kernel(TreeView<T> tv, const T* __restrict__ queries,
                                                                     int64_t nq, int shift,
                                                                     uint32_t* __restrict__ key,
                                                                     uint32_t* __restrict__ row) {
    const int64_t i = (int64_t)blockIdx.x * kPresortBlock + threadIdx.x;
    if (i >= nq) return;
    Pt<T, D> q;
    load_query<T, D>(queries, i, q);
    int32_t node = 0;
    int a = 0;
    for (int level = 0; level < tv.levels; ++level) {
        node = 2 * node + (ax_get(q, a) < tv.split_at(node) ? 1 : 2);
        a = ax_next<D>(a);
    }
    key[i] = (uint32_t)((node - tv.n_nodes) >> shift);
    row[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------------------------
// Morton keys: the Morton code of a query's cell in a 2^B-per-axis grid over a box (B*D <= 24).
// The forward presort uses the queries' own bounding box and the finest grid, so queries
// outside the tree's cloud are grouped by where they really are.  The backward's canonical
// order uses the reference points' bounding box and B from canonical_morton_bits(Q, D)
// (knn_gpu.hpp), the same rule in the single-tree and the batched backward.
// ---------------------------------------------------------------------------------------------
// Bits per axis of the forward presort's grid (measured on MI355X): d = 3, 8 bits (a 24-bit key,
// still three radix passes; plain workload traversal 7.92 ms against 8.23 with 7 bits, depth 13
// 20.29 against 21.33); d = 8, 3 bits (24-bit key; 90.4 ms per 500K queries against 94.0 with 2).
#ifndef TK_MORTON_BITS_3D
#define TK_MORTON_BITS_3D 8
#endif
#ifndef TK_MORTON_BITS_8D
#define TK_MORTON_BITS_8D 3
#endif
template <int D>
struct MortonBits {
    static constexpr int kBits = D == 1 ? 20 : D == 2 ? 10 : D == 3 ? TK_MORTON_BITS_3D : TK_MORTON_BITS_8D;
    static constexpr int kBuckets = 1 << (kBits * D);
};

// Morton code of q's cell; lo/hi are the box, converted to float.  Queries outside the box fall
// into the boundary cells.
template <typename T, int D>
__device__ __forceinline__ uint32_t morton_cell(const float* lo, const float* hi, int bits, const T* __restrict__ q) {
    const float cells = (float)(1 << bits);
    uint32_t code = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const float scale = hi[a] > lo[a] ? cells / (hi[a] - lo[a]) : 0.0f;
        const float x = ((float)q[a] - lo[a]) * scale;
        const uint32_t g = (uint32_t)fminf(fmaxf(x, 0.0f), cells - 1.0f);
        for (int b = 0; b < bits; ++b) code |= ((g >> b) & 1u) << (b * D + a);
    }
    return code;
}

// Hilbert index of q's cell in the same grid as morton_cell: the cell's integer coordinates go
// through Skilling's transpose transform (J. Skilling, "Programming the Hilbert curve", AIP Conf.
// Proc. 707, 2004: undo the excess work, then Gray-encode), and the transposed coordinates are
// interleaved most significant bit first.  Consecutive indices are adjacent cells, which the
// Z-order's jumps between blocks of cells are not.
template <typename T, int D>
__device__ __forceinline__ uint32_t hilbert_cell(const float* lo, const float* hi, int bits, const T* __restrict__ q) {
    const float cells = (float)(1 << bits);
    uint32_t x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const float scale = hi[a] > lo[a] ? cells / (hi[a] - lo[a]) : 0.0f;
        const float v = ((float)q[a] - lo[a]) * scale;
        x[a] = (uint32_t)fminf(fmaxf(v, 0.0f), cells - 1.0f);
    }
    if (bits > 0) {
        const uint32_t m = 1u << (bits - 1);
        for (uint32_t qb = m; qb > 1; qb >>= 1) {
            const uint32_t p = qb - 1;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                if (x[a] & qb) {
                    x[0] ^= p;
                } else {
                    const uint32_t t = (x[0] ^ x[a]) & p;
                    x[0] ^= t;
                    x[a] ^= t;
                }
            }
        }
#pragma unroll
        for (int a = 1; a < D; ++a) x[a] ^= x[a - 1];
        uint32_t t = 0;
        for (uint32_t qb = m; qb > 1; qb >>= 1)
            if (x[D - 1] & qb) t ^= qb - 1;
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] ^= t;
    }
    uint32_t code = 0;
    for (int b = bits - 1; b >= 0; --b) {
#pragma unroll
        for (int a = 0; a < D; ++a) code = (code << 1) | ((x[a] >> b) & 1u);
    }
    return code;
}

// Order-preserving map of floating-point values onto unsigned integers (for atomicMin/Max).
__device__ __forceinline__ uint32_t ordered_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ unsigned long long ordered_bits(double f) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(f);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}
__device__ __forceinline__ double from_ordered(unsigned long long u) {
    return __longlong_as_double((long long)((u >> 63) ? (u & ~(1ull << 63)) : ~u));
}

template <typename T>
using OrderedT = typename std::conditional<sizeof(T) == 4, uint32_t, unsigned long long>::type;

constexpr int kBoxBlock = 256;

// Bounding box of n points: box[0..D) = min, box[D..2D) = max, as ordered integers; the caller
// initialises it to (all ones, zero).  Reduced per wavefront, then per block in LDS, then one
// atomic per block and value.
template <typename T, int D>
__global__ __launch_bounds__(kBoxBlock) void bbox_kernel(const T* __restrict__ pts, int64_t n,
                                                         OrderedT<T>* __restrict__ box) {
    constexpr int kWaves = kBoxBlock / 64;
    __shared__ T s_lo[kWaves][D], s_hi[kWaves][D];
    T lo[D], hi[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        lo[a] = pos_inf<T>();
        hi[a] = -pos_inf<T>();
    }
    for (int64_t i = (int64_t)blockIdx.x * kBoxBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBoxBlock) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const T v = pts[i * D + a];
            lo[a] = fmin(lo[a], v);
            hi[a] = fmax(hi[a], v);
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            lo[a] = fmin(lo[a], __shfl_xor(lo[a], off, 64));
            hi[a] = fmax(hi[a], __shfl_xor(hi[a], off, 64));
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
            s_lo[w][a] = lo[a];
            s_hi[w][a] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < D) {
        const int a = threadIdx.x;
        T l = s_lo[0][a], h = s_hi[0][a];
        for (int v = 1; v < kWaves; ++v) {
            l = fmin(l, s_lo[v][a]);
            h = fmax(h, s_hi[v][a]);
        }
        atomicMin(&box[a], ordered_bits(l));
        atomicMax(&box[D + a], ordered_bits(h));
    }
}

// Cell key over a box held as ordered integers (from bbox_kernel): the Hilbert index when hilbert
// (the forward presort at d <= 3: plain workload traversal 7.78 ms against 7.88 with the Morton
// code, fp64 workload 1.41 against 1.49; at d = 8 the Morton code is faster, 90.6 against 92.2 ms),
// else the Morton code (always for the backward's canonical order).
template <typename T, int D>
__global__ __launch_bounds__(kPresortBlock) void morton_bucket_kernel(const T* __restrict__ queries, int64_t nq,
                                                                       const OrderedT<T>* __restrict__ box, int bits,
                                                                       uint32_t* __restrict__ key,
                                                                       uint32_t* __restrict__ row,
                                                                       bool hilbert) {
    const int64_t i = (int64_t)blockIdx.x * kPresortBlock + threadIdx.x;
    if (i >= nq) return;
    float lo[D], hi[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        lo[a] = (float)from_ordered(box[a]);
        hi[a] = (float)from_ordered(box[D + a]);
    }
    key[i] = hilbert ? hilbert_cell<T, D>(lo, hi, bits, queries + i * D) : morton_cell<T, D>(lo, hi, bits, queries + i * D);
    row[i] = (uint32_t)i;
}

}  // namespace tk
