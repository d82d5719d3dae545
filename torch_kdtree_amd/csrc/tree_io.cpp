// Tree serialization: the arrays of a built tree (the reference's read-only getters,
// src/interface.cpp:45-58, plus the node and leaf arrays), so that a static cloud's tree is built
// once and loaded afterwards.
//
// Format, little endian:
//   "TKDTREE\0"  magic, 8 bytes
//   u32 format version (1), u32 dtype (0 float32, 1 float64), u32 dims, u32 levels, i64 n_points
//   T[2^levels - 1]      split values, heap order
//   i32[2^levels + 1]    leaf offsets
//   T[n_points * dims]   points in structured order
//   i32[n_points]        original index of each structured position
//   u64                  FNV-1a 64 of every byte before it
// Loading checks the checksum and the structure (offsets monotone from 0 to n_points, each leaf
// as large as the build makes it, a permutation of indices) before the tree is used; the axes and
// point boxes are recomputed.
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tree.hpp"

namespace tk {
namespace {

constexpr char kMagic[8] = {'T', 'K', 'D', 'T', 'R', 'E', 'E', '\0'};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeader = 8 + 4 * 4 + 8;

uint64_t fnv1a(const unsigned char* p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

struct Writer {
    std::vector<unsigned char>& out;
    void put(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        out.insert(out.end(), b, b + n);
    }
    template <typename V>
    void put_v(V v) {
        put(&v, sizeof(V));
    }
};

struct Reader {
    const unsigned char* p;
    size_t n, pos = 0;
    void get(void* dst, size_t bytes) {
        if (pos + bytes > n) throw std::runtime_error("tree file: truncated");
        std::memcpy(dst, p + pos, bytes);
        pos += bytes;
    }
    template <typename V>
    V get_v() {
        V v;
        get(&v, sizeof(V));
        return v;
    }
};

}  // namespace

std::vector<unsigned char> serialize_tree(const Tree& t) {
    std::vector<unsigned char> out;
    const size_t ts = dtype_size(t.dtype);
    out.reserve(kHeader + t.split.size() + 4 * t.leaf_off.size() + t.points.size() + 4 * t.shuffled.size() + 8);
    Writer w{out};
    w.put(kMagic, 8);
    w.put_v<uint32_t>(kVersion);
    w.put_v<uint32_t>((uint32_t)t.dtype);
    w.put_v<uint32_t>((uint32_t)t.dims);
    w.put_v<uint32_t>((uint32_t)t.levels);
    w.put_v<int64_t>((int64_t)t.n_points);
    w.put(t.split.data(), ts * (size_t)t.n_nodes);
    w.put(t.leaf_off.data(), 4 * t.leaf_off.size());
    w.put(t.points.data(), t.points.size());
    w.put(t.shuffled.data(), 4 * t.shuffled.size());
    w.put_v<uint64_t>(fnv1a(out.data(), out.size()));
    return out;
}

Tree* deserialize_tree(const unsigned char* buf, size_t n) {
    if (n < kHeader + 8) throw std::runtime_error("tree file: truncated");
    uint64_t stored;
    std::memcpy(&stored, buf + n - 8, 8);
    if (fnv1a(buf, n - 8) != stored) throw std::runtime_error("tree file: checksum mismatch");
    Reader r{buf, n - 8};
    char magic[8];
    r.get(magic, 8);
    if (std::memcmp(magic, kMagic, 8) != 0) throw std::runtime_error("tree file: not a tree file");
    if (r.get_v<uint32_t>() != kVersion) throw std::runtime_error("tree file: unsupported format version");
    const uint32_t dtype = r.get_v<uint32_t>();
    const uint32_t dims = r.get_v<uint32_t>();
    const uint32_t levels = r.get_v<uint32_t>();
    const int64_t n_points = r.get_v<int64_t>();
    if (dtype != kF32 && dtype != kF64) throw std::runtime_error("tree file: unknown dtype");
    if (!dims_supported((int)dims)) throw std::runtime_error("tree file: unsupported dims");
    if (levels < 1 || (int)levels > kMaxLevels) throw std::runtime_error("tree file: levels out of range");
    if (n_points < (int64_t(1) << (levels - 1)) || n_points >= (int64_t(1) << 31))
        throw std::runtime_error("tree file: point count does not fit the levels");
    std::unique_ptr<Tree> t(new Tree());
    t->dtype = (int)dtype;
    t->dims = (int)dims;
    t->levels = (int)levels;
    t->n_points = (int32_t)n_points;
    t->n_nodes = total_nodes((int)levels);
    t->n_leaves = (int32_t)(1u << levels);
    const size_t ts = dtype_size(t->dtype);
    t->split.resize(ts * (size_t)t->n_nodes);
    t->leaf_off.resize((size_t)t->n_leaves + 1);
    t->points.resize(ts * (size_t)dims * (size_t)n_points);
    t->shuffled.resize((size_t)n_points);
    r.get(t->split.data(), t->split.size());
    r.get(t->leaf_off.data(), 4 * t->leaf_off.size());
    r.get(t->points.data(), t->points.size());
    r.get(t->shuffled.data(), 4 * t->shuffled.size());
    if (r.pos != r.n) throw std::runtime_error("tree file: trailing bytes");
    // structure: leaf sizes as the median split makes them, offsets from 0 to n, a permutation
    if (t->leaf_off[0] != 0 || t->leaf_off[t->n_leaves] != t->n_points)
        throw std::runtime_error("tree file: bad leaf offsets");
    const int64_t lo_size = n_points >> levels, hi_size = (n_points + t->n_leaves - 1) >> levels;
    for (int32_t j = 0; j < t->n_leaves; ++j) {
        const int64_t sz = (int64_t)t->leaf_off[j + 1] - t->leaf_off[j];
        if (sz < lo_size || sz > hi_size) throw std::runtime_error("tree file: bad leaf offsets");
    }
    std::vector<unsigned char> seen((size_t)n_points, 0);
    for (int32_t v : t->shuffled) {
        if (v < 0 || v >= n_points || seen[(size_t)v]) throw std::runtime_error("tree file: indices are not a permutation");
        seen[(size_t)v] = 1;
    }
    finish_tree(*t);
    return t.release();
}

void save_tree(const Tree& t, const char* path) {
    const std::vector<unsigned char> bytes = serialize_tree(t);
    std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "wb"), std::fclose);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path + " for writing");
    if (std::fwrite(bytes.data(), 1, bytes.size(), f.get()) != bytes.size())
        throw std::runtime_error(std::string("cannot write ") + path);
    if (std::fflush(f.get()) != 0) throw std::runtime_error(std::string("cannot write ") + path);
}

Tree* load_tree(const char* path) {
    std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "rb"), std::fclose);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    std::vector<unsigned char> bytes;
    unsigned char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f.get())) > 0) bytes.insert(bytes.end(), chunk, chunk + got);
    if (std::ferror(f.get())) throw std::runtime_error(std::string("cannot read ") + path);
    return deserialize_tree(bytes.data(), bytes.size());
}

}  // namespace tk
