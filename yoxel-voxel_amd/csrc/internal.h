// Internal declarations shared by the host C++ and the HIP translation units of
// libyoxel. Nothing here is part of the C ABI (include/yoxel.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "yoxel.h"

// Host copy of an octree in the reference's VoxNode layout
// (reaction/report/main.tex:36-56): 10 u32 per node, canonical ids are indices.
struct yx_scene {
    std::vector<uint32_t> nodes;
    uint32_t root = YX_EMPTY_NODE;
    uint64_t num_leaves = 0;
    int depth = 0;
};

namespace yx {

// Deepest supported leaf level (root children are level 1). Bounds the
// per-ray traversal stack.
constexpr int kMaxDepth = 16;

int fail(int code, const char* fmt, ...);
void clear_error();

// Worker threads for host-side builds: min(hardware threads, OMP_NUM_THREADS /
// YX_THREADS if set, 16).
int host_threads();

// Splitmix64 (Steele, Lea & Flood 2014): seeding and hashing.
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        return mix64(s);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// Improved gradient noise (Perlin, SIGGRAPH 2002) with a seeded permutation.
struct Perlin {
    int perm[512];
    explicit Perlin(uint64_t seed);
    double noise2(double x, double y) const;
    double noise3(double x, double y, double z) const;
    double fbm2(double x, double y, int octaves) const;
    double fbm3(double x, double y, double z, int octaves) const;
};

// Unified GPU record array built from a VoxNode tree (breadth-first, each
// node's present children stored contiguously in slot order).
//   node record: x = index of the first child record,
//                y = child_mask | leaf_mask << 8
//   leaf record: x = leaf payload (RGBA8), y = canonical id of the parent node
struct GpuTree {
    std::vector<uint32_t> recs;  // 2 u32 per record
    int depth = 0;               // deepest leaf level
    int node_levels = 0;         // deepest node level + 1 (root only: 1)
    bool empty = true;
};
int build_gpu_tree(const yx_scene& s, GpuTree& out);

// 3D Morton code, x in bit 0 of each triple (child slot convention).
inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    auto spread = [](uint64_t v) {
        v &= 0x1fffffull;
        v = (v | (v << 32)) & 0x1f00000000ffffull;
        v = (v | (v << 16)) & 0x1f0000ff0000ffull;
        v = (v | (v << 8)) & 0x100f00f00f00f00full;
        v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
        v = (v | (v << 2)) & 0x1249249249249249ull;
        return v;
    };
    return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

// Builds the canonical breadth-first VoxNode tree from sorted unique leaf
// Morton codes at `depth`; payload(code) gives the leaf payload.
void build_tree_from_codes(const std::vector<uint64_t>& codes, int depth, uint64_t color_seed,
                           yx_scene& out);

// Leaf colour of a synthetic scene: a hash of the Morton code, RGBA8.
inline uint32_t leaf_color(uint64_t code, uint64_t seed) {
    return (uint32_t)mix64(code ^ (seed * 0x9E3779B97F4A7C15ull));
}

}  // namespace yx
