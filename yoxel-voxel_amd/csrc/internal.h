// Internal declarations shared by the host C++ and the HIP translation units of
// libyoxel. Nothing here is part of the C ABI (include/yoxel.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "yoxel.h"

namespace yx {

// An edit in voxel units of the finest level (BuildRange with a source): GROW
// unions the source's solid voxels, CLEAR subtracts them.
// kind 0, SphereSource: the ball of voxel centres within r of (cx, cy, cz);
//   `color` paints grown voxels.
// kind 1, a volume source (IsoSource / RawSource): source voxel (a, b, c)
//   covers finest voxels vx + a s .. vx + (a + 1) s - 1 (and so on per axis);
//   it is solid where EditSet::vsolid[off + a + sx (b + sy c)] is set, and
//   grown voxels take EditSet::vcol at the same index.
struct EditSphere {
    int mode;
    double cx, cy, cz, r;
    uint32_t color;
    int xlo, xhi, ylo, yhi;  // columns the edit can touch
    int zlo, zhi;            // and its z range
    int kind;
    int vx, vy, vz, s;
    int sx, sy, sz;
    uint64_t off;
};
// Edits in order and, per cell of 64 x 64 columns of the (x, y) plane, the ids
// of the edits whose column range meets the cell, ascending (CSR form). A
// column or voxel only ever needs its cell's edits.
struct EditSet {
    static constexpr int kCellShift = 6;
    std::vector<EditSphere> edits;
    std::vector<uint32_t> vcol;   // volume sources' voxel colours
    std::vector<uint8_t> vsolid;  // and solid flags
    int cells = 0;                   // cells per axis
    std::vector<uint32_t> cell_off;  // cells * cells + 1
    std::vector<uint32_t> cell_ids;
    void index(int n);
    bool empty() const { return edits.empty(); }
    const uint32_t* begin(int i, int j) const {
        return cell_ids.data() + cell_off[(size_t)(j >> kCellShift) * cells + (size_t)(i >> kCellShift)];
    }
    const uint32_t* end(int i, int j) const {
        return cell_ids.data() + cell_off[(size_t)(j >> kCellShift) * cells + (size_t)(i >> kCellShift) + 1];
    }
};
}  // namespace yx

// Host copy of an octree in the reference's VoxNode layout
// (reaction/report/main.tex:36-56): 10 u32 per node, canonical ids are indices.
// Edits (yx_scene_write_nodes / yx_scene_set_root) bump `version` and stamp
// the 256-node pages they touch with it, so a device copy can be brought up to
// date page by page (the page versions of HomoStorage / CudaSVO::Update,
// reaction/report/main.tex:67-71).
constexpr uint32_t kScenePageNodes = 256;
uint64_t next_scene_uid();

struct yx_scene {
    std::vector<uint32_t> nodes;
    uint32_t root = YX_EMPTY_NODE;
    uint64_t num_leaves = 0;
    mutable int depth = 0;                 // deepest leaf level, computed on demand after edits
    mutable bool depth_stale = false;      // edited since depth was computed
    uint64_t uid = next_scene_uid();       // never reused by another scene object
    uint64_t version = 1;                  // 1 at creation, +1 per edit
    std::vector<uint64_t> page_version;    // per page; pages past the end: 1
    uint64_t page_ver(size_t p) const { return p < page_version.size() ? page_version[p] : 1u; }
    // synthetic scenes: the recipe and the edits applied so far (sphere and
    // volume sources, yx_scene_build_range*); cleared by node writes
    bool has_recipe = false;
    yx_scene_desc recipe{};
    yx::EditSet edits;
    std::vector<uint32_t> free_nodes;      // blanked node ids, a min-heap (in-place edits)
    std::vector<int32_t> hf_top;           // heightfield top per column, filled on demand
};

namespace yx {

// Deepest supported leaf level (root children are level 1). Bounds the
// per-ray traversal stack.
constexpr int kMaxDepth = 16;

// Appends sphere edits, converted to the finest level of a depth-D grid, to
// `out` (not re-indexed); YX_ERR_INVALID_ARG for a bad mode, level or radius,
// with `out` unchanged.
int add_sphere_edits(const yx_sphere_edit* e, size_t n, int depth, EditSet& out);
// Appends a volume source (IsoSource / RawSource) the same way.
int add_volume_edit(const yx_volume_source& v, int depth, EditSet& out);
// The synthetic scene of a recipe with edits, built on the GPU (svo_build.hip).
int gpu_build_synthetic(const yx_scene_desc& d, const EditSet& edits, yx_scene& out, void* stream);

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
// node's present children stored contiguously in slot order). `base` is the
// index the array will start at in a shared pool (0 for a scene of its own).
//   node record: x = pool index of the first child record,
//                y = child_mask | leaf_mask << 8
//   leaf record: x = leaf payload (RGBA8), y = canonical id of the parent node
struct GpuTree {
    std::vector<uint32_t> recs;  // 2 u32 per record
    int depth = 0;               // deepest leaf level
    int node_levels = 0;         // deepest node level + 1 (root only: 1)
    bool empty = true;
    // level_records[l]: records of levels 0..l (the root is level 0); the
    // breadth-first order puts them first, so they are a prefix of `recs`
    std::vector<uint32_t> level_records;
    // with_lod: per record, the mean RGBA8 of the leaves under a node record
    // (VoxNode::data as the subtree mean); a leaf record holds its own colour
    std::vector<uint32_t> lod;
};
int build_gpu_tree(const yx_scene& s, GpuTree& out, uint32_t base = 0, bool with_lod = false);

// The same records and LOD colours built on the GPU from a device copy of the
// VoxNode array (10 u32 per node): recs (2 u32 per record) and lod on success;
// empty trees give none.
struct GpuLayout {
    uint32_t* recs = nullptr;
    uint32_t* lod = nullptr;
    size_t n_recs = 0;
    int depth = 0;
    int node_levels = 0;
    std::vector<uint32_t> level_records;
    bool empty = true;
};
// The re-layout's device workspace, kept by its owner between calls.
struct RelayoutWork;
RelayoutWork* relayout_work_create();
void relayout_work_free(RelayoutWork* w);
// *recs_io / *lod_io (capacity *cap_io records) are reused when large enough,
// else replaced; they are written only after the tree has been validated.
int gpu_relayout(const uint32_t* d_vox, uint32_t nn, uint32_t root, void* stream, RelayoutWork* w,
                 uint32_t** recs_io, uint32_t** lod_io, size_t* cap_io, GpuLayout& out);

// Pixel-ray basis of a pinhole camera: dir0, du, dv as 9 floats; the ray of
// pixel (x, y) is normalize(dir0 + du x + dv y).
void camera_basis(const float dir[3], const float up[3], float fov, int w, int h, float basis[9]);
// Camera frame: right, down (= -up'), forward as 9 floats.
void camera_frame(const float dir[3], const float up[3], float frame[9]);

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
                           yx_scene& out, const EditSet* edits = nullptr);

// Leaf colour of a synthetic scene: a hash of the Morton code, RGBA8.
inline uint32_t leaf_color(uint64_t code, uint64_t seed) {
    return (uint32_t)mix64(code ^ (seed * 0x9E3779B97F4A7C15ull));
}

}  // namespace yx

namespace yx {

// Coordinate descent over the stack kernel's tuning space, shared by the
// single-scene renderer and the batch. `time_ms(t, &ms)` renders with tuning
// t and returns a status (a candidate that is rejected as invalid is skipped).
// Each axis in turn is set to the value with the lowest time while the others
// stay fixed; sweeps repeat until one changes nothing (at most 3). A new value
// must beat the current one by 1 % to replace it, so timing noise does not
// move the result. Returns the best tuning found in *best.
// With stage_levels the number of top tree levels staged in LDS is one more
// axis (single-scene launches only).
template <class TimeFn>
int tune_coordinate_descent(const yx_tuning& start, TimeFn time_ms, yx_tune_trial* trials, int32_t cap,
                            int32_t* n_trials, yx_tuning* best, bool stage_levels = false) {
    struct Seen {
        yx_tuning t;
        float ms;
    };
    std::vector<Seen> seen;
    int32_t n = 0;
    auto same = [](const yx_tuning& a, const yx_tuning& b) {
        return a.kernel == b.kernel && a.block == b.block && a.tile_w == b.tile_w && a.tile_h == b.tile_h &&
               a.refill_min == b.refill_min && a.queues == b.queues && a.window == b.window &&
               a.waves_per_cu == b.waves_per_cu && a.lds_levels == b.lds_levels;
    };
    auto eval = [&](const yx_tuning& t, float* ms) -> int {
        for (const Seen& s : seen)
            if (same(s.t, t)) {
                *ms = s.ms;
                return YX_OK;
            }
        const int rc = time_ms(t, ms);
        if (rc == YX_ERR_INVALID_ARG || rc == YX_ERR_RANGE) {
            *ms = -1.0f;
        } else if (rc != YX_OK) {
            return rc;
        }
        seen.push_back({t, *ms});
        if (trials && n < cap) trials[n] = {t, *ms};
        ++n;
        return YX_OK;
    };
    static const int tiles[][2] = {{8, 8}, {16, 4}, {4, 16}, {16, 8}, {8, 16}, {16, 16}, {32, 8}};
    static const int refills[] = {24, 32, 40, 48, 56, 64};
    static const int windows[] = {0, 2, 4};
    static const int blocks[] = {64, 128, 256};
    static const int queues[] = {1, 8, 32};
    static const int stages[] = {0, 2, 3, 4, 5};
    yx_tuning cur = start;
    cur.kernel = YX_KERNEL_PERSISTENT;
    float cur_ms = 0.0f;
    int rc = eval(cur, &cur_ms);
    if (rc != YX_OK) return rc;
    if (cur_ms < 0.0f) return fail(YX_ERR_INVALID_ARG, "the starting tuning is not valid");
    auto try_axis = [&](auto apply, int count) -> int {
        yx_tuning pick = cur;
        float pick_ms = cur_ms;
        for (int i = 0; i < count; ++i) {
            yx_tuning t = cur;
            apply(t, i);
            float ms = 0.0f;
            const int e = eval(t, &ms);
            if (e != YX_OK) return e;
            if (ms >= 0.0f && ms < pick_ms * 0.99f) {
                pick = t;
                pick_ms = ms;
            }
        }
        const bool moved = !same(pick, cur);
        cur = pick;
        cur_ms = pick_ms;
        return moved ? 1 : 0;
    };
    for (int sweep = 0; sweep < 3; ++sweep) {
        int moved = 0, e;
        if ((e = try_axis([](yx_tuning& t, int i) { t.tile_w = tiles[i][0]; t.tile_h = tiles[i][1]; }, 7)) < 0)
            return e;
        moved += e;
        if ((e = try_axis([](yx_tuning& t, int i) { t.refill_min = refills[i]; }, 6)) < 0) return e;
        moved += e;
        if ((e = try_axis([](yx_tuning& t, int i) { t.window = windows[i]; }, 3)) < 0) return e;
        moved += e;
        if ((e = try_axis([](yx_tuning& t, int i) { t.block = blocks[i]; }, 3)) < 0) return e;
        moved += e;
        if ((e = try_axis([](yx_tuning& t, int i) { t.queues = queues[i]; }, 3)) < 0) return e;
        moved += e;
        if (stage_levels) {
            if ((e = try_axis([](yx_tuning& t, int i) { t.lds_levels = stages[i]; }, 5)) < 0) return e;
            moved += e;
        }
        if (moved == 0) break;
    }
    if (n_trials) *n_trials = n;
    *best = cur;
    return YX_OK;
}

}  // namespace yx
