// SVO ray-cast kernels for gfx950 and the renderer side of the C ABI.
//
// One frame = primary rays for every pixel (ray generation of
// RendererBase::InitRayDir, cell/renderer_base.h:50-61, and
// cell/ppu_renderer.cpp:56), exact octree traversal (RecTrace,
// cell/ppu_renderer.cpp:18-41), an optional shadow ray per lit hit, and the
// Lambert shade, fused in one kernel (the reference GPU pipeline runs
// InitEyeRays -> Trace -> ShadeSimple as three launches, trace_cuda.py:101-108).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include "internal.h"
#include "timing.h"
#include "trace_device.h"
#include "trace_persist.h"
#include "trace_stack.h"
#include "ssna.h"

namespace yx {

struct FrameArgs {
    const uint2* recs;
    float pos[3];
    float basis[9];
    float light[3];  // unit direction toward the light, as given
    int light_mode;
    int payload;  // YX_PAYLOAD_*
    int w, h;
    yx_outputs out;
};

template <int MAXD>
__global__ __launch_bounds__(256) void frame_kernel(FrameArgs a) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= a.w || y >= a.h) return;
    const size_t offs = (size_t)y * (size_t)a.w + (size_t)x;

    Ray3 d = pixel_dir(a.basis, x, y);
    d.x = adjust_coord(d.x);
    d.y = adjust_coord(d.y);
    d.z = adjust_coord(d.z);

    uint32_t hit_id = YX_MISS, rgba = 0, visits = 0;
    int vx = -1, vy = -1, vz = -1;
    float t = -1.0f;
    uint8_t shadow = 2;
    HitInfo h;
    if (a.recs && trace_ray<MAXD>(a.recs, a.pos, d, false, h, visits)) {
        const uint2 leaf = a.recs[h.leaf_rec];
        hit_id = leaf.y * 8u + (uint32_t)h.child;
        vx = h.vx;
        vy = h.vy;
        vz = h.vz;
        t = max3(h.cell.t1x, h.cell.t1y, h.cell.t1z);
        const int e = entry_axis(h.cell);
        const float de = e == 0 ? d.x : (e == 1 ? d.y : d.z);
        float lam;
        if (a.light_mode == YX_LIGHT_HEADLIGHT) {
            lam = fabsf(de);
        } else {
            const float le = e == 0 ? a.light[0] : (e == 1 ? a.light[1] : a.light[2]);
            lam = (de < 0) ? le : -le;
            lam = dmax(lam, 0.0f);
        }
        uint32_t color = leaf.x;
        if (a.payload == YX_PAYLOAD_C16N16) {  // the leaf's own colour and normal
            lam = a.light_mode == YX_LIGHT_HEADLIGHT ? lambert_c16n16(leaf.x, -d.x, -d.y, -d.z)
                                                     : lambert_c16n16(leaf.x, a.light[0], a.light[1], a.light[2]);
            color = rgba_from_565(leaf.x & 0xffffu);
        }
        bool lit = true;
        if (a.light_mode == YX_LIGHT_SUN_SHADOW && lam > 0.0f) {
            float p[3];
            p[0] = a.pos[0] + d.x * t;
            p[1] = a.pos[1] + d.y * t;
            p[2] = a.pos[2] + d.z * t;
            const float half = __builtin_bit_cast(float, (uint32_t)(127 - (h.level + 1)) << 23);
            if (de < 0)
                p[e] = p[e] + half;
            else
                p[e] = p[e] - half;
            Ray3 l;
            l.x = adjust_coord(a.light[0]);
            l.y = adjust_coord(a.light[1]);
            l.z = adjust_coord(a.light[2]);
            HitInfo hs;
            lit = !trace_ray<MAXD>(a.recs, p, l, true, hs, visits);
            shadow = lit ? 0 : 1;
        }
        const float k = lit ? 0.1f + 0.9f * lam : 0.1f;
        rgba = shade_rgba(color, k);
    }
    if (a.out.hit_id) a.out.hit_id[offs] = hit_id;
    if (a.out.voxel) {
        a.out.voxel[3 * offs + 0] = vx;
        a.out.voxel[3 * offs + 1] = vy;
        a.out.voxel[3 * offs + 2] = vz;
    }
    if (a.out.t) a.out.t[offs] = t;
    if (a.out.rgba) a.out.rgba[offs] = rgba;
    if (a.out.shadow) a.out.shadow[offs] = shadow;
    if (a.out.visits) a.out.visits[offs] = visits;
}

}  // namespace yx

using namespace yx;

// ---- renderer object ------------------------------------------------------

static constexpr int kMaxQueues = 64;
static constexpr int kCounterWords = kMaxQueues + 1;

struct yx_renderer {
    float pos[3] = {0.0f, 0.0f, 0.0f};
    float dir[3] = {1.0f, 0.0f, 0.0f};
    float up[3] = {0.0f, 0.0f, 1.0f};  // RendererBase default m_up(0, 0, 1)
    float fov = 70.0f;                 // RendererBase default m_fov(70.0)
    int w = 640, h = 480;              // RendererBase::SetResolution(640, 480)
    float light[3] = {0.0f, 0.0f, 1.0f};
    int light_mode = YX_LIGHT_HEADLIGHT;
    int payload = YX_PAYLOAD_RGBA8;  // how leaf payloads are read when shading
    bool has_scene = false;
    uint2* d_recs = nullptr;
    size_t n_recs = 0;
    int depth = 0;        // deepest leaf level
    int node_levels = 0;  // deepest node level + 1
    std::vector<uint32_t> level_records;  // records of levels 0..l (a prefix)
    uint32_t* d_frame = nullptr;  // RenderFrame's device colour buffer
    std::vector<uint32_t> h_frame;
    uint32_t* d_counters = nullptr;  // ray queue heads + retired-wave count
    int device = -1;
    int num_cus = 0;
    size_t max_lds = 0;  // dynamic LDS bytes a workgroup may use
    // measured best on MI355X for the 1080p depth-11 frame (stack kernel, one
    // wave per workgroup, no register window, refill at 64 waiting lanes: a
    // wave takes rays only when none of its lanes traces, so the lanes that
    // start together - a whole tile of primary rays, or the shadow rays of
    // the last tile's lit hits - trace in phase)
    yx_tuning tune = {YX_KERNEL_PERSISTENT, 64, 8, 8, 64, 1, 0, 0, 0};
    // shading pipeline (SVORenderer::Render): per-record mean leaf colours,
    // settings with SVORenderer's defaults (demo/SVORenderer.cpp:6-16), and the
    // z / colour work buffers (m_zbuf[2], demo/SVORenderer.h:59)
    uint32_t* d_lod = nullptr;
    float detail_coef = 1.0f;
    bool ssna = true;
    bool show_normals = false;
    yx_light_params lights[YX_MAX_LIGHTS] = {};
    float* d_z[2] = {nullptr, nullptr};
    uint32_t* d_albedo = nullptr;
    size_t work_pixels = 0;
    // device copy of a scene's VoxNode array, kept in sync page by page
    // (CudaSVO::Update): which scene object and version it holds
    uint32_t* d_vox = nullptr;
    size_t vox_cap = 0, vox_n = 0;  // nodes
    uint64_t mirror_uid = 0, mirror_version = 0;
    uint32_t last_pages = 0;
    uint64_t last_bytes = 0;
    RelayoutWork* rl_work = nullptr;  // GPU re-layout workspace
    size_t recs_cap = 0;              // records d_recs / d_lod can hold
    // staging of scattered dirty pages: page table + node bytes, pinned on
    // the host and its device twin
    uint8_t* stage_host = nullptr;
    uint8_t* d_stage = nullptr;
    size_t stage_cap = 0;
};

#define YX_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return fail(YX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// RendererBase::InitRayDir (cell/renderer_base.h:50-61).
static void init_ray_dir(const yx_renderer* r, float* out) { camera_basis(r->dir, r->up, r->fov, r->w, r->h, out); }

// Compute units and LDS limit of the current device (no allocation).
static int query_device(yx_renderer* r) {
    if (r->num_cus > 0) return YX_OK;
    int dev = 0;
    YX_HIP(hipGetDevice(&dev));
    int cus = 0;
    YX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int lds_max = 0;
    YX_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    r->max_lds = (size_t)lds_max;
    r->num_cus = cus;
    return YX_OK;
}

// query_device, plus the renderer's own queue counters.
static int ensure_device(yx_renderer* r) {
    if (r->device >= 0) return YX_OK;
    int rc = query_device(r);
    if (rc != YX_OK) return rc;
    int dev = 0;
    YX_HIP(hipGetDevice(&dev));
    hipError_t e = hipMalloc((void**)&r->d_counters, kCounterWords * sizeof(uint32_t));
    if (e != hipSuccess) {
        r->d_counters = nullptr;
        return fail(YX_ERR_OOM, "hipMalloc counters: %s", hipGetErrorString(e));
    }
    YX_HIP(hipMemset(r->d_counters, 0, kCounterWords * sizeof(uint32_t)));
    r->device = dev;
    return YX_OK;
}

// Persistent launch: as many workgroups as fit on the device at once (or the
// tuned waves per CU); every wave loops over the ray queues until they drain.
static int launch_persistent(yx_renderer* r, const void* fn, void (*kern)(PersistArgs), int block,
                             PersistArgs& a, size_t lds, hipStream_t stream) {
    int per_cu = r->tune.waves_per_cu > 0 ? std::max(1, r->tune.waves_per_cu / (block / 64)) : 0;
    if (per_cu == 0) {
        int occ = 0;
        YX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, block, lds));
        per_cu = std::max(1, occ);
    }
    uint32_t grid = (uint32_t)(per_cu * r->num_cus);
    // no more waves than 64-ray packets: a small image would otherwise start
    // waves that find the queue empty
    const uint32_t waves_per_block = (uint32_t)(block / 64);
    const uint32_t useful = (a.total / 64u + waves_per_block - 1) / waves_per_block;
    grid = std::max(1u, std::min(grid, useful));
    a.n_waves = grid * waves_per_block;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, stream, a);
    YX_HIP(hipGetLastError());
    return YX_OK;
}

#define YX_LAUNCH(KERNEL, BLOCK) \
    launch_persistent(r, (const void*)KERNEL, KERNEL, BLOCK, a, lds, stream)

// `counters`: the persistent kernels' queue heads and retired-wave count
// (kCounterWords words, zero on entry; the kernels leave them zero); null
// means the renderer's own, allocated on first use.
static int launch_frame(yx_renderer* r, const yx_outputs& out, hipStream_t stream, uint32_t* counters = nullptr) {
    float basis[9];
    init_ray_dir(r, basis);
    if (r->tune.kernel == YX_KERNEL_PER_PIXEL || !r->d_recs) {
        FrameArgs a;
        std::memset(&a, 0, sizeof(a));
        a.recs = r->d_recs;
        for (int i = 0; i < 3; ++i) {
            a.pos[i] = r->pos[i];
            a.light[i] = r->light[i];
        }
        std::memcpy(a.basis, basis, sizeof(basis));
        a.light_mode = r->light_mode;
        a.payload = r->payload;
        a.w = r->w;
        a.h = r->h;
        a.out = out;
        dim3 grid((r->w + 15) / 16, (r->h + 15) / 16);
        // the per-lane stack holds node levels 0 .. node_levels - 2
        if (r->node_levels <= 9)
            hipLaunchKernelGGL(frame_kernel<8>, grid, dim3(256), 0, stream, a);
        else if (r->node_levels <= 13)
            hipLaunchKernelGGL(frame_kernel<12>, grid, dim3(256), 0, stream, a);
        else
            hipLaunchKernelGGL(frame_kernel<kMaxDepth>, grid, dim3(256), 0, stream, a);
        YX_HIP(hipGetLastError());
        return YX_OK;
    }
    int rc = counters ? query_device(r) : ensure_device(r);
    if (rc != YX_OK) return rc;
    PersistArgs a;
    std::memset(&a, 0, sizeof(a));
    a.recs = r->d_recs;
    for (int i = 0; i < 3; ++i) {
        a.pos[i] = r->pos[i];
        a.light[i] = r->light[i];
    }
    std::memcpy(a.basis, basis, sizeof(basis));
    a.light_mode = r->light_mode;
    a.payload = r->payload;
    a.w = r->w;
    a.h = r->h;
    a.out = out;
    a.counters = counters ? counters : r->d_counters;
    const yx_tuning& t = r->tune;
    a.tile_wl = __builtin_ctz((unsigned)t.tile_w);
    a.tile_hl = __builtin_ctz((unsigned)t.tile_h);
    a.tiles_x = (r->w + t.tile_w - 1) / t.tile_w;
    const int tiles_y = (r->h + t.tile_h - 1) / t.tile_h;
    const uint32_t tile = (uint32_t)(t.tile_w * t.tile_h);
    a.total = (uint32_t)a.tiles_x * (uint32_t)tiles_y * tile;
    a.nq = t.queues;
    const uint32_t n_tiles = (uint32_t)a.tiles_x * (uint32_t)tiles_y;
    a.per_q = queue_rays(n_tiles, tile, t.queues);  // stack kernel: interleaved tile groups
    a.refill_min = t.refill_min;
    // stack rows: levels 0 .. node_levels - 2 are pushed
    a.levels = std::max(1, r->node_levels - 1);
    int kernel = t.kernel;
    int block = t.block;
    if (kernel == YX_KERNEL_PERSISTENT) {
        // LDS stack rows per lane (stack_lane_bytes); shrink the workgroup, then
        // fall back to the replay kernel (same results) when they do not fit
        // W == 0 rows hold record indices below 2^29; larger pools use a window
        int win = t.window <= 0 ? 0 : (t.window <= 2 ? 2 : (t.window <= 4 ? 4 : 6));
        if (win == 0 && r->n_recs >= kStackMaxRecords) win = 2;
        const size_t lane_bytes = stack_lane_bytes(a.levels, win);
        size_t lds = lane_bytes * (size_t)block;
        while (lds > r->max_lds && block > 64) {
            block /= 2;
            lds = lane_bytes * (size_t)block;
        }
        // top tree levels staged in LDS (window 0 only): the records of levels
        // 0..lds_levels-1 (a prefix of the breadth-first array), as many
        // levels as fit
        if (t.lds_levels > 0 && win == 0 && !r->level_records.empty()) {
            int lv = std::min<int>(t.lds_levels, (int)r->level_records.size()) - 1;
            while (lv >= 0 && lds + (size_t)r->level_records[(size_t)lv] * 8u > r->max_lds) --lv;
            if (lv >= 0) {
                a.lds_recs = r->level_records[(size_t)lv];
                lds += ((size_t)a.lds_recs * 8u + 15u) & ~(size_t)15u;
                if (block == 64) return YX_LAUNCH((stack_kernel<64, 0, false, true>), 64);
                if (block == 128) return YX_LAUNCH((stack_kernel<128, 0, false, true>), 128);
                return YX_LAUNCH((stack_kernel<256, 0, false, true>), 256);
            }
        }
        if (lds <= r->max_lds) {
            if (block == 64) {
                if (win == 0) return YX_LAUNCH((stack_kernel<64, 0, false>), 64);
                if (win == 2) return YX_LAUNCH((stack_kernel<64, 2, false>), 64);
                if (win == 4) return YX_LAUNCH((stack_kernel<64, 4, false>), 64);
                return YX_LAUNCH((stack_kernel<64, 6, false>), 64);
            }
            if (block == 128) {
                if (win == 0) return YX_LAUNCH((stack_kernel<128, 0, false>), 128);
                if (win == 2) return YX_LAUNCH((stack_kernel<128, 2, false>), 128);
                if (win == 4) return YX_LAUNCH((stack_kernel<128, 4, false>), 128);
                return YX_LAUNCH((stack_kernel<128, 6, false>), 128);
            }
            if (win == 0) return YX_LAUNCH((stack_kernel<256, 0, false>), 256);
            if (win == 2) return YX_LAUNCH((stack_kernel<256, 2, false>), 256);
            if (win == 4) return YX_LAUNCH((stack_kernel<256, 4, false>), 256);
            return YX_LAUNCH((stack_kernel<256, 6, false>), 256);
        }
        kernel = YX_KERNEL_REPLAY;
        block = 256;
    }
    // the replay kernel's queues are contiguous bands of tiles
    a.per_q = ((n_tiles + (uint32_t)t.queues - 1u) / (uint32_t)t.queues) * tile;
    a.levels = std::max(1, r->node_levels);
    block = block >= 512 ? 512 : 256;
    size_t lds = (size_t)a.levels * (size_t)block * sizeof(uint2);
    if (lds > r->max_lds && block == 512) {  // deep trees: same results with smaller workgroups
        block = 256;
        lds = (size_t)a.levels * (size_t)block * sizeof(uint2);
    }
    if (lds > r->max_lds)
        return fail(YX_ERR_RANGE, "per-level LDS rows need %zu bytes per workgroup, device allows %zu", lds,
                    r->max_lds);
    if (block == 256) {
        if (t.window == 0) return YX_LAUNCH((persist_kernel<0, 256>), 256);
        if (t.window == 1) return YX_LAUNCH((persist_kernel<1, 256>), 256);
        return YX_LAUNCH((persist_kernel<2, 256>), 256);
    }
    if (t.window == 0) return YX_LAUNCH((persist_kernel<0, 512>), 512);
    if (t.window == 1) return YX_LAUNCH((persist_kernel<1, 512>), 512);
    return YX_LAUNCH((persist_kernel<2, 512>), 512);  // windows above 2 use 2
}

extern "C" int yx_renderer_create(yx_renderer** out) {
    clear_error();
    if (!out) return fail(YX_ERR_INVALID_ARG, "null argument");
    yx_renderer* r = new (std::nothrow) yx_renderer();
    if (!r) return fail(YX_ERR_OOM, "out of host memory");
    *out = r;
    return YX_OK;
}

extern "C" void yx_renderer_destroy(yx_renderer* r) {
    if (!r) return;
    if (r->d_recs) (void)hipFree(r->d_recs);
    if (r->d_frame) (void)hipFree(r->d_frame);
    if (r->d_counters) (void)hipFree(r->d_counters);
    if (r->d_lod) (void)hipFree(r->d_lod);
    for (float* z : r->d_z)
        if (z) (void)hipFree(z);
    if (r->d_albedo) (void)hipFree(r->d_albedo);
    if (r->d_vox) (void)hipFree(r->d_vox);
    if (r->stage_host) (void)hipHostFree(r->stage_host);
    if (r->d_stage) (void)hipFree(r->d_stage);
    relayout_work_free(r->rl_work);
    delete r;
}

extern "C" int yx_renderer_set_tuning(yx_renderer* r, const yx_tuning* t) {
    clear_error();
    if (!r || !t) return fail(YX_ERR_INVALID_ARG, "null argument");
    auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
    if (t->kernel < YX_KERNEL_PER_PIXEL || t->kernel > YX_KERNEL_REPLAY)
        return fail(YX_ERR_INVALID_ARG, "unknown kernel %d", t->kernel);
    if (t->block != 64 && t->block != 128 && t->block != 256 && t->block != 512)
        return fail(YX_ERR_INVALID_ARG, "block must be 64, 128, 256 or 512");
    if (!pow2(t->tile_w) || !pow2(t->tile_h) || t->tile_w * t->tile_h > 4096)
        return fail(YX_ERR_INVALID_ARG, "tile %dx%d: powers of two, at most 4096 pixels", t->tile_w, t->tile_h);
    if (t->refill_min < 1 || t->refill_min > 64) return fail(YX_ERR_INVALID_ARG, "refill_min must be 1..64");
    if (t->queues < 1 || t->queues > kMaxQueues) return fail(YX_ERR_INVALID_ARG, "queues must be 1..%d", kMaxQueues);
    if (t->window < 0 || t->window > 6) return fail(YX_ERR_INVALID_ARG, "window must be 0..6");
    if (t->waves_per_cu < 0 || t->waves_per_cu > 32) return fail(YX_ERR_INVALID_ARG, "waves_per_cu must be 0..32");
    if (t->lds_levels < 0 || t->lds_levels > kMaxDepth)
        return fail(YX_ERR_INVALID_ARG, "lds_levels must be 0..%d", kMaxDepth);
    r->tune = *t;
    return YX_OK;
}

extern "C" int yx_renderer_get_tuning(const yx_renderer* r, yx_tuning* t) {
    clear_error();
    if (!r || !t) return fail(YX_ERR_INVALID_ARG, "null argument");
    *t = r->tune;
    return YX_OK;
}

extern "C" int yx_renderer_set_scene_host(yx_renderer* r, const yx_scene* s) {
    clear_error();
    if (!r || !s) return fail(YX_ERR_INVALID_ARG, "null argument");
    GpuTree t;
    int rc;
    try {
        rc = build_gpu_tree(*s, t, 0, true);
    } catch (const std::bad_alloc&) {
        return fail(YX_ERR_OOM, "out of host memory in the re-layout");
    }
    if (rc != YX_OK) return rc;
    if (r->d_recs) {
        YX_HIP(hipFree(r->d_recs));
        r->d_recs = nullptr;
    }
    if (r->d_lod) {
        YX_HIP(hipFree(r->d_lod));
        r->d_lod = nullptr;
    }
    r->n_recs = 0;
    r->depth = 0;
    r->node_levels = 0;
    if (!t.empty) {
        const size_t bytes = t.recs.size() * sizeof(uint32_t);
        hipError_t e = hipMalloc((void**)&r->d_recs, bytes);
        if (e != hipSuccess) {
            r->d_recs = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%zu) for the scene: %s", bytes, hipGetErrorString(e));
        }
        YX_HIP(hipMemcpy(r->d_recs, t.recs.data(), bytes, hipMemcpyHostToDevice));
        const size_t lod_bytes = t.lod.size() * sizeof(uint32_t);
        e = hipMalloc((void**)&r->d_lod, lod_bytes);
        if (e != hipSuccess) {
            r->d_lod = nullptr;
            (void)hipFree(r->d_recs);
            r->d_recs = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%zu) for the LOD colours: %s", lod_bytes, hipGetErrorString(e));
        }
        YX_HIP(hipMemcpy(r->d_lod, t.lod.data(), lod_bytes, hipMemcpyHostToDevice));
        r->n_recs = t.recs.size() / 2;
        r->recs_cap = r->n_recs;
        r->depth = t.depth;
        r->node_levels = t.node_levels;
        r->level_records = t.level_records;
    } else {
        r->level_records.clear();
    }
    r->has_scene = true;
    return YX_OK;
}

// SetScene: the whole VoxNode array goes up once and is laid out on the GPU
// (the path of yx_renderer_update_scene for a scene object the renderer has
// not seen), then the call waits for the records.
extern "C" int yx_renderer_set_scene(yx_renderer* r, const yx_scene* s) {
    clear_error();
    if (!r || !s) return fail(YX_ERR_INVALID_ARG, "null argument");
    // nothing to lay out: the host path needs no device
    if (s->root == YX_EMPTY_NODE || s->nodes.empty()) return yx_renderer_set_scene_host(r, s);
    r->mirror_uid = 0;  // upload every page
    const int rc = yx_renderer_update_scene(r, s, nullptr);
    if (rc != YX_OK) return rc;
    YX_HIP(hipStreamSynchronize(nullptr));
    return YX_OK;
}

extern "C" int yx_renderer_set_view_pos(yx_renderer* r, const float pos[3]) {
    clear_error();
    if (!r || !pos) return fail(YX_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < 3; ++i) r->pos[i] = pos[i];
    return YX_OK;
}

extern "C" int yx_renderer_set_view_dir(yx_renderer* r, const float dir[3]) {
    clear_error();
    if (!r || !dir) return fail(YX_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < 3; ++i) r->dir[i] = dir[i];
    return YX_OK;
}

extern "C" int yx_renderer_set_view_up(yx_renderer* r, const float up[3]) {
    clear_error();
    if (!r || !up) return fail(YX_ERR_INVALID_ARG, "null argument");
    for (int i = 0; i < 3; ++i) r->up[i] = up[i];
    return YX_OK;
}

extern "C" int yx_renderer_set_resolution(yx_renderer* r, int32_t w, int32_t h) {
    clear_error();
    if (!r) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (w <= 0 || h <= 0) return fail(YX_ERR_INVALID_ARG, "resolution %dx%d must be positive", w, h);
    if ((int64_t)w * (int64_t)h > (int64_t)1 << 28) return fail(YX_ERR_RANGE, "resolution %dx%d too large", w, h);
    if (w != r->w || h != r->h) {
        if (r->d_frame) {
            YX_HIP(hipFree(r->d_frame));
            r->d_frame = nullptr;
        }
        r->h_frame.clear();
    }
    r->w = w;
    r->h = h;
    return YX_OK;
}

extern "C" int yx_renderer_get_resolution(const yx_renderer* r, int32_t* w, int32_t* h) {
    clear_error();
    if (!r || !w || !h) return fail(YX_ERR_INVALID_ARG, "null argument");
    *w = r->w;
    *h = r->h;
    return YX_OK;
}

extern "C" int yx_renderer_set_fov(yx_renderer* r, float fov) {
    clear_error();
    if (!r) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (!(fov > 0.0f && fov < 180.0f)) return fail(YX_ERR_INVALID_ARG, "fov %g outside (0, 180)", fov);
    r->fov = fov;
    return YX_OK;
}

extern "C" int yx_renderer_set_light(yx_renderer* r, const float dir[3], int32_t mode) {
    clear_error();
    if (!r || !dir) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (mode < YX_LIGHT_HEADLIGHT || mode > YX_LIGHT_SUN_SHADOW)
        return fail(YX_ERR_INVALID_ARG, "unknown light mode %d", mode);
    for (int i = 0; i < 3; ++i) r->light[i] = dir[i];
    r->light_mode = mode;
    return YX_OK;
}

extern "C" int yx_renderer_set_payload(yx_renderer* r, int32_t format) {
    clear_error();
    if (!r) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (format != YX_PAYLOAD_RGBA8 && format != YX_PAYLOAD_C16N16)
        return fail(YX_ERR_INVALID_ARG, "payload format %d", format);
    r->payload = format;
    return YX_OK;
}

extern "C" int yx_renderer_ray_basis(const yx_renderer* r, float basis[9]) {
    clear_error();
    if (!r || !basis) return fail(YX_ERR_INVALID_ARG, "null argument");
    init_ray_dir(r, basis);
    return YX_OK;
}

extern "C" int yx_render(yx_renderer* r, const yx_outputs* out, void* stream) {
    clear_error();
    if (!r || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (!r->has_scene) return fail(YX_ERR_NO_SCENE, "no scene set");
    return launch_frame(r, *out, (hipStream_t)stream);
}

// ---- allocation-free rendering -------------------------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops that forbid it): the persistent kernels' only scratch
// - the ray-queue heads and the retired-wave count, kScratchWords words
// (260 bytes) - lives in this file-scope array instead of a renderer-owned
// allocation. Each device holds its own copy from module load until the
// process ends; every call on a device shares it. Host code uses the device
// address that hipGetSymbolAddress returns for the current device, never the
// array's name.
static constexpr int kScratchWords = kCounterWords;
static constexpr int kScratchDevices = 64;
static __device__ __align__(256) uint32_t scratch[kScratchWords];
static std::mutex g_scratch_mu;
static uint32_t* g_scratch_ptr[kScratchDevices];  // per device, looked up once

static int scratch_address(uint32_t** out) {
    int dev = 0;
    YX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        return fail(YX_ERR_RANGE, "device %d: the allocation-free path serves devices 0..%d", dev,
                    kScratchDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(YX_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[dev];
    return YX_OK;
}

__global__ void zero_words_kernel(uint32_t* p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0u;
}

extern "C" int yx_render_noalloc(yx_renderer* r, const yx_outputs* out, void* stream) {
    clear_error();
    if (!r || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (!r->has_scene) return fail(YX_ERR_NO_SCENE, "no scene set");
    if (r->tune.queues + 1 > kScratchWords)
        return fail(YX_ERR_RANGE, "%d ray queues need %d scratch words; the allocation-free path holds %d",
                    r->tune.queues, r->tune.queues + 1, kScratchWords);
    uint32_t* s = nullptr;
    int rc = scratch_address(&s);
    if (rc != YX_OK) return rc;
    const hipStream_t st = (hipStream_t)stream;
    // the queue heads must start at zero (the last call's kernels left them
    // zero; a fill kernel rather than a memset node keeps the call capturable)
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(128), 0, st, s, kScratchWords);
    YX_HIP(hipGetLastError());
    return launch_frame(r, *out, st, s);
}

extern "C" int yx_render_frame(yx_renderer* r, const uint32_t** colors) {
    clear_error();
    if (!r || !colors) return fail(YX_ERR_INVALID_ARG, "null argument");
    *colors = nullptr;
    if (!r->has_scene) return fail(YX_ERR_NO_SCENE, "no scene set");
    const size_t n = (size_t)r->w * (size_t)r->h;
    if (!r->d_frame) {
        hipError_t e = hipMalloc((void**)&r->d_frame, n * sizeof(uint32_t));
        if (e != hipSuccess) {
            r->d_frame = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc frame: %s", hipGetErrorString(e));
        }
    }
    try {
        r->h_frame.resize(n);
    } catch (const std::bad_alloc&) {
        return fail(YX_ERR_OOM, "out of host memory");
    }
    yx_outputs o;
    std::memset(&o, 0, sizeof(o));
    o.rgba = r->d_frame;
    int rc = launch_frame(r, o, nullptr);
    if (rc != YX_OK) return rc;
    YX_HIP(hipMemcpy(r->h_frame.data(), r->d_frame, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    *colors = r->h_frame.data();
    return YX_OK;
}

extern "C" int yx_renderer_scene_bytes(const yx_renderer* r, uint64_t* bytes) {
    clear_error();
    if (!r || !bytes) return fail(YX_ERR_INVALID_ARG, "null argument");
    *bytes = (uint64_t)r->n_recs * 12u;  // 8-byte records + 4-byte LOD colours
    return YX_OK;
}

extern "C" int yx_renderer_autotune(yx_renderer* r, const yx_outputs* out, int32_t reps, yx_tune_trial* trials,
                                    int32_t cap, int32_t* n_trials, void* stream) {
    clear_error();
    if (!r || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (!r->has_scene) return fail(YX_ERR_NO_SCENE, "no scene set");
    if (reps < 1) return fail(YX_ERR_INVALID_ARG, "reps must be >= 1");
    const yx_tuning keep = r->tune;
    const hipStream_t st = (hipStream_t)stream;
    auto time_ms = [&](const yx_tuning& t, float* ms) -> int {
        int rc = yx_renderer_set_tuning(r, &t);
        if (rc != YX_OK) return rc;
        return time_frames([&]() { return launch_frame(r, *out, st); }, reps, st, ms);
    };
    yx_tuning best = keep;
    const int rc = tune_coordinate_descent(keep, time_ms, trials, cap, n_trials, &best, true);
    r->tune = rc == YX_OK ? best : keep;
    return rc;
}

// ---- shading pipeline (SVORenderer::Render, demo/SVORenderer.cpp:95-149) ---

extern "C" int yx_renderer_set_detail_coef(yx_renderer* r, float coef) {
    clear_error();
    if (!r) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (!(coef >= 0.0f) || !std::isfinite(coef)) return fail(YX_ERR_INVALID_ARG, "detail coef must be finite and >= 0");
    r->detail_coef = coef;
    return YX_OK;
}

extern "C" int yx_renderer_get_detail_coef(const yx_renderer* r, float* coef) {
    clear_error();
    if (!r || !coef) return fail(YX_ERR_INVALID_ARG, "null argument");
    *coef = r->detail_coef;
    return YX_OK;
}

extern "C" int yx_renderer_set_ssna(yx_renderer* r, int32_t enable) {
    clear_error();
    if (!r) return fail(YX_ERR_INVALID_ARG, "null argument");
    r->ssna = enable != 0;
    return YX_OK;
}

extern "C" int yx_renderer_get_ssna(const yx_renderer* r, int32_t* enable) {
    clear_error();
    if (!r || !enable) return fail(YX_ERR_INVALID_ARG, "null argument");
    *enable = r->ssna ? 1 : 0;
    return YX_OK;
}

extern "C" int yx_renderer_set_show_normals(yx_renderer* r, int32_t enable) {
    clear_error();
    if (!r) return fail(YX_ERR_INVALID_ARG, "null argument");
    r->show_normals = enable != 0;
    return YX_OK;
}

extern "C" int yx_renderer_set_point_light(yx_renderer* r, int32_t i, const yx_light_params* light) {
    clear_error();
    if (!r || !light) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (i < 0 || i >= YX_MAX_LIGHTS) return fail(YX_ERR_INVALID_ARG, "light index %d not in 0..%d", i, YX_MAX_LIGHTS - 1);
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(light->pos[k]) || !std::isfinite(light->diffuse[k]) || !std::isfinite(light->specular[k]) ||
            !std::isfinite(light->attenuation[k]))
            return fail(YX_ERR_INVALID_ARG, "light %d: non-finite parameter", i);
    // the attenuation 1 / (c0 + c1 d + c2 d^2) stays finite and positive
    if (!(light->attenuation[0] > 0.0f) || light->attenuation[1] < 0.0f || light->attenuation[2] < 0.0f)
        return fail(YX_ERR_INVALID_ARG, "light %d: attenuation needs c0 > 0, c1 >= 0, c2 >= 0", i);
    r->lights[i] = *light;
    r->lights[i].enabled = light->enabled != 0;
    return YX_OK;
}

// Frame constants of the pipeline (demo/SVORenderer.cpp:104-105,131-139).
struct PipelineParams {
    float detail, pixel_ang, pix, vox, edge, zlimit[kBlurPasses];
};

static PipelineParams pipeline_params(const yx_renderer* r) {
    PipelineParams p;
    const float deg2rad = (float)(3.14159265358979323846 / (double)180.0f);
    const float w = (float)r->w;
    p.detail = (r->detail_coef * ((r->fov / 2.0f) * deg2rad)) / w;
    p.pixel_ang = (r->fov * deg2rad) / w;
    const float da = tanf((r->fov / 2.0f) * deg2rad) / w;
    p.pix = 2.0f * da;
    p.vox = ldexpf(1.0f, -r->depth);
    p.edge = 4.0f * p.vox;
    float blur = 3.0f;
    for (int i = 0; i < kBlurPasses; ++i) {
        p.zlimit[i] = (5.0f * p.vox) / (p.pixel_ang * blur);
        blur += 3.0f;
    }
    return p;
}

// SVORenderer::InitBlur (demo/SVORenderer.cpp:55-79): exp(-(tx^2 + ty^2)) with
// tx = 2 (x - h) / h over the 5x5 window, normalised by the sum.
static void blur_kernel_table(float* kern) {
    const float h = (float)(kBlurK / 2);
    const float scale = 2.0f;
    float sum = 0.0f;
    for (int y = 0; y < kBlurK; ++y)
        for (int x = 0; x < kBlurK; ++x) {
            float tx = scale * ((float)x - h) / h;
            float ty = scale * ((float)y - h) / h;
            tx = tx * tx;
            ty = ty * ty;
            const float v = expf(-(tx + ty));
            kern[y * kBlurK + x] = v;
            sum += v;
        }
    for (int i = 0; i < kBlurK * kBlurK; ++i) kern[i] /= sum;
}

static int ensure_work_buffers(yx_renderer* r, size_t npix) {
    if (npix <= r->work_pixels && r->d_albedo) return YX_OK;
    for (float*& z : r->d_z)
        if (z) {
            YX_HIP(hipFree(z));
            z = nullptr;
        }
    if (r->d_albedo) {
        YX_HIP(hipFree(r->d_albedo));
        r->d_albedo = nullptr;
    }
    r->work_pixels = 0;
    hipError_t e = hipMalloc((void**)&r->d_z[0], npix * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&r->d_z[1], npix * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&r->d_albedo, npix * sizeof(uint32_t));
    if (e != hipSuccess) {
        for (float*& z : r->d_z)
            if (z) {
                (void)hipFree(z);
                z = nullptr;
            }
        if (r->d_albedo) (void)hipFree(r->d_albedo);
        r->d_albedo = nullptr;
        return fail(YX_ERR_OOM, "hipMalloc for the pipeline's %zu-pixel work buffers: %s", npix, hipGetErrorString(e));
    }
    r->work_pixels = npix;
    return YX_OK;
}

extern "C" int yx_render_pipeline(yx_renderer* r, const yx_pipeline_outputs* out, void* stream) {
    clear_error();
    if (!r || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (!r->has_scene) return fail(YX_ERR_NO_SCENE, "no scene set");
    int rc = ensure_device(r);
    if (rc != YX_OK) return rc;
    const size_t npix = (size_t)r->w * (size_t)r->h;
    rc = ensure_work_buffers(r, npix);
    if (rc != YX_OK) return rc;
    const hipStream_t st = (hipStream_t)stream;
    const PipelineParams pp = pipeline_params(r);
    float basis[9];
    init_ray_dir(r, basis);

    // 1. LOD trace (Run_Trace): persistent stack kernel, primary rays only
    PersistArgs a;
    std::memset(&a, 0, sizeof(a));
    a.recs = r->d_recs;  // null for an empty scene: every ray misses
    for (int i = 0; i < 3; ++i) a.pos[i] = r->pos[i];
    std::memcpy(a.basis, basis, sizeof(basis));
    a.w = r->w;
    a.h = r->h;
    a.out.t = out->t;
    a.out.visits = out->visits;
    a.counters = r->d_counters;
    const yx_tuning& t = r->tune;
    a.tile_wl = __builtin_ctz((unsigned)t.tile_w);
    a.tile_hl = __builtin_ctz((unsigned)t.tile_h);
    a.tiles_x = (r->w + t.tile_w - 1) / t.tile_w;
    const int tiles_y = (r->h + t.tile_h - 1) / t.tile_h;
    const uint32_t tile = (uint32_t)(t.tile_w * t.tile_h);
    a.total = (uint32_t)a.tiles_x * (uint32_t)tiles_y * tile;
    a.nq = t.queues;
    a.per_q = queue_rays((uint32_t)a.tiles_x * (uint32_t)tiles_y, tile, t.queues);
    a.refill_min = t.refill_min;
    a.levels = std::max(1, r->node_levels - 1);
    a.lod = r->d_lod;
    a.detail = pp.detail;
    a.z = r->d_z[0];
    a.albedo = out->albedo ? out->albedo : r->d_albedo;
    a.hit_level = out->level;
    int block = std::min(256, std::max(64, t.block));
    if (r->n_recs >= kStackMaxRecords)
        return fail(YX_ERR_RANGE, "the LOD trace holds record indices below 2^29; the scene has %zu records",
                    r->n_recs);
    const size_t lane_bytes = stack_lane_bytes(a.levels, 0);
    size_t lds = lane_bytes * (size_t)block;
    while (lds > r->max_lds && block > 64) {
        block /= 2;
        lds = lane_bytes * (size_t)block;
    }
    if (lds > r->max_lds)
        return fail(YX_ERR_RANGE, "LOD trace stack needs %zu bytes of LDS per workgroup, device allows %zu", lds,
                    r->max_lds);
    if (block == 64)
        rc = launch_persistent(r, (const void*)stack_kernel<64, 0, false, false, true>,
                               stack_kernel<64, 0, false, false, true>, 64, a, lds, st);
    else if (block == 128)
        rc = launch_persistent(r, (const void*)stack_kernel<128, 0, false, false, true>,
                               stack_kernel<128, 0, false, false, true>, 128, a, lds, st);
    else
        rc = launch_persistent(r, (const void*)stack_kernel<256, 0, false, false, true>,
                               stack_kernel<256, 0, false, false, true>, 256, a, lds, st);
    if (rc != YX_OK) return rc;

    // 2. depth-limited blur passes (Run_BlurZ), ping-pong between the z buffers
    int cur = 0;
    if (r->ssna) {
        BlurArgs b;
        b.w = r->w;
        b.h = r->h;
        b.edge = pp.edge;
        blur_kernel_table(b.kern);
        const dim3 grid((unsigned)((r->w + kBlurTX - 1) / kBlurTX), (unsigned)((r->h + kBlurTY - 1) / kBlurTY));
        for (int i = 0; i < kBlurPasses; ++i) {
            b.zin = r->d_z[cur];
            b.zout = r->d_z[1 - cur];
            b.zlimit = pp.zlimit[i];
            hipLaunchKernelGGL(blur_z_kernel, grid, dim3(kBlurThreads), 0, st, b);
            YX_HIP(hipGetLastError());
            cur = 1 - cur;
        }
    }

    // 3. Phong shading with normals from z (Run_ShadeSimple)
    ShadeArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.z = r->d_z[cur];
    sa.albedo = a.albedo;
    sa.rgba = out->rgba;
    sa.zout = out->z;
    sa.w = r->w;
    sa.h = r->h;
    std::memcpy(sa.basis, basis, sizeof(basis));
    for (int i = 0; i < 3; ++i) sa.eye[i] = r->pos[i];
    camera_frame(r->dir, r->up, sa.frame);
    sa.pix = pp.pix;
    sa.ambient = 0.1f;  // rp.ambient, demo/SVORenderer.cpp:113
    sa.show_normals = r->show_normals ? 1 : 0;
    for (int i = 0; i < YX_MAX_LIGHTS; ++i) sa.lights[i] = r->lights[i];
    const dim3 sgrid((unsigned)((r->w + kShadeTX - 1) / kShadeTX), (unsigned)((r->h + kShadeTY - 1) / kShadeTY));
    hipLaunchKernelGGL(shade_simple_kernel, sgrid, dim3(kShadeTX * kShadeTY), 0, st, sa);
    YX_HIP(hipGetLastError());
    return YX_OK;
}

// ---- incremental scene sync (CudaSVO::Update, reaction/report/main.tex:67-71) ----

// Pages of `s` the renderer's device copy lacks: all of them for another scene
// object, else those edited after the synced version and those past its end.
static void dirty_pages(const yx_renderer* r, const yx_scene* s, std::vector<std::pair<size_t, size_t>>* runs,
                        uint32_t* pages, uint64_t* bytes) {
    const size_t nn = s->nodes.size() / 10;
    const size_t np = (nn + kScenePageNodes - 1) / kScenePageNodes;
    const bool all = r->mirror_uid != s->uid;
    const size_t have = all ? 0 : r->vox_n;
    uint32_t cnt = 0;
    uint64_t b = 0;
    size_t run_begin = 0;
    bool in_run = false;
    for (size_t p = 0; p <= np; ++p) {
        bool d = false;
        if (p < np) {
            const size_t n0 = p * kScenePageNodes, n1 = std::min(nn, n0 + kScenePageNodes);
            d = all || n1 > have || s->page_ver(p) > r->mirror_version;
            if (d) {
                ++cnt;
                b += (uint64_t)(n1 - n0) * 40u;
            }
        }
        if (d && !in_run) {
            run_begin = p;
            in_run = true;
        } else if (!d && in_run) {
            if (runs) runs->push_back({run_begin * kScenePageNodes, std::min(nn, p * kScenePageNodes)});
            in_run = false;
        }
    }
    if (pages) *pages = cnt;
    if (bytes) *bytes = b;
}

// One dirty page: `nodes` VoxNodes from staging node `src` to scene node `dst`.
struct PageCopy {
    uint32_t dst, src, nodes, pad;
};

// One workgroup per page, 8-byte words (node offsets are multiples of 40 B).
__global__ __launch_bounds__(256) void page_scatter_kernel(const PageCopy* tab, const uint2* src, uint2* dst) {
    const PageCopy c = tab[blockIdx.x];
    const uint2* s = src + (size_t)c.src * 5;
    uint2* d = dst + (size_t)c.dst * 5;
    for (uint32_t w = threadIdx.x; w < c.nodes * 5u; w += 256) d[w] = s[w];
}

// Many scattered runs: gather the pages into pinned memory, send them with
// one copy and scatter them on the GPU. Returns YX_OK once the copy is queued
// (the caller synchronises before the staging is reused).
static int upload_scattered(yx_renderer* r, const yx_scene* s, const std::vector<std::pair<size_t, size_t>>& runs,
                            uint32_t pages, uint64_t bytes, hipStream_t st) {
    const size_t tab_bytes = ((size_t)pages * sizeof(PageCopy) + 255) & ~(size_t)255;
    const size_t need = tab_bytes + bytes;
    if (need > r->stage_cap) {
        const size_t cap = std::max(need, r->stage_cap + r->stage_cap / 2);
        if (r->stage_host) YX_HIP(hipHostFree(r->stage_host));
        if (r->d_stage) YX_HIP(hipFree(r->d_stage));
        r->stage_host = nullptr;
        r->d_stage = nullptr;
        r->stage_cap = 0;
        if (hipHostMalloc((void**)&r->stage_host, cap, hipHostMallocDefault) != hipSuccess) {
            r->stage_host = nullptr;
            return fail(YX_ERR_OOM, "hipHostMalloc(%zu) for page staging", cap);
        }
        if (hipMalloc((void**)&r->d_stage, cap) != hipSuccess) {
            r->d_stage = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%zu) for page staging", cap);
        }
        r->stage_cap = cap;
    }
    PageCopy* tab = reinterpret_cast<PageCopy*>(r->stage_host);
    uint8_t* data = r->stage_host + tab_bytes;
    uint32_t k = 0;
    size_t at = 0;  // nodes staged so far
    for (const auto& run : runs)
        for (size_t n0 = run.first; n0 < run.second; n0 += kScenePageNodes) {
            const size_t n1 = std::min(run.second, n0 + kScenePageNodes);
            tab[k++] = PageCopy{(uint32_t)n0, (uint32_t)at, (uint32_t)(n1 - n0), 0u};
            std::memcpy(data + at * 40, s->nodes.data() + n0 * 10, (n1 - n0) * 40);
            at += n1 - n0;
        }
    if (k != pages || at * 40 != bytes) return fail(YX_ERR_HIP, "page staging: %u of %u pages", k, pages);
    YX_HIP(hipMemcpyAsync(r->d_stage, r->stage_host, need, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(page_scatter_kernel, dim3(pages), dim3(256), 0, st,
                       reinterpret_cast<const PageCopy*>(r->d_stage),
                       reinterpret_cast<const uint2*>(r->d_stage + tab_bytes), reinterpret_cast<uint2*>(r->d_vox));
    YX_HIP(hipGetLastError());
    return YX_OK;
}

extern "C" int yx_renderer_count_changed_pages(const yx_renderer* r, const yx_scene* s, uint32_t* pages,
                                               uint64_t* bytes) {
    clear_error();
    if (!r || !s || !pages || !bytes) return fail(YX_ERR_INVALID_ARG, "null argument");
    dirty_pages(r, s, nullptr, pages, bytes);
    return YX_OK;
}

extern "C" int yx_renderer_update_scene(yx_renderer* r, const yx_scene* s, void* stream) {
    clear_error();
    if (!r || !s) return fail(YX_ERR_INVALID_ARG, "null argument");
    int rc = ensure_device(r);
    if (rc != YX_OK) return rc;
    const hipStream_t st = (hipStream_t)stream;
    const size_t nn = s->nodes.size() / 10;
    std::vector<std::pair<size_t, size_t>> runs;
    uint32_t pages = 0;
    uint64_t bytes = 0;
    dirty_pages(r, s, &runs, &pages, &bytes);
    const bool same = r->mirror_uid == s->uid;
    if (nn > r->vox_cap) {
        const size_t cap = std::max(nn, r->vox_cap + r->vox_cap / 2);
        uint32_t* d = nullptr;
        if (hipMalloc((void**)&d, cap * 40u) != hipSuccess)
            return fail(YX_ERR_OOM, "hipMalloc(%zu) for the scene's node copy", cap * 40u);
        if (same && r->vox_n) {
            const hipError_t e = hipMemcpyAsync(d, r->d_vox, r->vox_n * 40u, hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) {
                (void)hipFree(d);
                return fail(YX_ERR_HIP, "copying the node copy: %s", hipGetErrorString(e));
            }
            YX_HIP(hipStreamSynchronize(st));
        }
        if (r->d_vox) YX_HIP(hipFree(r->d_vox));
        r->d_vox = d;
        r->vox_cap = cap;
    }
    // until the copy below completes the device copy is in no known state
    r->mirror_uid = 0;
    if (runs.size() >= 8 && bytes <= (64ull << 20)) {
        if ((rc = upload_scattered(r, s, runs, pages, bytes, st)) != YX_OK) return rc;
    } else {
        for (const auto& run : runs)
            YX_HIP(hipMemcpyAsync(r->d_vox + run.first * 10, s->nodes.data() + run.first * 10,
                                  (run.second - run.first) * 40u, hipMemcpyHostToDevice, st));
    }
    // the copies read the scene's host memory: complete them before returning
    YX_HIP(hipStreamSynchronize(st));
    r->vox_n = nn;
    r->mirror_uid = s->uid;
    r->mirror_version = s->version;
    r->last_pages = pages;
    r->last_bytes = bytes;
    if (!r->rl_work && !(r->rl_work = relayout_work_create())) return fail(YX_ERR_OOM, "out of host memory");
    GpuLayout L;
    uint32_t* recs = reinterpret_cast<uint32_t*>(r->d_recs);
    uint32_t* lod = r->d_lod;
    size_t cap = r->d_recs ? r->recs_cap : 0;
    rc = gpu_relayout(r->d_vox, (uint32_t)nn, s->root, (void*)st, r->rl_work, &recs, &lod, &cap, L);
    r->d_recs = reinterpret_cast<uint2*>(recs);
    r->d_lod = lod;
    r->recs_cap = cap;
    if (rc != YX_OK) return rc;  // format errors leave the previous records in place
    if (L.empty) {
        // nothing to hit: no records (the launchers treat a null record array as empty)
        if (r->d_recs) YX_HIP(hipFree(r->d_recs));
        if (r->d_lod) YX_HIP(hipFree(r->d_lod));
        r->d_recs = nullptr;
        r->d_lod = nullptr;
        r->recs_cap = 0;
    }
    r->n_recs = L.empty ? 0 : L.n_recs;
    r->depth = L.empty ? 0 : L.depth;
    r->node_levels = L.empty ? 0 : L.node_levels;
    if (L.empty)
        r->level_records.clear();
    else
        r->level_records = L.level_records;
    r->has_scene = true;
    return YX_OK;
}

extern "C" int yx_renderer_last_update(const yx_renderer* r, uint32_t* pages, uint64_t* bytes) {
    clear_error();
    if (!r || !pages || !bytes) return fail(YX_ERR_INVALID_ARG, "null argument");
    *pages = r->last_pages;
    *bytes = r->last_bytes;
    return YX_OK;
}

extern "C" int yx_renderer_layout(const yx_renderer* r, uint32_t* recs, uint32_t* lod, uint64_t cap,
                                  uint64_t* n_recs) {
    clear_error();
    if (!r || !n_recs) return fail(YX_ERR_INVALID_ARG, "null argument");
    *n_recs = r->n_recs;
    if (!recs && !lod) return YX_OK;
    if (cap < r->n_recs) return fail(YX_ERR_RANGE, "capacity %llu below %zu records", (unsigned long long)cap, r->n_recs);
    if (r->n_recs == 0) return YX_OK;
    if (recs) YX_HIP(hipMemcpy(recs, r->d_recs, r->n_recs * 8u, hipMemcpyDeviceToHost));
    if (lod) YX_HIP(hipMemcpy(lod, r->d_lod, r->n_recs * 4u, hipMemcpyDeviceToHost));
    return YX_OK;
}
