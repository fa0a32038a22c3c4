// Batched SVO ray casting: many scenes x many cameras in one persistent launch.
//
// The reference renders one scene and one view per call (SVORenderer::Render,
// demo/SVORenderer.cpp:95-149, driven per view by demo/Demo.cpp:138-174). Here
// every scene of the batch is re-laid out into one shared record pool (node
// records hold absolute pool indices, so the traversal loop is the
// single-scene one) and every (scene, camera) job owns a range of ray tiles in
// one ray space that the persistent stack kernel drains from its atomic
// queues. Per job the device table holds the camera basis, light, root record
// and output offset. A frame is replayed from a hipGraph captured once per
// (outputs, tuning, job layout).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>

#include "internal.h"
#include "timing.h"
#include "trace_stack.h"

using namespace yx;

#define YX_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(YX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

static constexpr int kBatchQueues = 64;

struct yx_batch {
    // scenes
    uint2* d_pool = nullptr;
    size_t n_pool = 0;
    std::vector<uint32_t> roots;     // pool index of each scene's root, or kNoRoot
    std::vector<int> node_levels;    // per scene
    // jobs
    std::vector<yx_batch_job> jobs;
    std::vector<uint64_t> pix_begin;
    uint64_t total_pixels = 0;
    BatchJob* d_jobs = nullptr;
    uint32_t* d_job_tiles = nullptr;
    size_t jobs_cap = 0;
    bool table_dirty = true;  // device job table out of date
    uint32_t total_tiles = 0;
    int table_tile_w = 0, table_tile_h = 0;
    int levels = 1;  // stack rows: deepest node level of any job's scene
    uint64_t layout_gen = 1;  // changes with scenes, job scenes / resolutions, tile shape, payload format
    int payload = YX_PAYLOAD_RGBA8;  // how every scene's leaf payloads are read
    // device
    uint32_t* d_counters = nullptr;
    int device = -1, num_cus = 0;
    size_t max_lds = 0;
    yx_tuning tune = {YX_KERNEL_PERSISTENT, 64, 8, 8, 64, 1, 0, 0, 0};
    // graph of one frame
    bool use_graph = true;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    yx_outputs g_out{};
    yx_tuning g_tune{};
    uint64_t g_layout = 0;
};

static void drop_graph(yx_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->g_layout = 0;
}

static int batch_device(yx_batch* b) {
    if (b->device >= 0) return YX_OK;
    int dev = 0, cus = 0, lds = 0;
    YX_HIP(hipGetDevice(&dev));
    YX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    YX_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    hipError_t e = hipMalloc((void**)&b->d_counters, (kBatchQueues + 1) * sizeof(uint32_t));
    if (e != hipSuccess) {
        b->d_counters = nullptr;
        return fail(YX_ERR_OOM, "hipMalloc counters: %s", hipGetErrorString(e));
    }
    YX_HIP(hipMemset(b->d_counters, 0, (kBatchQueues + 1) * sizeof(uint32_t)));
    YX_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    b->device = dev;
    b->num_cus = cus;
    b->max_lds = (size_t)lds;
    return YX_OK;
}

extern "C" int yx_batch_create(yx_batch** out) {
    clear_error();
    if (!out) return fail(YX_ERR_INVALID_ARG, "null argument");
    yx_batch* b = new (std::nothrow) yx_batch();
    if (!b) return fail(YX_ERR_OOM, "out of host memory");
    *out = b;
    return YX_OK;
}

extern "C" void yx_batch_destroy(yx_batch* b) {
    if (!b) return;
    drop_graph(b);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->d_pool) (void)hipFree(b->d_pool);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_job_tiles) (void)hipFree(b->d_job_tiles);
    if (b->d_counters) (void)hipFree(b->d_counters);
    delete b;
}

extern "C" int yx_batch_set_scenes(yx_batch* b, const yx_scene* const* scenes, int32_t n) {
    clear_error();
    if (!b || (!scenes && n > 0) || n < 0) return fail(YX_ERR_INVALID_ARG, "null argument");
    for (int32_t i = 0; i < n; ++i)
        if (!scenes[i]) return fail(YX_ERR_INVALID_ARG, "scene %d is null", i);
    int rc = batch_device(b);
    if (rc != YX_OK) return rc;
    // re-layout every scene at its offset in the pool
    std::vector<GpuTree> trees((size_t)n);
    std::vector<uint32_t> roots((size_t)n, kNoRoot);
    std::vector<int> levels((size_t)n, 0);
    uint64_t total = 0;
    try {
        for (int32_t i = 0; i < n; ++i) {
            if (total > 0xFFFFFFF0ull) return fail(YX_ERR_RANGE, "more than 2^32 records in the batch");
            rc = build_gpu_tree(*scenes[i], trees[(size_t)i], (uint32_t)total);
            if (rc != YX_OK) return rc;
            if (!trees[(size_t)i].empty) {
                roots[(size_t)i] = (uint32_t)total;
                levels[(size_t)i] = trees[(size_t)i].node_levels;
                total += trees[(size_t)i].recs.size() / 2;
            }
        }
    } catch (const std::bad_alloc&) {
        return fail(YX_ERR_OOM, "out of host memory in the re-layout");
    }
    if (total > 0xFFFFFFF0ull) return fail(YX_ERR_RANGE, "more than 2^32 records in the batch");
    YX_HIP(hipDeviceSynchronize());  // earlier frames may still read the old pool
    drop_graph(b);
    if (b->d_pool) {
        YX_HIP(hipFree(b->d_pool));
        b->d_pool = nullptr;
    }
    b->n_pool = 0;
    if (total) {
        hipError_t e = hipMalloc((void**)&b->d_pool, total * sizeof(uint2));
        if (e != hipSuccess) {
            b->d_pool = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%llu) for the scene pool: %s",
                        (unsigned long long)(total * sizeof(uint2)), hipGetErrorString(e));
        }
        for (int32_t i = 0; i < n; ++i) {
            const GpuTree& t = trees[(size_t)i];
            if (t.empty) continue;
            YX_HIP(hipMemcpy(b->d_pool + roots[(size_t)i], t.recs.data(), t.recs.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice));
        }
    }
    b->n_pool = (size_t)total;
    b->roots = std::move(roots);
    b->node_levels = std::move(levels);
    b->jobs.clear();
    b->pix_begin.clear();
    b->total_pixels = 0;
    b->table_dirty = true;
    ++b->layout_gen;
    return YX_OK;
}

extern "C" int yx_batch_set_jobs(yx_batch* b, const yx_batch_job* jobs, int32_t n, uint64_t* total_pixels) {
    clear_error();
    if (!b || (!jobs && n > 0) || n < 0) return fail(YX_ERR_INVALID_ARG, "null argument");
    uint64_t total = 0;
    std::vector<uint64_t> begin((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        const yx_batch_job& j = jobs[i];
        if (j.scene < 0 || (size_t)j.scene >= b->roots.size())
            return fail(YX_ERR_INVALID_ARG, "job %d: scene %d not in the batch (%zu scenes)", i, j.scene,
                        b->roots.size());
        if (j.width <= 0 || j.height <= 0 || (int64_t)j.width * j.height > ((int64_t)1 << 28))
            return fail(YX_ERR_INVALID_ARG, "job %d: resolution %dx%d", i, j.width, j.height);
        if (!(j.fov > 0.0f && j.fov < 180.0f)) return fail(YX_ERR_INVALID_ARG, "job %d: fov %g", i, j.fov);
        if (j.light_mode < YX_LIGHT_HEADLIGHT || j.light_mode > YX_LIGHT_SUN_SHADOW)
            return fail(YX_ERR_INVALID_ARG, "job %d: light mode %d", i, j.light_mode);
        begin[(size_t)i] = total;
        total += (uint64_t)j.width * (uint64_t)j.height;
    }
    if (total >= 0xFFFFFFFFull) return fail(YX_ERR_RANGE, "%llu output pixels: the batch allows < 2^32",
                                            (unsigned long long)total);
    bool same_layout = (size_t)n == b->jobs.size();
    for (int32_t i = 0; same_layout && i < n; ++i)
        same_layout = jobs[i].scene == b->jobs[(size_t)i].scene && jobs[i].width == b->jobs[(size_t)i].width &&
                      jobs[i].height == b->jobs[(size_t)i].height;
    b->jobs.assign(jobs, jobs + n);
    b->pix_begin = std::move(begin);
    b->total_pixels = total;
    b->table_dirty = true;
    if (!same_layout) ++b->layout_gen;
    if (total_pixels) *total_pixels = total;
    return YX_OK;
}

extern "C" int yx_batch_job_offset(const yx_batch* b, int32_t job, uint64_t* first_pixel) {
    clear_error();
    if (!b || !first_pixel) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (job < 0 || (size_t)job >= b->jobs.size()) return fail(YX_ERR_INVALID_ARG, "job %d out of range", job);
    *first_pixel = b->pix_begin[(size_t)job];
    return YX_OK;
}

static int check_batch_tuning(const yx_tuning* t) {
    auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
    if (t->kernel != YX_KERNEL_PERSISTENT) return fail(YX_ERR_INVALID_ARG, "batches use the persistent stack kernel");
    if (t->block != 64 && t->block != 128 && t->block != 256)
        return fail(YX_ERR_INVALID_ARG, "block must be 64, 128 or 256");
    if (!pow2(t->tile_w) || !pow2(t->tile_h) || t->tile_w * t->tile_h > 4096)
        return fail(YX_ERR_INVALID_ARG, "tile %dx%d: powers of two, at most 4096 pixels", t->tile_w, t->tile_h);
    if (t->refill_min < 1 || t->refill_min > 64) return fail(YX_ERR_INVALID_ARG, "refill_min must be 1..64");
    if (t->queues < 1 || t->queues > kBatchQueues)
        return fail(YX_ERR_INVALID_ARG, "queues must be 1..%d", kBatchQueues);
    if (t->window != 0 && t->window != 2 && t->window != 4)
        return fail(YX_ERR_INVALID_ARG, "window must be 0, 2 or 4");
    if (t->waves_per_cu < 0 || t->waves_per_cu > 32) return fail(YX_ERR_INVALID_ARG, "waves_per_cu must be 0..32");
    if (t->lds_levels != 0) return fail(YX_ERR_INVALID_ARG, "batches stage no tree levels in LDS (lds_levels 0)");
    return YX_OK;
}

extern "C" int yx_batch_set_tuning(yx_batch* b, const yx_tuning* t) {
    clear_error();
    if (!b || !t) return fail(YX_ERR_INVALID_ARG, "null argument");
    const int rc = check_batch_tuning(t);
    if (rc != YX_OK) return rc;
    if (t->tile_w != b->tune.tile_w || t->tile_h != b->tune.tile_h) b->table_dirty = true;
    b->tune = *t;
    return YX_OK;
}

extern "C" int yx_batch_get_tuning(const yx_batch* b, yx_tuning* t) {
    clear_error();
    if (!b || !t) return fail(YX_ERR_INVALID_ARG, "null argument");
    *t = b->tune;
    return YX_OK;
}

extern "C" int yx_batch_set_graph(yx_batch* b, int32_t use_graph) {
    clear_error();
    if (!b) return fail(YX_ERR_INVALID_ARG, "null argument");
    b->use_graph = use_graph != 0;
    return YX_OK;
}

extern "C" int yx_batch_set_payload(yx_batch* b, int32_t format) {
    clear_error();
    if (!b) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (format != YX_PAYLOAD_RGBA8 && format != YX_PAYLOAD_C16N16)
        return fail(YX_ERR_INVALID_ARG, "unknown payload format %d", format);
    if (format != b->payload) {
        b->payload = format;
        ++b->layout_gen;  // the captured launch holds the old format
    }
    return YX_OK;
}

extern "C" int yx_batch_scene_bytes(const yx_batch* b, uint64_t* bytes) {
    clear_error();
    if (!b || !bytes) return fail(YX_ERR_INVALID_ARG, "null argument");
    *bytes = (uint64_t)b->n_pool * sizeof(uint2);
    return YX_OK;
}

// Rebuilds the device job table for the current cameras and tile shape.
static int upload_jobs(yx_batch* b) {
    if (!b->table_dirty) return YX_OK;
    const int tw = b->tune.tile_w, th = b->tune.tile_h;
    const size_t n = b->jobs.size();
    std::vector<BatchJob> host(n);
    std::vector<uint32_t> tiles(n + 1);
    uint64_t t0 = 0;
    int levels = 1;
    for (size_t i = 0; i < n; ++i) {
        const yx_batch_job& j = b->jobs[i];
        BatchJob& d = host[i];
        std::memset(&d, 0, sizeof(d));
        for (int k = 0; k < 3; ++k) {
            d.pos[k] = j.pos[k];
            d.light[k] = j.light[k];
        }
        camera_basis(j.dir, j.up, j.fov, j.width, j.height, d.basis);
        d.light_mode = j.light_mode;
        d.root = b->roots[(size_t)j.scene];
        d.w = j.width;
        d.h = j.height;
        d.tiles_x = (j.width + tw - 1) / tw;
        d.pix_begin = (uint32_t)b->pix_begin[i];
        tiles[i] = (uint32_t)t0;
        t0 += (uint64_t)d.tiles_x * (uint64_t)((j.height + th - 1) / th);
        if (d.root != kNoRoot) levels = std::max(levels, b->node_levels[(size_t)j.scene] - 1);
    }
    if (t0 * (uint64_t)(tw * th) >= 0xFFFFFFFFull) return fail(YX_ERR_RANGE, "too many rays for one launch");
    tiles[n] = (uint32_t)t0;
    if (n + 1 > b->jobs_cap) {
        if (b->d_jobs) YX_HIP(hipFree(b->d_jobs));
        if (b->d_job_tiles) YX_HIP(hipFree(b->d_job_tiles));
        b->d_jobs = nullptr;
        b->d_job_tiles = nullptr;
        b->jobs_cap = 0;
        const size_t cap = std::max<size_t>(n + 1, 64);
        hipError_t e = hipMalloc((void**)&b->d_jobs, cap * sizeof(BatchJob));
        if (e == hipSuccess) e = hipMalloc((void**)&b->d_job_tiles, cap * sizeof(uint32_t));
        if (e != hipSuccess) return fail(YX_ERR_OOM, "hipMalloc job table: %s", hipGetErrorString(e));
        b->jobs_cap = cap;
        ++b->layout_gen;  // the table moved: captured launches point at the old one
    }
    YX_HIP(hipDeviceSynchronize());  // earlier frames may still read the table
    if (n) YX_HIP(hipMemcpy(b->d_jobs, host.data(), n * sizeof(BatchJob), hipMemcpyHostToDevice));
    YX_HIP(hipMemcpy(b->d_job_tiles, tiles.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (b->total_tiles != (uint32_t)t0 || b->table_tile_w != tw || b->table_tile_h != th || b->levels != levels)
        ++b->layout_gen;
    b->total_tiles = (uint32_t)t0;
    b->table_tile_w = tw;
    b->table_tile_h = th;
    b->levels = levels;
    b->table_dirty = false;
    return YX_OK;
}

// Launch configuration and kernel of one batched frame.
struct BatchLaunch {
    const void* fn;
    void (*kern)(PersistArgs);
    int block;
    uint32_t grid;
    size_t lds;
    PersistArgs a;
};

#define YX_BATCH_KERNEL(B, W)                                            \
    do {                                                                 \
        L.fn = (const void*)stack_kernel<B, W, true>;                    \
        L.kern = stack_kernel<B, W, true>;                               \
    } while (0)

static int plan_launch(yx_batch* b, const yx_outputs& out, BatchLaunch& L) {
    const yx_tuning& t = b->tune;
    PersistArgs& a = L.a;
    std::memset(&a, 0, sizeof(a));
    a.recs = b->d_pool;
    a.out = out;
    a.counters = b->d_counters;
    a.tile_wl = __builtin_ctz((unsigned)t.tile_w);
    a.tile_hl = __builtin_ctz((unsigned)t.tile_h);
    const uint32_t tile = (uint32_t)(t.tile_w * t.tile_h);
    a.total = b->total_tiles * tile;
    a.nq = t.queues;
    a.per_q = queue_rays(b->total_tiles, tile, t.queues);
    a.refill_min = t.refill_min;
    a.levels = b->levels;
    a.payload = b->payload;
    a.jobs = b->d_jobs;
    a.job_tiles = b->d_job_tiles;
    a.n_jobs = (int)b->jobs.size();
    a.tiles_x = 1;  // per job in the table
    a.w = a.h = 1;
    // W == 0 rows hold record indices below 2^29; larger pools use a window
    const int win = (t.window == 0 && b->n_pool >= kStackMaxRecords) ? 2 : t.window;
    const size_t lane_bytes = stack_lane_bytes(b->levels, win);
    int block = t.block;
    size_t lds = lane_bytes * (size_t)block;
    while (lds > b->max_lds && block > 64) {
        block /= 2;
        lds = lane_bytes * (size_t)block;
    }
    if (lds > b->max_lds)
        return fail(YX_ERR_RANGE, "traversal stack rows need %zu bytes of LDS, device allows %zu", lds, b->max_lds);
    if (block == 64) {
        if (win == 0) YX_BATCH_KERNEL(64, 0);
        else if (win == 2) YX_BATCH_KERNEL(64, 2);
        else YX_BATCH_KERNEL(64, 4);
    } else if (block == 128) {
        if (win == 0) YX_BATCH_KERNEL(128, 0);
        else if (win == 2) YX_BATCH_KERNEL(128, 2);
        else YX_BATCH_KERNEL(128, 4);
    } else {
        if (win == 0) YX_BATCH_KERNEL(256, 0);
        else if (win == 2) YX_BATCH_KERNEL(256, 2);
        else YX_BATCH_KERNEL(256, 4);
    }
    int per_cu = t.waves_per_cu > 0 ? std::max(1, t.waves_per_cu / (block / 64)) : 0;
    if (per_cu == 0) {
        int occ = 0;
        YX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, L.fn, block, lds));
        per_cu = std::max(1, occ);
    }
    L.block = block;
    L.lds = lds;
    L.grid = (uint32_t)(per_cu * b->num_cus);
    a.n_waves = L.grid * (uint32_t)(block / 64);
    return YX_OK;
}

static bool same_outputs(const yx_outputs& x, const yx_outputs& y) {
    return x.hit_id == y.hit_id && x.voxel == y.voxel && x.t == y.t && x.rgba == y.rgba && x.shadow == y.shadow &&
           x.visits == y.visits;
}

static bool same_tuning(const yx_tuning& x, const yx_tuning& y) { return std::memcmp(&x, &y, sizeof(x)) == 0; }

static int render_batch(yx_batch* b, const yx_outputs& out, hipStream_t stream) {
    if (b->jobs.empty()) return YX_OK;
    int rc = batch_device(b);
    if (rc != YX_OK) return rc;
    rc = upload_jobs(b);
    if (rc != YX_OK) return rc;
    if (b->total_tiles == 0) return YX_OK;
    if (b->use_graph && b->exec && b->g_layout == b->layout_gen && same_outputs(b->g_out, out) &&
        same_tuning(b->g_tune, b->tune)) {
        YX_HIP(hipGraphLaunch(b->exec, stream));
        return YX_OK;
    }
    BatchLaunch L;
    rc = plan_launch(b, out, L);
    if (rc != YX_OK) return rc;
    if (!b->use_graph) {
        hipLaunchKernelGGL(L.kern, dim3(L.grid), dim3(L.block), L.lds, stream, L.a);
        YX_HIP(hipGetLastError());
        return YX_OK;
    }
    drop_graph(b);
    // capture the frame's launch on the library's own stream, replay on the caller's
    YX_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(L.kern, dim3(L.grid), dim3(L.block), L.lds, b->cap_stream, L.a);
    const hipError_t le = hipGetLastError();
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(b->cap_stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(YX_ERR_HIP, "graph capture: %s", hipGetErrorString(le != hipSuccess ? le : ce));
    }
    b->graph = g;
    YX_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->g_out = out;
    b->g_tune = b->tune;
    b->g_layout = b->layout_gen;
    YX_HIP(hipGraphLaunch(b->exec, stream));
    return YX_OK;
}

extern "C" int yx_render_batch(yx_batch* b, const yx_outputs* out, void* stream) {
    clear_error();
    if (!b || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    return render_batch(b, *out, (hipStream_t)stream);
}

extern "C" int yx_batch_autotune(yx_batch* b, const yx_outputs* out, int32_t reps, yx_tune_trial* trials,
                                 int32_t cap, int32_t* n_trials, void* stream) {
    clear_error();
    if (!b || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (reps < 1) return fail(YX_ERR_INVALID_ARG, "reps must be >= 1");
    if (b->jobs.empty()) return fail(YX_ERR_INVALID_ARG, "no jobs to tune on");
    const yx_tuning keep = b->tune;
    const hipStream_t st = (hipStream_t)stream;
    auto time_ms = [&](const yx_tuning& t, float* ms) -> int {
        int rc = yx_batch_set_tuning(b, &t);
        if (rc != YX_OK) return rc;
        return time_frames([&]() { return render_batch(b, *out, st); }, reps, st, ms);
    };
    yx_tuning best = keep;
    const int rc = tune_coordinate_descent(keep, time_ms, trials, cap, n_trials, &best);
    const int rs = yx_batch_set_tuning(b, rc == YX_OK ? &best : &keep);
    return rc != YX_OK ? rc : rs;
}
