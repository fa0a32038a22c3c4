// Persistent SVO ray-cast kernel for gfx950.
//
// Work distribution: one atomic ray queue per slice of the image (workgroup b
// serves queue b % nq; with nq = 8 the waves of one XCD share one L2 over a
// contiguous band of tiles), with stealing from the other queues once a queue
// is drained. Each wave keeps 64 independent rays in flight; when at least
// `refill_min` lanes have finished (or none is running) the wave ballots its
// idle lanes, takes that many ray ids with one atomic, and restarts them in
// place. A primary ray that hits turns into its pixel's shadow ray. The last
// wave to retire resets the queues for the next launch.
//
// Traversal state without a per-level t stack. The reference recursion
// (RecTrace, cell/ppu_renderer.cpp:18-41) keeps each level's (t1, t2) on the
// call stack. Here a lane keeps the current level's cell, a window of the W
// nearest ancestors' cells in registers, each level's node record in LDS, and
// per level the FindFirstChild result and current child (3 bits each). A pop
// beyond the window rebuilds the ancestor cell exactly by replaying, per axis
// and from the root interval, the operations the descent performed: midpoint
// 0.5f*(t1+t2), then the lower half (t1, tm), the upper half chosen by
// FindFirstChild (tm, t2), or the upper half reached by GoNext
// (tm, tm + (tm - t1)). These are the same float operations in the same order
// as FindFirstChild / GoNext (cell/spu/trace_spu.cpp:48-93), so the rebuilt
// cell is bit-identical to the one the recursion would hold.
#pragma once

#include "trace_device.h"

namespace yx {

enum : int { kNeed = 0, kPrimary = 1, kShadow = 2, kRetired = 3 };

constexpr uint32_t kNoRoot = 0xFFFFFFFFu;  // job whose scene has no nodes: every ray misses

// One (scene, camera) job of a batched launch. The scenes share one record
// pool (node records hold absolute pool indices), so a job only names its
// root record; its rays are the tiles [tile_begin, next job's tile_begin) of
// the launch's ray space and its outputs the pixels from pix_begin on.
struct BatchJob {
    float pos[3];
    float basis[9];
    float light[3];
    int light_mode;
    uint32_t root;
    int w, h, tiles_x;
    uint32_t pix_begin;
    uint32_t pad[3];
};
static_assert(sizeof(BatchJob) == 96, "BatchJob layout");

struct PersistArgs {
    const uint2* recs;
    float pos[3];
    float basis[9];
    float light[3];
    int light_mode;
    int payload;  // YX_PAYLOAD_*: how leaf payloads are shaded (a batch's scenes share it)
    int w, h;
    yx_outputs out;
    uint32_t* counters;  // [0, nq): queue heads; [nq]: retired-wave count
    uint32_t total;      // padded ray count (whole tiles)
    uint32_t per_q;      // rays per queue (multiple of the tile size)
    uint32_t n_waves;    // waves in the launch
    int nq;
    int tile_wl, tile_hl;  // log2 tile width / height
    int tiles_x;
    int refill_min;
    int levels;  // LDS rows per lane (node levels of the tree)
    // stack kernel: records [0, lds_recs) (the top tree levels) staged in LDS
    uint32_t lds_recs;
    // batched launches only
    const BatchJob* jobs;
    const uint32_t* job_tiles;  // first tile of each job, n_jobs + 1 entries (last = all tiles)
    int n_jobs;
    // LOD trace of the shading pipeline (stack kernel, LOD = true)
    const uint32_t* lod;  // mean leaf colour per node record
    float detail;         // stop when 2^-level < entry t * detail
    float* z;             // view depth t / |dir0 + du x + dv y|, +inf on a miss
    uint32_t* albedo;     // leaf or LOD colour, 0 on a miss
    uint8_t* hit_level;   // level of the hit cell, 0 on a miss
};

struct WinEntry {
    Cell c;
    uint2 rec;
};

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// One axis of the exact replay: the child interval of slot bit `up` given the
// parent interval and the FindFirstChild bit `ff`.
__device__ __forceinline__ void replay_axis(float& t1, float& t2, bool up, bool ff) {
    const float tm = 0.5f * (t1 + t2);
    const float d = tm - t1;
    const float g = tm + d;
    const float n1 = up ? tm : t1;
    const float n2 = up ? (ff ? t2 : g) : tm;
    t1 = n1;
    t2 = n2;
}

// The same step when the FindFirstChild bits are not kept: they follow from
// the node cell itself, as FindFirstChild computed them (tEnter = max(t1)
// against each midpoint, the same float operations), so the result is
// bit-identical.
__device__ __forceinline__ void replay_level_ffc(Cell& s, uint32_t h) {
    const float te = max3(s.t1x, s.t1y, s.t1z);
    const float mx = 0.5f * (s.t1x + s.t2x), my = 0.5f * (s.t1y + s.t2y), mz = 0.5f * (s.t1z + s.t2z);
    const uint32_t f = (te > mx ? 1u : 0u) | (te > my ? 2u : 0u) | (te > mz ? 4u : 0u);
    replay_axis(s.t1x, s.t2x, (h & 1u) != 0, (f & 1u) != 0);
    replay_axis(s.t1y, s.t2y, (h & 2u) != 0, (f & 2u) != 0);
    replay_axis(s.t1z, s.t2z, (h & 4u) != 0, (f & 4u) != 0);
}

__device__ __forceinline__ void replay_level(Cell& s, uint32_t f, uint32_t h) {
    replay_axis(s.t1x, s.t2x, (h & 1u) != 0, (f & 1u) != 0);
    replay_axis(s.t1y, s.t2y, (h & 2u) != 0, (f & 2u) != 0);
    replay_axis(s.t1z, s.t2z, (h & 4u) != 0, (f & 4u) != 0);
}

__device__ __forceinline__ uint32_t bits3(uint64_t path, int level) {
    return (uint32_t)(path >> (3 * level)) & 7u;
}

__device__ __forceinline__ uint64_t set3(uint64_t path, int level, uint32_t v) {
    const int s = 3 * level;
    return (path & ~((uint64_t)7 << s)) | ((uint64_t)v << s);
}

template <int W, int BLOCK>
__global__ __launch_bounds__(BLOCK) void persist_kernel(PersistArgs a) {
    extern __shared__ uint2 sh_rec[];  // [levels][BLOCK]
    const int tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint2* __restrict__ recs = a.recs;

    int q = (int)(blockIdx.x % (uint32_t)a.nq);
    int qleft = a.nq;  // queues not yet found drained (wave-uniform)

    int status = kNeed;
    bool fresh = false;
    uint32_t pix = 0;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, dx = 1.0f, dy = 1.0f, dz = 1.0f;
    Cell root, c;
    root.t1x = root.t1y = root.t1z = root.t2x = root.t2y = root.t2z = 0.0f;
    c = root;
    uint32_t flags = 0;
    int level = 0;
    int ch = 0;
    uint2 node = make_uint2(0u, 0u);
    uint64_t chpath = 0, ffcpath = 0;
    WinEntry win[W > 0 ? W : 1];
    int wcnt = 0;
    uint32_t visits = 0;
    uint32_t color = 0;  // primary hit payload, shaded when the shadow ray ends
    float lam = 0.0f;

    for (;;) {
        // ---- refill idle lanes -------------------------------------------------
        const uint64_t need = __ballot(status == kNeed);
        const uint64_t busy = __ballot(status == kPrimary || status == kShadow);
        if (need != 0 && (busy == 0 || __popcll(need) >= a.refill_min)) {
            if (qleft == 0) {
                if (status == kNeed) status = kRetired;
            } else {
                const uint32_t cnt = (uint32_t)__popcll(need);
                const int leader = __ffsll((unsigned long long)need) - 1;
                uint32_t got = 0;
                if ((int)lane == leader) got = atomicAdd(&a.counters[q], cnt);
                got = __shfl(got, leader);
                const uint32_t qbeg = (uint32_t)q * a.per_q;
                const uint32_t qend = min(a.total, qbeg + a.per_q);
                if (qbeg + got + cnt >= qend) {
                    q = (q + 1 == a.nq) ? 0 : q + 1;
                    --qleft;
                }
                if (status == kNeed && qbeg + got < qend) {
                    const uint32_t my = qbeg + got + lanes_below(need);
                    if (my < qend) {
                        const int tl = a.tile_wl + a.tile_hl;
                        const uint32_t tile = my >> tl, within = my & ((1u << tl) - 1u);
                        const int tx = (int)(tile % (uint32_t)a.tiles_x);
                        const int ty = (int)(tile / (uint32_t)a.tiles_x);
                        const int x = (tx << a.tile_wl) + (int)(within & ((1u << a.tile_wl) - 1u));
                        const int y = (ty << a.tile_hl) + (int)(within >> a.tile_wl);
                        if (x < a.w && y < a.h) {
                            pix = (uint32_t)y * (uint32_t)a.w + (uint32_t)x;
                            const Ray3 d = pixel_dir(a.basis, x, y);
                            dx = adjust_coord(d.x);
                            dy = adjust_coord(d.y);
                            dz = adjust_coord(d.z);
                            ox = a.pos[0];
                            oy = a.pos[1];
                            oz = a.pos[2];
                            visits = 0;
                            status = kPrimary;
                            fresh = true;
                        }
                    }
                }
            }
        }
        if (__ballot(status != kRetired) == 0) break;

        // ---- start fresh rays (primary or shadow) -------------------------------
        if (fresh) {
            fresh = false;
            const float p[3] = {ox, oy, oz};
            Ray3 d;
            d.x = dx;
            d.y = dy;
            d.z = dz;
            bool ok = recs != nullptr && setup_trace(p, d, root, flags) &&
                      min3(root.t2x, root.t2y, root.t2z) > 0;
            if (ok) {
                ++visits;
                c = root;
                node = recs[0];
                level = 0;
                wcnt = 0;
                ch = find_first_child(c);
                ffcpath = (uint64_t)(uint32_t)ch;
                chpath = 0;
            } else if (status == kPrimary) {
                // primary miss
                if (a.out.hit_id) a.out.hit_id[pix] = YX_MISS;
                if (a.out.voxel) {
                    a.out.voxel[3 * (size_t)pix] = -1;
                    a.out.voxel[3 * (size_t)pix + 1] = -1;
                    a.out.voxel[3 * (size_t)pix + 2] = -1;
                }
                if (a.out.t) a.out.t[pix] = -1.0f;
                if (a.out.rgba) a.out.rgba[pix] = 0u;
                if (a.out.shadow) a.out.shadow[pix] = 2;
                if (a.out.visits) a.out.visits[pix] = visits;
                status = kNeed;
            } else {
                // shadow ray that misses the cube: lit
                if (a.out.rgba) a.out.rgba[pix] = shade_rgba(color, 0.1f + 0.9f * lam);
                if (a.out.shadow) a.out.shadow[pix] = 0;
                if (a.out.visits) a.out.visits[pix] = visits;
                status = kNeed;
            }
        }

        // ---- one traversal step ---------------------------------------------------
        if (status == kPrimary || status == kShadow) {
            const int real = ch ^ (int)flags;
            const uint32_t cm = node.y & 0xffu, lm = (node.y >> 8) & 0xffu;
            const uint32_t below = (cm | lm) & ((1u << real) - 1u);
            const float mt2 = min3(c.t2x, c.t2y, c.t2z);
            const bool is_leaf = (lm >> real) & 1u;
            if (is_leaf && (status == kPrimary || mt2 > 0)) {
                if (status == kShadow) {
                    // occluded
                    if (a.out.rgba) a.out.rgba[pix] = shade_rgba(color, 0.1f);
                    if (a.out.shadow) a.out.shadow[pix] = 1;
                    if (a.out.visits) a.out.visits[pix] = visits;
                    status = kNeed;
                } else {
                    const uint2 leaf = recs[node.x + (uint32_t)__builtin_popcount(below)];
                    // voxel coordinates from the path of child slots
                    int vx = 0, vy = 0, vz = 0;
                    for (int j = 0; j < level; ++j) {
                        const uint32_t s = bits3(chpath, j) ^ flags;
                        vx = (vx << 1) | (int)(s & 1u);
                        vy = (vy << 1) | (int)((s >> 1) & 1u);
                        vz = (vz << 1) | (int)((s >> 2) & 1u);
                    }
                    vx = (vx << 1) | (real & 1);
                    vy = (vy << 1) | ((real >> 1) & 1);
                    vz = (vz << 1) | ((real >> 2) & 1);
                    const float t = max3(c.t1x, c.t1y, c.t1z);
                    if (a.out.hit_id) a.out.hit_id[pix] = leaf.y * 8u + (uint32_t)real;
                    if (a.out.voxel) {
                        a.out.voxel[3 * (size_t)pix] = vx;
                        a.out.voxel[3 * (size_t)pix + 1] = vy;
                        a.out.voxel[3 * (size_t)pix + 2] = vz;
                    }
                    if (a.out.t) a.out.t[pix] = t;
                    const int e = entry_axis(c);
                    const float de = e == 0 ? dx : (e == 1 ? dy : dz);
                    if (a.light_mode == YX_LIGHT_HEADLIGHT) {
                        lam = fabsf(de);
                    } else {
                        const float le = e == 0 ? a.light[0] : (e == 1 ? a.light[1] : a.light[2]);
                        lam = dmax((de < 0) ? le : -le, 0.0f);
                    }
                    color = leaf.x;
                    if (a.payload == YX_PAYLOAD_C16N16) {  // the leaf's own colour and normal
                        lam = a.light_mode == YX_LIGHT_HEADLIGHT
                                  ? lambert_c16n16(leaf.x, -dx, -dy, -dz)
                                  : lambert_c16n16(leaf.x, a.light[0], a.light[1], a.light[2]);
                        color = rgba_from_565(leaf.x & 0xffffu);
                    }
                    if (a.light_mode == YX_LIGHT_SUN_SHADOW && lam > 0.0f) {
                        // hit point stepped half a cell back across the entry
                        // face (a - b == a + (-b) exactly)
                        const float half = __builtin_bit_cast(float, (uint32_t)(127 - (level + 2)) << 23);
                        const float off = (de < 0) ? half : -half;
                        float px = ox + dx * t, py = oy + dy * t, pz = oz + dz * t;
                        px = (e == 0) ? px + off : px;
                        py = (e == 1) ? py + off : py;
                        pz = (e == 2) ? pz + off : pz;
                        ox = px;
                        oy = py;
                        oz = pz;
                        dx = adjust_coord(a.light[0]);
                        dy = adjust_coord(a.light[1]);
                        dz = adjust_coord(a.light[2]);
                        status = kShadow;
                        fresh = true;
                    } else {
                        if (a.out.rgba) a.out.rgba[pix] = shade_rgba(color, 0.1f + 0.9f * lam);
                        if (a.out.shadow) a.out.shadow[pix] = 2;
                        if (a.out.visits) a.out.visits[pix] = visits;
                        status = kNeed;
                    }
                }
            } else if (!is_leaf && ((cm >> real) & 1u) && mt2 > 0) {
                // descend into the child node
                ++visits;
                sh_rec[level * BLOCK + tid] = node;
                chpath = set3(chpath, level, (uint32_t)ch);
                if (W > 0) {
#pragma unroll
                    for (int k = (W > 0 ? W : 1) - 1; k > 0; --k) win[k] = win[k - 1];
                    win[0].c = c;
                    win[0].rec = node;
                    wcnt = min(wcnt + 1, W);
                }
                node = recs[node.x + (uint32_t)__builtin_popcount(below)];
                ++level;
                ch = find_first_child(c);
                ffcpath = set3(ffcpath, level, (uint32_t)ch);
            } else {
                // next sibling, popping finished levels
                while (!go_next(ch, c)) {
                    if (level == 0) {
                        if (status == kPrimary) {
                            if (a.out.hit_id) a.out.hit_id[pix] = YX_MISS;
                            if (a.out.voxel) {
                                a.out.voxel[3 * (size_t)pix] = -1;
                                a.out.voxel[3 * (size_t)pix + 1] = -1;
                                a.out.voxel[3 * (size_t)pix + 2] = -1;
                            }
                            if (a.out.t) a.out.t[pix] = -1.0f;
                            if (a.out.rgba) a.out.rgba[pix] = 0u;
                            if (a.out.shadow) a.out.shadow[pix] = 2;
                        } else {
                            if (a.out.rgba) a.out.rgba[pix] = shade_rgba(color, 0.1f + 0.9f * lam);
                            if (a.out.shadow) a.out.shadow[pix] = 0;
                        }
                        if (a.out.visits) a.out.visits[pix] = visits;
                        status = kNeed;
                        break;
                    }
                    --level;
                    ch = (int)bits3(chpath, level);
                    if (W > 0 && wcnt > 0) {
                        c = win[0].c;
                        node = win[0].rec;
#pragma unroll
                        for (int k = 0; k + 1 < (W > 0 ? W : 1); ++k) win[k] = win[k + 1];
                        --wcnt;
                    } else {
                        // exact replay from the root interval to `level`; the
                        // cells of the W levels above it refill the window
                        Cell s = root;
                        for (int j = 0; j <= level; ++j) {
                            replay_level(s, bits3(ffcpath, j), bits3(chpath, j));
#pragma unroll
                            for (int k = 0; k < W; ++k)
                                if (j == level - 1 - k) win[k].c = s;
                        }
                        c = s;
                        node = sh_rec[level * BLOCK + tid];
                        if (W > 0) {
                            wcnt = min(level, W);
#pragma unroll
                            for (int k = 0; k < W; ++k)
                                if (k < wcnt) win[k].rec = sh_rec[(level - 1 - k) * BLOCK + tid];
                        }
                    }
                }
            }
        }
    }

    // last wave out resets the queues for the next launch
    if (lane == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&a.counters[a.nq], 1u);
        if (done + 1 == a.n_waves) {
            for (int k = 0; k < a.nq; ++k) atomicExch(&a.counters[k], 0u);
            atomicExch(&a.counters[a.nq], 0u);
            __threadfence();
        }
    }
}

}  // namespace yx
