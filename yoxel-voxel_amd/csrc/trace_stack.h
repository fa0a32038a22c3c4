// Persistent SVO ray-cast kernel with a per-lane traversal stack.
//
// Work distribution: atomic ray queues (workgroup b serves queue b % nq and
// steals from the others once it drains); the last wave to retire resets the
// queues for the next launch.
//
// Each lane runs one ray at a time: a primary ray, then, after a lit hit, its
// pixel's shadow ray. The wave loop does one traversal step per tracing lane
// per iteration - test the current child and descend, or GoNext, or pop one
// level - with the exact state the reference recursion keeps on its call stack
// (RecTrace, cell/ppu_renderer.cpp:18-41): per level the cell (t1, t2) and the
// node record. Without a register window (W = 0, the default) LDS holds every
// level's node record and the cells of even levels; a pop to an odd level
// rebuilds the cell from the level above with one exact replay step
// (FindFirstChild / GoNext float operations, as in trace_persist.h). With
// W > 0 the W most recent levels sit in registers (a shift window) and the
// older ones in 32-byte LDS rows. Everything that happens once per ray - taking
// ray ids, ray setup with its nine divisions, hit decoding, shading, the shadow
// ray spawn, output writes - is parked and done for many lanes at once, only
// when at least `refill_min` lanes wait (or none is tracing), so the wave does
// not pay for one lane's rare event in most iterations.
#pragma once

#include "trace_persist.h"

namespace yx {

// Wave-cooperative refill: idle lanes of the wave take consecutive ray ids
// from queue `q` with one atomic. Returns true in a lane that got a pixel.
struct QueueCursor {
    int q;
    int left;  // queues not yet found drained (wave-uniform)
};

// Queue q owns the groups of 2^kQueueGroupLog consecutive tiles whose group
// index is q modulo the queue count, in order. Every queue therefore walks the
// whole image from the first tile row to the last at about the same pace (the
// waves of all queues share one region of the scene in the caches), while one
// queue's workgroups - one XCD's, since workgroup b serves queue b % nq and
// workgroups are dealt to the 8 XCDs in turn - write whole output lines: 16
// tiles of 8 pixels span 128 pixels of a row, a 128-byte line even for the
// one-byte shadow flags. One queue is the plain tile order.
constexpr int kQueueGroupLog = 4;

// Host side: rays per queue for `tiles` tiles of `tile` rays over nq queues.
inline uint32_t queue_rays(uint32_t tiles, uint32_t tile, int nq) {
    const uint32_t groups = (tiles + (1u << kQueueGroupLog) - 1u) >> kQueueGroupLog;
    const uint32_t per_q = (groups + (uint32_t)nq - 1u) / (uint32_t)nq;
    return (per_q << kQueueGroupLog) * tile;
}

// In a batched launch the ray id's tile is looked up in the job table (binary
// search over the jobs' first tiles) and x, y are within that job's image.
template <bool BATCH>
__device__ __forceinline__ bool take_ray(const PersistArgs& a, QueueCursor& qc, uint64_t need, bool idle,
                                         uint32_t lane, int& job, int& x, int& y) {
    const uint32_t cnt = (uint32_t)__popcll(need);
    const int leader = __ffsll((unsigned long long)need) - 1;
    uint32_t got = 0;
    if ((int)lane == leader) got = atomicAdd(&a.counters[qc.q], cnt);
    got = __shfl(got, leader);
    const uint32_t q = (uint32_t)qc.q;
    if (got + cnt >= a.per_q) {
        qc.q = (qc.q + 1 == a.nq) ? 0 : qc.q + 1;
        --qc.left;
    }
    if (!idle || got >= a.per_q) return false;
    const uint32_t pos = got + lanes_below(need);  // ray position within queue q
    if (pos >= a.per_q) return false;
    const int tl = a.tile_wl + a.tile_hl;
    const uint32_t tq = pos >> tl;  // tile position within the queue
    uint32_t tile = ((((tq >> kQueueGroupLog) * (uint32_t)a.nq) + q) << kQueueGroupLog) |
                    (tq & ((1u << kQueueGroupLog) - 1u));
    if (tile >= (a.total >> tl)) return false;  // padding of the last group
    const uint32_t within = pos & ((1u << tl) - 1u);
    int tiles_x = a.tiles_x, w = a.w, h = a.h;
    if constexpr (BATCH) {
        // job_tiles[lo] <= tile < job_tiles[hi]
        int lo = 0, hi = a.n_jobs;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (a.job_tiles[mid] <= tile)
                lo = mid;
            else
                hi = mid;
        }
        job = lo;
        tile -= a.job_tiles[lo];
        tiles_x = a.jobs[lo].tiles_x;
        w = a.jobs[lo].w;
        h = a.jobs[lo].h;
    }
    const int tx = (int)(tile % (uint32_t)tiles_x);
    const int ty = (int)(tile / (uint32_t)tiles_x);
    x = (tx << a.tile_wl) + (int)(within & ((1u << a.tile_wl) - 1u));
    y = (ty << a.tile_hl) + (int)(within >> a.tile_wl);
    return x < w && y < h;
}

__device__ __forceinline__ void write_miss(const yx_outputs& o, uint32_t pix, uint32_t visits) {
    if (o.hit_id) o.hit_id[pix] = YX_MISS;
    if (o.voxel) {
        o.voxel[3 * (size_t)pix] = -1;
        o.voxel[3 * (size_t)pix + 1] = -1;
        o.voxel[3 * (size_t)pix + 2] = -1;
    }
    if (o.t) o.t[pix] = -1.0f;
    if (o.rgba) o.rgba[pix] = 0u;
    if (o.shadow) o.shadow[pix] = 2;
    if (o.visits) o.visits[pix] = visits;
}

__device__ __forceinline__ void write_miss_lod(const PersistArgs& a, uint32_t pix, uint32_t visits) {
    if (a.out.t) a.out.t[pix] = -1.0f;
    if (a.z) a.z[pix] = __builtin_inff();
    if (a.albedo) a.albedo[pix] = 0u;
    if (a.hit_level) a.hit_level[pix] = 0;
    if (a.out.visits) a.out.visits[pix] = visits;
}

__device__ __forceinline__ void write_final(const yx_outputs& o, uint32_t pix, uint32_t color, float k,
                                            uint8_t shadow, uint32_t visits) {
    if (o.rgba) o.rgba[pix] = shade_rgba(color, k);
    if (o.shadow) o.shadow[pix] = shadow;
    if (o.visits) o.visits[pix] = visits;
}

// LDS bytes per lane of stack_kernel<BLOCK, W, ...> when `levels` levels are
// pushed: W == 0 keeps every level's node record index (4 B) and the cells
// (24 B) of even levels; W > 0 keeps 32-byte rows for the levels older than
// the window.
// Record indices the W == 0 stack rows hold (index << 3 | child slot).
constexpr size_t kStackMaxRecords = (size_t)1 << 29;

inline size_t stack_lane_bytes(int levels, int W) {
    if (W == 0) return (size_t)4 * (size_t)levels + (size_t)24 * (size_t)((levels + 1) / 2);
    return (size_t)32 * (size_t)(levels - W > 1 ? levels - W : 1);
}

// lane states
enum : int {
    sIdle = 0,     // needs a ray id
    sStartP = 1,   // primary ray assigned, not set up
    sStartS = 2,   // shadow ray assigned, not set up
    sTraceP = 3,   // tracing a primary ray
    sTraceS = 4,   // tracing a shadow ray
    sHitP = 5,     // primary ray stopped on a leaf
    sMissP = 6,    // primary ray left the cube
    sEndS = 7,     // shadow ray finished (occluded flag)
    sRetired = 8,
};

// W: stack entries of the W levels above the current one held in registers
// (a shift window); older levels live in LDS rows indexed by level.
// BATCH: rays come from the (scene, camera) jobs of a.jobs (camera, light,
// root record and output offset per job) instead of the single view in `a`.
// STAGE: node records [0, a.lds_recs) are read from an LDS copy (single scene).
// LOD: the trace of the shading pipeline - primary rays only, a node below the
// root whose cell edge 2^-level is smaller than its entry distance times
// a.detail (and which has a non-null child) ends the ray with its mean colour
// a.lod[record]; outputs t, z, albedo, hit level and visits.
template <int BLOCK, int W, bool BATCH, bool STAGE = false, bool LOD = false>
__global__ __launch_bounds__(BLOCK) void stack_kernel(PersistArgs a) {
    // dynamic LDS: the staged top-level records (single scene), then the
    // traversal stack rows
    extern __shared__ uint4 sh_mem[];
    uint2* sh_rec = reinterpret_cast<uint2*>(sh_mem);
    uint4* sh_stk = sh_mem + ((a.lds_recs * 8u + 15u) >> 4);
    const int tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint2* __restrict__ recs = a.recs;

    QueueCursor qc;
    qc.q = (int)(blockIdx.x % (uint32_t)a.nq);
    qc.left = a.nq;

    // records [0, lds_recs) are the top tree levels (breadth-first order)
    if constexpr (STAGE) {
        for (uint32_t i = (uint32_t)tid; i < a.lds_recs; i += BLOCK) sh_rec[i] = recs[i];
        __syncthreads();
    }
    auto rec_at = [&](uint32_t i) -> uint2 {
        if constexpr (!STAGE) {
            return recs[i];
        } else {
            uint2 v;
            if (i < a.lds_recs)
                v = sh_rec[i];
            else
                v = recs[i];
            return v;
        }
    };

    int status = sIdle;
    bool adv = false, occluded = false;
    uint32_t pix = 0;
    int job = 0;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, dx = 1.0f, dy = 1.0f, dz = 1.0f;
    Cell c;
    c.t1x = c.t1y = c.t1z = c.t2x = c.t2y = c.t2z = 0.0f;
    uint32_t flags = 0;
    int level = 0, ch = 0;
    uint2 node = make_uint2(0u, 0u);
    uint64_t chpath = 0;
    uint32_t visits = 0, color = 0;
    float lam = 0.0f;
    WinEntry win[W > 0 ? W : 1];
    float qlen = 1.0f;      // LOD: |dir0 + du x + dv y| of the lane's pixel
    bool lod_stop = false;  // LOD: the ray ended on a node
    uint32_t lod_rec = 0;   // LOD: that node's record

    // W == 0: every level's node record index and child slot as one word,
    // index << 3 | slot ([levels][BLOCK] uint32; a pop reads the record again,
    // mostly from the cache that served it on the way down; record indices
    // stay below 2^29), and the cell of even levels only ([(levels + 1) / 2]
    // [BLOCK] as uint4 + uint2). A pop to an odd level rebuilds its cell from
    // the level above by one exact replay step, so a lane needs 4 B + 12 B per
    // level instead of 32 B.
    uint32_t cur = 0;      // record index of `node` (W == 0)
    // W == 0: bit l set when level l of the current path can still move to a
    // sibling once its child is exhausted, i.e. when the exit axis of the
    // level's cell (argmin t2, GoNext's tie order) is clear in its child slot.
    // The recursion hands the child a copy of (t1, t2) and finds its own cell
    // unchanged on return, so the bit, taken when the level descends, is the
    // outcome of the GoNext the recursion runs there after the child returns.
    // A failed GoNext therefore returns in one step to the deepest such level
    // and advances there; the levels in between would have failed in turn.
    uint32_t advmask = 0;
    uint32_t* sh_idx = reinterpret_cast<uint32_t*>(sh_stk);
    uint4* sh_ca = reinterpret_cast<uint4*>(sh_idx + (size_t)a.levels * BLOCK);
    uint2* sh_cb = reinterpret_cast<uint2*>(sh_ca + (size_t)((a.levels + 1) >> 1) * BLOCK);
    auto push0 = [&](int lv, const Cell& cc, uint32_t ri, int slot) {
        sh_idx[(size_t)lv * BLOCK + tid] = (ri << 3) | (uint32_t)slot;
        if ((lv & 1) == 0) {
            const size_t r = (size_t)(lv >> 1) * BLOCK + tid;
            sh_ca[r] = make_uint4(__float_as_uint(cc.t1x), __float_as_uint(cc.t1y), __float_as_uint(cc.t1z),
                                  __float_as_uint(cc.t2x));
            sh_cb[r] = make_uint2(__float_as_uint(cc.t2y), __float_as_uint(cc.t2z));
        }
    };
    auto pop0 = [&](int lv, Cell& cc, uint2& nn, int& slot) {
        const uint32_t v = sh_idx[(size_t)lv * BLOCK + tid];
        cur = v >> 3;
        slot = (int)(v & 7u);
        nn = rec_at(cur);
        const size_t r = (size_t)(lv >> 1) * BLOCK + tid;
        const uint4 r0 = sh_ca[r];
        const uint2 r1 = sh_cb[r];
        cc.t1x = __uint_as_float(r0.x);
        cc.t1y = __uint_as_float(r0.y);
        cc.t1z = __uint_as_float(r0.z);
        cc.t2x = __uint_as_float(r0.w);
        cc.t2y = __uint_as_float(r1.x);
        cc.t2z = __uint_as_float(r1.y);
        // odd level: the row holds the cell of level lv - 1, i.e. this level's
        // node interval; replay FindFirstChild / GoNext to the current child
        if (lv & 1) replay_level_ffc(cc, (uint32_t)slot);
    };
    auto lds_store = [&](int row, const Cell& cc, uint2 nn) {
        uint4* p = sh_stk + (size_t)(2 * row) * BLOCK + tid;
        p[0] = make_uint4(__float_as_uint(cc.t1x), __float_as_uint(cc.t1y), __float_as_uint(cc.t1z),
                          __float_as_uint(cc.t2x));
        p[BLOCK] = make_uint4(__float_as_uint(cc.t2y), __float_as_uint(cc.t2z), nn.x, nn.y);
    };
    auto lds_load = [&](int row, Cell& cc, uint2& nn) {
        const uint4* p = sh_stk + (size_t)(2 * row) * BLOCK + tid;
        const uint4 r0 = p[0], r1 = p[BLOCK];
        cc.t1x = __uint_as_float(r0.x);
        cc.t1y = __uint_as_float(r0.y);
        cc.t1z = __uint_as_float(r0.z);
        cc.t2x = __uint_as_float(r0.w);
        cc.t2y = __uint_as_float(r1.x);
        cc.t2z = __uint_as_float(r1.y);
        nn = make_uint2(r1.z, r1.w);
    };

    for (;;) {
        const bool tracing = status == sTraceP || status == sTraceS;
        const uint64_t tr_mask = __ballot(tracing);
        const uint64_t wait_mask = __ballot(!tracing && status != sRetired);
        if (wait_mask != 0 && (tr_mask == 0 || __popcll(wait_mask) >= a.refill_min)) {
            // ---- batched once-per-ray work ----------------------------------
            if (LOD && status == sHitP) {
                const float t = max3(c.t1x, c.t1y, c.t1z);
                uint32_t col;
                int lv = level;
                if (lod_stop) {
                    col = a.lod[lod_rec];
                } else {
                    const int real = ch ^ (int)flags;
                    const uint32_t cm = node.y & 0xffu, lm = (node.y >> 8) & 0xffu;
                    const uint32_t below = (cm | lm) & ((1u << real) - 1u);
                    col = rec_at(node.x + (uint32_t)__builtin_popcount(below)).x;
                    lv = level + 1;
                }
                if (a.out.t) a.out.t[pix] = t;
                if (a.z) a.z[pix] = t / qlen;
                if (a.albedo) a.albedo[pix] = col;
                if (a.hit_level) a.hit_level[pix] = (uint8_t)lv;
                if (a.out.visits) a.out.visits[pix] = visits;
                status = sIdle;
            } else if (LOD && status == sMissP) {
                write_miss_lod(a, pix, visits);
                status = sIdle;
            } else if (status == sHitP) {
                const int real = ch ^ (int)flags;
                const uint32_t cm = node.y & 0xffu, lm = (node.y >> 8) & 0xffu;
                const uint32_t below = (cm | lm) & ((1u << real) - 1u);
                const uint2 leaf = rec_at(node.x + (uint32_t)__builtin_popcount(below));
                int vx = 0, vy = 0, vz = 0;
                for (int j = 0; j < level; ++j) {
                    uint32_t s;
                    if constexpr (W == 0)
                        s = (sh_idx[(size_t)j * BLOCK + tid] & 7u) ^ flags;
                    else
                        s = bits3(chpath, j) ^ flags;
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
                const float* light = a.light;
                int light_mode = a.light_mode;
                if constexpr (BATCH) {
                    light = a.jobs[job].light;
                    light_mode = a.jobs[job].light_mode;
                }
                if (light_mode == YX_LIGHT_HEADLIGHT) {
                    lam = fabsf(de);
                } else {
                    const float le = e == 0 ? light[0] : (e == 1 ? light[1] : light[2]);
                    lam = dmax((de < 0) ? le : -le, 0.0f);
                }
                color = leaf.x;
                if (a.payload == YX_PAYLOAD_C16N16) {  // the leaf's own colour and normal
                    lam = light_mode == YX_LIGHT_HEADLIGHT ? lambert_c16n16(leaf.x, -dx, -dy, -dz)
                                                           : lambert_c16n16(leaf.x, light[0], light[1], light[2]);
                    color = rgba_from_565(leaf.x & 0xffffu);
                }
                if (light_mode == YX_LIGHT_SUN_SHADOW && lam > 0.0f) {
                    // hit point stepped half a cell back across the entry face
                    const float half = __builtin_bit_cast(float, (uint32_t)(127 - (level + 2)) << 23);
                    const float off = (de < 0) ? half : -half;
                    const float px = ox + dx * t, py = oy + dy * t, pz = oz + dz * t;
                    ox = (e == 0) ? px + off : px;
                    oy = (e == 1) ? py + off : py;
                    oz = (e == 2) ? pz + off : pz;
                    dx = adjust_coord(light[0]);
                    dy = adjust_coord(light[1]);
                    dz = adjust_coord(light[2]);
                    status = sStartS;
                } else {
                    write_final(a.out, pix, color, 0.1f + 0.9f * lam, 2, visits);
                    status = sIdle;
                }
            } else if (status == sMissP) {
                write_miss(a.out, pix, visits);
                status = sIdle;
            } else if (status == sEndS) {
                if (occluded)
                    write_final(a.out, pix, color, 0.1f, 1, visits);
                else
                    write_final(a.out, pix, color, 0.1f + 0.9f * lam, 0, visits);
                status = sIdle;
            }
            // shadow rays and primary rays go in separate packets: while this
            // block starts shadow rays, the idle lanes wait instead of taking
            // the next primary rays, so a packet is either the shadow rays of
            // one packet's lit hits or a whole run of 64 primary rays (a full
            // 8x8 tile); mixed packets ran 4 % longer on the 1080p frame
            const uint64_t need = __ballot(status == sStartS) != 0 ? 0ull : __ballot(status == sIdle);
            if (need != 0) {
                if (qc.left == 0) {
                    if (status == sIdle) status = sRetired;
                } else {
                    int x = 0, y = 0;
                    if (take_ray<BATCH>(a, qc, need, status == sIdle, lane, job, x, y)) {
                        const float* basis = a.basis;
                        const float* pos = a.pos;
                        if constexpr (BATCH) {
                            const BatchJob& jb = a.jobs[job];
                            pix = jb.pix_begin + (uint32_t)y * (uint32_t)jb.w + (uint32_t)x;
                            basis = jb.basis;
                            pos = jb.pos;
                        } else {
                            pix = (uint32_t)y * (uint32_t)a.w + (uint32_t)x;
                        }
                        Ray3 d;
                        if constexpr (LOD) {
                            d = pixel_dir_n(basis, x, y, qlen);
                            lod_stop = false;
                        } else {
                            d = pixel_dir(basis, x, y);
                        }
                        dx = adjust_coord(d.x);
                        dy = adjust_coord(d.y);
                        dz = adjust_coord(d.z);
                        ox = pos[0];
                        oy = pos[1];
                        oz = pos[2];
                        visits = 0;
                        status = sStartP;
                    }
                }
            }
            if (status == sStartP || status == sStartS) {
                const float p[3] = {ox, oy, oz};
                Ray3 d;
                d.x = dx;
                d.y = dy;
                d.z = dz;
                uint32_t root = 0;
                if constexpr (BATCH) root = a.jobs[job].root;
                const bool ok = recs != nullptr && root != kNoRoot && setup_trace(p, d, c, flags) &&
                                min3(c.t2x, c.t2y, c.t2z) > 0;
                if (ok) {
                    ++visits;
                    node = rec_at(root);
                    if constexpr (W == 0) cur = root;
                    level = 0;
                    adv = false;
                    ch = find_first_child(c);
                    status = status == sStartP ? sTraceP : sTraceS;
                } else if (status == sStartP) {
                    if constexpr (LOD)
                        write_miss_lod(a, pix, visits);
                    else
                        write_miss(a.out, pix, visits);
                    status = sIdle;
                } else {
                    write_final(a.out, pix, color, 0.1f + 0.9f * lam, 0, visits);  // lit
                    status = sIdle;
                }
            }
        }
        if (__ballot(status != sRetired) == 0) break;
        if (status != sTraceP && status != sTraceS) continue;

        // ---- one traversal step ---------------------------------------------
        if (!adv) {
            const int real = ch ^ (int)flags;
            const uint32_t cm = node.y & 0xffu, lm = (node.y >> 8) & 0xffu;
            const float mt2 = min3(c.t2x, c.t2y, c.t2z);
            const bool is_leaf = (lm >> real) & 1u;
            if (is_leaf && (status == sTraceP || mt2 > 0)) {
                if (status == sTraceP) {
                    status = sHitP;
                } else {
                    occluded = true;
                    status = sEndS;
                }
                continue;
            }
            if (!is_leaf && ((cm >> real) & 1u) && mt2 > 0) {
                // push this level, descend into the child node
                ++visits;
                if constexpr (W > 0) {
                    if (level >= W) lds_store(level - W, win[W - 1].c, win[W - 1].rec);
#pragma unroll
                    for (int k = W - 1; k > 0; --k) win[k] = win[k - 1];
                    win[0].c = c;
                    win[0].rec = node;
                    chpath = set3(chpath, level, (uint32_t)ch);
                } else {
                    push0(level, c, cur, ch);
                    const uint32_t bit = 1u << level;
                    advmask = (((uint32_t)ch >> exit_axis(c)) & 1u) ? (advmask & ~bit) : (advmask | bit);
                }
                const uint32_t below = (cm | lm) & ((1u << real) - 1u);
                const uint32_t crec = node.x + (uint32_t)__builtin_popcount(below);
                node = rec_at(crec);
                if constexpr (W == 0) cur = crec;
                ++level;
                if constexpr (LOD) {
                    // c is this node's cell; 2^-level as a float is exact
                    const float edge = __builtin_bit_cast(float, (uint32_t)(127 - level) << 23);
                    if (node.y != 0u && edge < max3(c.t1x, c.t1y, c.t1z) * a.detail) {
                        lod_stop = true;
                        lod_rec = crec;
                        status = sHitP;
                        continue;
                    }
                }
                ch = find_first_child(c);
                continue;
            }
            adv = true;
        }
        if (go_next(ch, c)) {
            adv = false;
            continue;
        }
        // the current level is exhausted: W == 0 returns to the deepest level
        // that can advance and advances there; W > 0 pops one level
        uint32_t up = 0;
        if constexpr (W == 0) {
            up = advmask & ((1u << level) - 1u);
        } else {
            up = level;
        }
        if (up == 0) {
            if (status == sTraceP) {
                status = sMissP;
            } else {
                occluded = false;
                status = sEndS;
            }
        } else if constexpr (W == 0) {
            level = 31 - __builtin_clz(up);
            pop0(level, c, node, ch);
            go_next(ch, c);  // succeeds: the level's exit axis is clear in ch
            adv = false;
        } else {
            --level;
            c = win[0].c;
            node = win[0].rec;
#pragma unroll
            for (int k = 0; k + 1 < W; ++k) win[k] = win[k + 1];
            if (level >= W) lds_load(level - W, win[W - 1].c, win[W - 1].rec);
            ch = (int)bits3(chpath, level);
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
