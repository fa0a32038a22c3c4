// Differentiable front-to-back compositing over brick voxels (gfx950).
//
// Forward model (reaction/trace/trace.cu:42-139 castRay / walkBrick, in the
// form of the thesis recurrence reaction/report/voxel.tex:59-87):
//   ray: pixel direction of RendererBase::InitRayDir (cell/renderer_base.h:
//        50-61), components below 1e-5 clamped with their sign (PrepareRay,
//        reaction/trace/trace.cu:21,31-33); unit-cube slab test with 1/dir
//        (intersect, :66-80), t_enter clamped at 0 (:104).
//   samples: t_k = k * dt with dt = one voxel (walkBrick's lattice
//        ceil(t_enter / dt) * dt, :44-48), t_enter <= t_k < t_exit; voxel =
//        floor(p * N) clamped to the grid; samples in unallocated bricks are
//        empty and skipped (castRay's mark test, :129).
//   per sample: sigma = softplus(raw), alpha = 1 - exp(-sigma),
//        C += T * alpha * rgb, T *= exp(-sigma) (walkBrick's accum += col *
//        (1 - accum.w), :58). No early exit.
// Adjoint, per sample k with g = dL/dC, S = C_final - sum_{j<=k} w_j rgb_j:
//   dL/drgb_k   = w_k g,   w_k = T_k alpha_k
//   dL/dsigma_k = g . (T_{k+1} rgb_k - S)
//   dL/draw_k   = dL/dsigma_k * sigmoid(raw_k)
// The backward recomputes the exact discrete sample sequence instead of
// storing per-ray lists, writes one (voxel key, float4 gradient) record per
// sample at the ray's exclusive-scan offset, and sorts the records' ids by
// the brick slot bits of the key with the stable LSD radix sort of prims.hip
// (8-bit digits; per-block digit histograms, a scan, a scatter that ranks
// equal digits in record order with wave ballots). One wave per slot then sums
// each voxel's records in emission order in LDS and writes the brick. Bricks
// above 8^3 sort by the whole key and sum each key's run instead. No atomics
// on gradients; the result is deterministic. The record total stays on the
// device: the workspace has a capacity, and a total beyond it is flagged and
// turns the gradient into NaN instead of a silently wrong one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>

#include "internal.h"
#include "prims.h"
#include "trace_device.h"

namespace yx {

// Words of the backward's small device block: record total, records used
// (min(total, capacity)), overflow flag, distinct voxels.
enum : int { kMiscTotal = 0, kMiscN = 1, kMiscOverflow = 2, kMiscDistinct = 3, kMiscWords = 64 };

struct CompArgs {
    const int32_t* table;
    const float4* params;
    const float* views;
    const uint8_t* super;  // per super-brick of 2^ls bricks per axis: 1 when any of its bricks is allocated
    uint32_t num_slots;
    int n, lb, g;  // grid, log2 brick, bricks per axis
    int lg;        // log2 of g
    int ls;        // log2 of the super-brick edge in bricks (min(2, lg))
    int n_views, w, h;
    float* rgb;
    float* trans;
    uint32_t* nsamp;
    // backward
    const float* rgb_in;
    const float* grad_rgb;
    const uint32_t* offsets;
    const uint32_t* nsamp_in;
    uint32_t total_records;
    uint32_t n_waves;
    uint32_t* keys;

    float4* vals;
};

typedef float V4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t lanes_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

struct CompRay {
    float ox, oy, oz, dx, dy, dz, rx, ry, rz;  // origin, direction, 1/direction
    int k, kend;                               // sample range [k, kend)
};

__device__ __forceinline__ float clamp_eps(float x) {
    const float eps = 1e-5f;
    return fabsf(x) < eps ? copysignf(eps, x) : x;
}

__device__ __forceinline__ bool comp_setup(const CompArgs& a, int view, int x, int y, CompRay& r) {
    const float* v = a.views + 12 * view;
    const Ray3 d = pixel_dir(v + 3, x, y);
    r.ox = v[0];
    r.oy = v[1];
    r.oz = v[2];
    r.dx = clamp_eps(d.x);
    r.dy = clamp_eps(d.y);
    r.dz = clamp_eps(d.z);
    r.rx = 1.0f / r.dx;
    r.ry = 1.0f / r.dy;
    r.rz = 1.0f / r.dz;
    const float lx = r.rx * (0.0f - r.ox), hx = r.rx * (1.0f - r.ox);
    const float ly = r.ry * (0.0f - r.oy), hy = r.ry * (1.0f - r.oy);
    const float lz = r.rz * (0.0f - r.oz), hz = r.rz * (1.0f - r.oz);
    float te = fmaxf(fminf(lx, hx), fmaxf(fminf(ly, hy), fminf(lz, hz)));
    const float tx = fminf(fmaxf(lx, hx), fminf(fmaxf(ly, hy), fmaxf(lz, hz)));
    te = fmaxf(te, 0.0f);
    if (tx < 0.0f || te > tx) return false;
    const float nf = (float)a.n;
    r.k = (int)ceilf(te * nf);
    r.kend = (int)ceilf(tx * nf);
    return r.k < r.kend;
}

__device__ __forceinline__ int vox(float o, float d, float t, float nf, int n) {
    const float p = o + t * d;
    int v = (int)floorf(p * nf);
    return min(max(v, 0), n - 1);
}

// Walk over the samples of one ray that fall in allocated bricks, in order,
// one step per call: step() either jumps over the samples of one empty brick
// or yields one sample (returns true with its parameter index and float4
// parameters); `live` turns false at the end of the ray. One step per call
// keeps the lanes of a wave in one flat loop: a lane crossing empty space does
// not hold up the others' samples.
// SMALL: at most 1024 bricks per axis (a power of two, 2^lg), so the brick
// index (bz << 2 lg | by << lg | bx, < 2^30) fits 32 bits and costs two
// shift-or instructions; otherwise it is formed with 64-bit multiplies.
template <bool SMALL>
struct SampleIter {
    using BrickId = typename std::conditional<SMALL, uint32_t, int64_t>::type;
    CompRay r;
    int k;
    bool live;

    BrickId cb;  // brick of the last table lookup (all ones / -1: none)
    int cslot;   // its slot
    uint32_t csb;  // super-brick of the last summary lookup (all ones: none)
    uint32_t cs;   // its summary byte (0: every brick in it is empty)

    // Advance by one step without loading the sample's parameters: returns
    // true with the parameter index of a sample, false after a skip or at the
    // end. The sequence does not depend on parameter values, so callers can
    // take several steps and then load their parameters together.
    __device__ __forceinline__ bool next(const CompArgs& a, uint32_t& idx) {
        idx = 0;
        if (!live) return false;
        if (k >= r.kend) {
            live = false;
            return false;
        }
        const float nf = (float)a.n, inv_n = 1.0f / nf;
        const int bmask = (1 << a.lb) - 1;
        const float t = (float)k * inv_n;
        const int vx = vox(r.ox, r.dx, t, nf, a.n), vy = vox(r.oy, r.dy, t, nf, a.n), vz = vox(r.oz, r.dz, t, nf, a.n);
        const int bx = vx >> a.lb, by = vy >> a.lb, bz = vz >> a.lb;
        BrickId bl;
        if constexpr (SMALL)
            bl = ((uint32_t)bz << (2 * a.lg)) | ((uint32_t)by << a.lg) | (uint32_t)bx;
        else
            bl = ((int64_t)bz * a.g + by) * a.g + bx;
        if (bl != cb) {  // consecutive samples mostly share a brick
            cb = bl;
            // the super-brick summary first (a small array that stays in the
            // cache): inside an empty super-brick the table is not read
            const int gs_l = a.lg - a.ls;
            const uint32_t sb = ((uint32_t)(bz >> a.ls) << (2 * gs_l)) | ((uint32_t)(by >> a.ls) << gs_l) |
                                (uint32_t)(bx >> a.ls);
            if (sb != csb) {
                csb = sb;
                cs = a.super[sb];
            }
            cslot = cs ? a.table[bl] : -1;
        }
        const int slot = cslot;
        if ((uint32_t)slot >= a.num_slots) {  // -1 (and anything out of range) is empty
            // jump to the last sample still in this brick, or in its whole
            // super-brick when that is empty (exact: the candidate is checked
            // and moved back while it left the box; the samples inside a box
            // are consecutive, since each voxel coordinate is monotonic in k)
            const int sh = cs ? a.lb : a.lb + a.ls;  // log2 of the box edge in voxels
            const int cx = vx >> sh, cy = vy >> sh, cz = vz >> sh;
            const float bsz = (float)(1 << sh) * inv_n;
            const float ex = ((r.dx > 0.0f ? (float)(cx + 1) : (float)cx) * bsz - r.ox) * r.rx;
            const float ey = ((r.dy > 0.0f ? (float)(cy + 1) : (float)cy) * bsz - r.oy) * r.ry;
            const float ez = ((r.dz > 0.0f ? (float)(cz + 1) : (float)cz) * bsz - r.oz) * r.rz;
            const float tb = fminf(ex, fminf(ey, ez));
            int kk = (int)ceilf(tb * nf) - 1;
            kk = min(max(kk, k), r.kend - 1);
            while (kk > k) {
                const float t2 = (float)kk * inv_n;
                const int wx = vox(r.ox, r.dx, t2, nf, a.n) >> sh, wy = vox(r.oy, r.dy, t2, nf, a.n) >> sh,
                          wz = vox(r.oz, r.dz, t2, nf, a.n) >> sh;
                if (wx == cx && wy == cy && wz == cz) break;
                --kk;
            }
            k = kk + 1;
            return false;
        }
        const uint32_t local = (uint32_t)((vx & bmask) + ((vy & bmask) << a.lb) + ((vz & bmask) << (2 * a.lb)));
        idx = ((uint32_t)slot << (3 * a.lb)) + local;
        ++k;
        return true;
    }

    __device__ __forceinline__ void start(const CompArgs& a, int view, int x, int y, bool valid) {
        live = valid && comp_setup(a, view, x, y, r);
        k = r.k;
        cb = (BrickId)-1;
        cslot = -1;
        csb = ~0u;
        cs = 1;
    }
};

// Steps taken per wave iteration in the forward and in the adjoint's
// emission; their parameter loads are issued together. Ten each measured
// best on composite_512 (forward 0.87 ms at 72 VGPRs against 0.95 ms with
// four and 0.94 ms with twelve; adjoint 3.04 against 3.18 ms with four).
constexpr int kCompBatchF = 10;
constexpr int kCompBatchB = 10;

// Opacity of one sample with sigma = softplus(raw), without forming sigma:
// alpha = 1 - exp(-sigma) = sigmoid(raw) and exp(-sigma) = sigmoid(-raw).
// With n = exp(-|raw|) in (0, 1] and r = 1 / (1 + n) = sigmoid(|raw|), one
// is r and the other n * r, so a sample costs one exponential and one
// reciprocal (no log1p / expm1, no cancellation for either sign of raw).
// alpha is also the chain factor dsigma/draw = sigmoid(raw) of the adjoint.
// Both are the hardware instructions (v_exp_f32 on -|raw| log2 e, v_rcp_f32
// on 1 + n in [1, 2], 1 ulp each): n carries a relative error below
// |raw| * 1e-7 from the rounded product, far inside the compositor's error
// budget against the FP64 model, and the forward and the adjoint share this
// function, so the adjoint's recomputed transmittance is the forward's to
// the bit.
__device__ __forceinline__ void opacity(float raw, float& alpha, float& keep) {
    const float n = __builtin_amdgcn_exp2f(-fabsf(raw) * 1.44269504f);
    const float r = __builtin_amdgcn_rcpf(1.0f + n);
    const float nr = n * r;
    alpha = raw >= 0.0f ? r : nr;
    keep = raw >= 0.0f ? nr : r;
}

// Thread t of a launch takes lane t % 64 of wave t / 64; each wave covers one
// 8x8 pixel tile (tiles row-major per view, views in order) so its rays stay
// coherent. Rays (outputs) are numbered view-major, then row-major.
__device__ __forceinline__ bool ray_of(const CompArgs& a, size_t t, int& view, int& x, int& y, size_t& ray) {
    const size_t wave = t >> 6;
    const int lane = (int)(t & 63);
    const int tx_n = (a.w + 7) >> 3, ty_n = (a.h + 7) >> 3;
    const size_t per_view = (size_t)tx_n * ty_n;
    if (wave >= per_view * (size_t)a.n_views) return false;
    view = (int)(wave / per_view);
    const int tile = (int)(wave - (size_t)view * per_view);
    const int ty = tile / tx_n, tx = tile - ty * tx_n;
    x = tx * 8 + (lane & 7);
    y = ty * 8 + (lane >> 3);
    if (x >= a.w || y >= a.h) return false;
    ray = ((size_t)view * a.h + y) * a.w + x;
    return true;
}

// Summary byte per super-brick (2^ls bricks per axis): 1 when any of its
// bricks holds a slot below num_slots. With 4^3 bricks per super-brick
// (ls = 2) a wave covers one super-brick, a lane per brick, and votes;
// smaller grids loop over the bricks in one thread.
__global__ __launch_bounds__(256) void comp_super_kernel(const int32_t* table, uint32_t num_slots, int lg, int ls,
                                                         uint8_t* super) {
    const int gs_l = lg - ls;
    const uint32_t nsb = 1u << (3 * gs_l);
    const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint32_t gmask = (1u << gs_l) - 1u;
    if (ls == 2) {
        if ((t >> 6) >= nsb) return;  // whole waves
        const uint32_t i = (uint32_t)(t >> 6), l = (uint32_t)(t & 63u);  // super-brick, brick within it
        const uint32_t sx = i & gmask, sy = (i >> gs_l) & gmask, sz = i >> (2 * gs_l);
        const uint64_t b = ((((uint64_t)(sz * 4 + (l >> 4))) << lg | (sy * 4 + ((l >> 2) & 3u))) << lg) |
                           (sx * 4 + (l & 3u));
        const uint64_t any = __ballot((uint32_t)table[b] < num_slots);
        if (l == 0) super[i] = any ? 1 : 0;
        return;
    }
    if (t >= nsb) return;
    const uint32_t sx = (uint32_t)t & gmask, sy = ((uint32_t)t >> gs_l) & gmask, sz = (uint32_t)t >> (2 * gs_l);
    const uint32_t e = 1u << ls;
    uint32_t any = 0;
    for (uint32_t z = 0; z < e; ++z)
        for (uint32_t y = 0; y < e; ++y)
            for (uint32_t x = 0; x < e; ++x) {
                const uint64_t b = ((((uint64_t)(sz * e + z)) << lg | (sy * e + y)) << lg) | (sx * e + x);
                any |= (uint32_t)table[b] < num_slots ? 1u : 0u;
            }
    super[t] = (uint8_t)any;
}

template <bool SMALL>
__global__ __launch_bounds__(256) void comp_forward_kernel(CompArgs a) {
    int view, x, y;
    size_t ray;
    if (!ray_of(a, (size_t)xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x, view, x, y, ray)) return;
    float cr = 0.0f, cg = 0.0f, cb = 0.0f, T = 1.0f;
    uint32_t n = 0;
    SampleIter<SMALL> it;
    it.start(a, view, x, y, true);
    // kCompBatchF steps per lane per iteration, their loads issued together;
    // the wave-wide vote keeps the loop flat (without it the skip steps get
    // folded into an inner loop that holds the wave until every lane has found
    // its next sample)
    for (;;) {
        uint32_t ix[kCompBatchF];
        bool got[kCompBatchF];
#pragma unroll
        for (int j = 0; j < kCompBatchF; ++j) got[j] = it.next(a, ix[j]);
        float4 q[kCompBatchF];
#pragma unroll
        for (int j = 0; j < kCompBatchF; ++j) q[j] = got[j] ? a.params[ix[j]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int j = 0; j < kCompBatchF; ++j) {
            if (got[j]) {
                float al, e;
                opacity(q[j].x, al, e);
                const float wgt = T * al;
                cr += wgt * q[j].y;
                cg += wgt * q[j].z;
                cb += wgt * q[j].w;
                T = T * e;
                ++n;
            }
        }
        if (__ballot(it.live) == 0) break;
    }
    if (a.rgb) {
        a.rgb[3 * ray] = cr;
        a.rgb[3 * ray + 1] = cg;
        a.rgb[3 * ray + 2] = cb;
    }
    if (a.trans) a.trans[ray] = T;
    if (a.nsamp) a.nsamp[ray] = n;
}

// Sum of the forward's per-ray sample counts over each wave's tile.
__global__ __launch_bounds__(256) void comp_wave_sums_kernel(CompArgs a, uint32_t* sums, uint32_t n_waves) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    int view, x, y;
    size_t ray = 0;
    const bool valid = ray_of(a, t, view, x, y, ray);
    uint32_t n = valid ? a.nsamp_in[ray] : 0u;
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if ((t & 63) == 0 && (t >> 6) < n_waves) sums[t >> 6] = n;
}

// One record per sample. A wave's records fill the contiguous range that the
// exclusive scan of the per-wave sample counts gives its tile; within it the
// lanes that emit in one iteration take consecutive slots (ballot rank), so
// stores coalesce and the record order is fixed by the inputs alone.
template <bool SMALL>
__global__ __launch_bounds__(256) void comp_backward_kernel(CompArgs a) {
    const size_t t = (size_t)xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const size_t wave = t >> 6;
    if (wave >= a.n_waves) return;  // whole wave past the end
    int view = 0, x = 0, y = 0;
    size_t ray = 0;
    const bool valid = ray_of(a, t, view, x, y, ray);
    // the forward's sample count: a ray without samples does not walk, and a
    // ray stops right after its last sample instead of walking the empty
    // bricks behind it (those steps emit nothing, so the records and their
    // positions are the same)
    uint32_t left = valid ? a.nsamp_in[ray] : 0u;
    SampleIter<SMALL> it;
    it.start(a, view, x, y, valid && left > 0u);
    float gr = 0.0f, gg = 0.0f, gb = 0.0f, fr = 0.0f, fg = 0.0f, fb = 0.0f;
    if (valid) {
        gr = a.grad_rgb[3 * ray];
        gg = a.grad_rgb[3 * ray + 1];
        gb = a.grad_rgb[3 * ray + 2];
        fr = a.rgb_in[3 * ray];
        fg = a.rgb_in[3 * ray + 1];
        fb = a.rgb_in[3 * ray + 2];
    }
    uint32_t cursor = a.offsets[wave];
    float pr = 0.0f, pg = 0.0f, pb = 0.0f, T = 1.0f;
    // kCompBatchB steps per iteration as in the forward; the records of step j
    // of every lane of the wave are appended (in lane order) before those of
    // step j + 1, so the record order is fixed
    for (;;) {
        uint32_t ix[kCompBatchB];
        bool got[kCompBatchB];
#pragma unroll
        for (int j = 0; j < kCompBatchB; ++j) {
            got[j] = it.next(a, ix[j]);
            if (got[j] && --left == 0u) it.live = false;
        }
        float4 q[kCompBatchB];
#pragma unroll
        for (int j = 0; j < kCompBatchB; ++j) q[j] = got[j] ? a.params[ix[j]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int j = 0; j < kCompBatchB; ++j) {
            const uint64_t mask = __ballot(got[j]);
            if (got[j]) {
                float al, e;
                opacity(q[j].x, al, e);
                const float wgt = T * al;
                pr += wgt * q[j].y;
                pg += wgt * q[j].z;
                pb += wgt * q[j].w;
                const float tn = T * e;
                const float sr = fr - pr, sg = fg - pg, sb = fb - pb;
                const float dsig = gr * (tn * q[j].y - sr) + gg * (tn * q[j].z - sg) + gb * (tn * q[j].w - sb);
                const uint32_t pos = cursor + lanes_rank(mask);
                if (pos < a.total_records) {
                    // streaming stores: the records are read back only by the
                    // sort, after this kernel, so they should not take the L2
                    // lines of the parameters and table this walk reads
                    const V4f v = {dsig * al, wgt * gr, wgt * gg, wgt * gb};
                    __builtin_nontemporal_store(ix[j], a.keys + pos);
                    __builtin_nontemporal_store(v, reinterpret_cast<V4f*>(a.vals) + pos);
                }
                T = tn;
            }
            cursor += (uint32_t)__popcll(mask);
        }
        if (__ballot(it.live) == 0) break;
    }
}

// Statistics only (not on the gradient path): number of distinct sorted keys
// among the *n_ptr records.
__global__ __launch_bounds__(256) void comp_count_runs_kernel(const uint32_t* keys, const uint32_t* n_ptr,
                                                              uint32_t* count) {
    __shared__ uint32_t part[4];
    const uint32_t n = *n_ptr;
    uint32_t c = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        c += (i == 0 || keys[i - 1] != keys[i]) ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(count, part[0] + part[1] + part[2] + part[3]);
}

__global__ void comp_zero_kernel(float4* g, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Sum each run of equal keys of the sorted records into grad[key]. A wave
// takes 64 consecutive records and does a segmented inclusive scan over them
// (fixed shuffle order, so the result is the same on every call); the last
// lane of each run piece writes. A run that leaves the wave is finished by
// the wave holding its first record, reading on serially; waves that only
// hold the continuation of an earlier run write nothing for it.
__global__ __launch_bounds__(256) void comp_reduce_runs_kernel(const uint32_t* keys, const uint32_t* ids,
                                                               const float4* vals, const uint32_t* n_ptr,
                                                               float4* grad, uint32_t nvox) {
    const uint32_t n = *n_ptr;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return;  // whole block past the records
    const int lane = (int)(threadIdx.x & 63);
    const bool in = i < n;
    const uint32_t key = in ? keys[i] : 0xFFFFFFFFu;
    float4 s = in ? vals[ids[i]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t up = __shfl_up(key, 1u);
    const uint32_t before = lane > 0 ? up : (i > 0 && in ? keys[i - 1] : ~key);
    const bool head = in && before != key;
    const uint64_t heads = __ballot(head);
    const uint64_t starts = heads | 1ull;
    const uint64_t below = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __builtin_clzll(starts & below);
    for (int off = 1; off < 64; off <<= 1) {
        const float ox = __shfl_up(s.x, (unsigned)off), oy = __shfl_up(s.y, (unsigned)off);
        const float oz = __shfl_up(s.z, (unsigned)off), ow = __shfl_up(s.w, (unsigned)off);
        if (lane - off >= start) {
            s.x += ox;
            s.y += oy;
            s.z += oz;
            s.w += ow;
        }
    }
    if (!in) return;
    const uint32_t down = __shfl_down(key, 1u);
    const bool last_in_wave = lane == 63 || i + 1 >= n;
    const uint32_t after = last_in_wave ? (i + 1 < n ? keys[i + 1] : ~key) : down;
    if (after == key && !last_in_wave) return;  // not the end of this run piece
    const bool owns = (heads >> start) & 1ull;  // the run's first record is in this wave
    if (!owns) return;
    for (uint32_t j = i + 1; j < n && keys[j] == key; ++j) {
        const float4 v = vals[ids[j]];
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
    }
    if (key < nvox) grad[key] = s;  // a record no sample wrote holds no valid key
}

// Record ranges per brick slot after the records are sorted by slot (key >>
// shift): start[s] is the first record of slot s, start[num_slots] = n. Thread
// i looks at the boundary between records i - 1 and i and fills the starts of
// every slot in between, so empty slots get the start of the next non-empty
// one. The slot is read from the sorted bit range and clamped below num_slots, so
// even records that no sample wrote (sample counts that do not match the
// forward's) cannot index past start[num_slots].
// Thread t takes records 4t .. 4t + 3 (one 16-byte load; `keys` is 16-byte
// aligned) and the boundary before each of them, up to record n.
__global__ __launch_bounds__(256) void comp_slot_starts_kernel(const uint32_t* keys, const uint32_t* n_ptr,
                                                               int shift, uint32_t slot_mask, uint32_t num_slots,
                                                               uint32_t* start) {
    const uint32_t n = *n_ptr;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i0 = t * 4u;
    if (i0 > n) return;
    auto slot_of = [&](uint32_t k) { return min((k >> shift) & slot_mask, num_slots - 1u); };
    uint32_t k4[4];
    if (i0 + 4u <= n) {
        const uint4 v = *reinterpret_cast<const uint4*>(keys + i0);
        k4[0] = v.x;
        k4[1] = v.y;
        k4[2] = v.z;
        k4[3] = v.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) k4[j] = i0 + (uint32_t)j < n ? keys[i0 + j] : 0u;
    }
    uint32_t lo = i0 == 0 ? 0u : slot_of(keys[i0 - 1]) + 1u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = i0 + (uint32_t)j;
        if (i > n) break;
        const uint32_t s = i < n ? slot_of(k4[j]) : num_slots;
        for (uint32_t q = lo; q <= s; ++q) start[q] = i;
        lo = s + 1u;
    }
}

__device__ __forceinline__ float4 shfl4(float4 v, int src) {
    return make_float4(__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src), __shfl(v.w, src));
}

// Sum the records of one brick slot per wave (records sorted by slot, stable,
// so a slot's records are in emission order). The slot's 2^VB voxel sums live
// in LDS. The wave takes 64 records at a time; lanes holding the same voxel
// find each other with VB ballots, the lowest of them adds the group's values
// in lane order and then adds that to the voxel's sum. Every sum is thus built
// in emission order, whatever the schedule: the result is the same on every
// call. The wave then writes the whole brick, untouched voxels as 0, and
// counts its distinct voxels.
template <int VB>
__global__ __launch_bounds__(256) void comp_reduce_slots_kernel(const uint32_t* keys, const uint32_t* ids,
                                                                const float4* vals, const uint32_t* start,
                                                                uint32_t num_slots, float4* grad,
                                                                uint32_t* distinct, const uint32_t* overflow) {
    constexpr int NV = 1 << VB;
    constexpr int NW = (NV + 31) / 32;  // touched-voxel bit words
    __shared__ float4 acc_all[4][NV];
    __shared__ uint32_t seen_all[4][NW];
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const uint32_t slot = blockIdx.x * 4u + (uint32_t)w;
    if (slot >= num_slots) return;  // whole wave; the waves of a block never synchronise
    float4* acc = acc_all[w];
    uint32_t* seen = seen_all[w];
    for (int v = lane; v < NV; v += 64) acc[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int v = lane; v < NW; v += 64) seen[v] = 0u;
    const uint32_t b = start[slot], e = start[slot + 1];
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t base = b; base < e; base += 64u) {
        const uint32_t i = base + (uint32_t)lane;
        const bool in = i < e;
        uint32_t v = 0;
        float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (in) {
            v = keys[i] & (uint32_t)(NV - 1);
            x = vals[ids[i]];
        }
        uint64_t peers = __ballot(in);
#pragma unroll
        for (int bit = 0; bit < VB; ++bit) {
            const bool on = (v >> bit) & 1u;
            const uint64_t m = __ballot(on);
            peers &= on ? m : ~m;
        }
        const bool leader = in && (peers & below) == 0;
        uint64_t rest = leader ? (peers & ~(below | (1ull << lane))) : 0ull;
        float4 s = x;
        while (__ballot(rest != 0ull) != 0ull) {
            const int src = rest != 0ull ? __ffsll((unsigned long long)rest) - 1 : lane;
            const float4 y = shfl4(x, src);
            if (rest != 0ull) {
                s.x += y.x;
                s.y += y.y;
                s.z += y.z;
                s.w += y.w;
                rest &= rest - 1ull;
            }
        }
        if (leader) {
            float4 t = acc[v];
            t.x += s.x;
            t.y += s.y;
            t.z += s.z;
            t.w += s.w;
            acc[v] = t;
            atomicOr(&seen[v >> 5], 1u << (v & 31u));
        }
    }
    float4* g = grad + (size_t)slot * NV;
    if (*overflow) {  // records were dropped: no gradient rather than a wrong one
        const float qn = __builtin_nanf("");
        for (int v = lane; v < NV; v += 64) g[v] = make_float4(qn, qn, qn, qn);
        return;
    }
    for (int v = lane; v < NV; v += 64) g[v] = acc[v];
    uint32_t cnt = 0;
    for (int k = lane; k < NW; k += 64) cnt += (uint32_t)__builtin_popcount(seen[k]);
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if (lane == 0 && cnt) atomicAdd(distinct, cnt);
}

// Record total against the workspace capacity: n = min(total, cap), the
// overflow flag, and the distinct-voxel counter cleared.
__global__ void comp_check_kernel(uint32_t* misc, uint32_t cap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t total = misc[kMiscTotal];
    misc[kMiscN] = total <= cap ? total : cap;
    misc[kMiscOverflow] = total > cap ? 1u : 0u;
    misc[kMiscDistinct] = 0u;
}

// After an overflow, the whole-key path's gradient is NaN (the per-slot path
// writes NaN bricks itself).
__global__ void comp_poison_kernel(const uint32_t* overflow, float4* g, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && *overflow) {
        const float qn = __builtin_nanf("");
        g[i] = make_float4(qn, qn, qn, qn);
    }
}

}  // namespace yx

using namespace yx;

#define YX_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(YX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

struct yx_compositor {
    void* ws = nullptr;  // one device allocation, carved into the buffers below
    size_t ws_bytes = 0;
    uint64_t cap = 0;           // records the workspace holds
    bool reserved = false;      // cap set by yx_compositor_reserve
    uint64_t sig[8] = {};       // geometry of the call cap was sized for
    uint32_t* misc = nullptr;   // kMiscWords device words (inside ws)
    uint32_t* h_misc = nullptr; // pinned copy of misc after each backward
    uint32_t* d_probe = nullptr;  // sizing: the record total of a new geometry
    const uint32_t* sorted_keys = nullptr;  // the last backward's sorted keys (workspace)
    bool counted = false;  // the last backward's slot reduction counted its distinct voxels
    bool ran = false;      // a backward has run since the workspace was laid out
    uint8_t* super = nullptr;  // super-brick summary (rebuilt by every call)
    size_t super_bytes = 0;
};

static int check_vol(const yx_volume_dev* v, int32_t n_views, int32_t w, int32_t h) {
    if (!v || !v->table || (!v->params && v->num_slots)) return fail(YX_ERR_INVALID_ARG, "null volume pointer");
    auto pow2 = [](int x) { return x > 0 && (x & (x - 1)) == 0; };
    if (!pow2(v->grid) || !pow2(v->brick) || v->brick > v->grid || v->grid > 4096)
        return fail(YX_ERR_INVALID_ARG, "grid %d and brick %d must be powers of two, brick <= grid <= 4096", v->grid,
                    v->brick);
    if ((uint64_t)v->num_slots * (uint64_t)v->brick * v->brick * v->brick > 0xFFFFFFFFull)
        return fail(YX_ERR_RANGE, "more than 2^32 voxel slots");
    if (n_views < 1 || n_views > 65535 || w < 1 || h < 1 || w > 16384 || h > 16384)
        return fail(YX_ERR_INVALID_ARG, "views %d of %dx%d", n_views, w, h);
    return YX_OK;
}

static CompArgs make_args(const yx_volume_dev* v, const float* views, int32_t n_views, int32_t w, int32_t h) {
    CompArgs a;
    std::memset(&a, 0, sizeof(a));
    a.table = v->table;
    a.params = (const float4*)v->params;
    a.num_slots = v->num_slots;
    a.views = views;
    a.n = v->grid;
    a.lb = __builtin_ctz((unsigned)v->brick);
    a.g = v->grid / v->brick;
    a.lg = __builtin_ctz((unsigned)a.g);
    a.ls = a.lg < 2 ? a.lg : 2;
    a.n_views = n_views;
    a.w = w;
    a.h = h;
    return a;
}

extern "C" int yx_camera_basis(const float pos[3], const float dir[3], const float up[3], float fov, int32_t w,
                               int32_t h, float view[12]) {
    clear_error();
    if (!pos || !dir || !up || !view) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (w <= 0 || h <= 0 || !(fov > 0.0f && fov < 180.0f)) return fail(YX_ERR_INVALID_ARG, "bad camera");
    for (int i = 0; i < 3; ++i) view[i] = pos[i];
    camera_basis(dir, up, fov, w, h, view + 3);
    return YX_OK;
}

extern "C" int yx_compositor_create(yx_compositor** out) {
    clear_error();
    if (!out) return fail(YX_ERR_INVALID_ARG, "null argument");
    yx_compositor* c = new (std::nothrow) yx_compositor();
    if (!c) return fail(YX_ERR_OOM, "out of host memory");
    *out = c;
    return YX_OK;
}

extern "C" void yx_compositor_destroy(yx_compositor* c) {
    if (!c) return;
    if (c->ws) (void)hipFree(c->ws);
    if (c->h_misc) (void)hipHostFree(c->h_misc);
    if (c->d_probe) (void)hipFree(c->d_probe);
    if (c->super) (void)hipFree(c->super);
    delete c;
}

extern "C" int yx_compositor_reserve(yx_compositor* c, uint64_t records) {
    clear_error();
    if (!c) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (records >= 0x7FFFFFFFull) return fail(YX_ERR_RANGE, "%llu records exceed 2^31", (unsigned long long)records);
    if (records > c->cap) c->cap = records;
    c->reserved = true;
    return YX_OK;
}

extern "C" int yx_compositor_last_stats(const yx_compositor* c, uint64_t* records, uint64_t* unique_keys) {
    clear_error();
    if (!c || !records || !unique_keys) return fail(YX_ERR_INVALID_ARG, "null argument");
    *records = 0;
    *unique_keys = 0;
    if (!c->ran || !c->misc) return YX_OK;
    // the backward ran on the caller's stream, which may not synchronise with
    // the null stream: wait for the device before reading its results
    YX_HIP(hipDeviceSynchronize());
    uint32_t m[kMiscWords];
    YX_HIP(hipMemcpy(m, c->misc, sizeof(m), hipMemcpyDeviceToHost));
    if (m[kMiscOverflow])
        return fail(YX_ERR_RANGE, "the last backward had %u records, beyond the workspace's %llu: its gradient is NaN",
                    m[kMiscTotal], (unsigned long long)c->cap);
    *records = m[kMiscTotal];
    if (c->counted) {
        *unique_keys = m[kMiscDistinct];
    } else if (c->sorted_keys && m[kMiscN]) {
        uint32_t* cnt = c->misc + kMiscDistinct;
        YX_HIP(hipMemset(cnt, 0, 4));
        hipLaunchKernelGGL(comp_count_runs_kernel, dim3(1024), dim3(256), 0, nullptr, c->sorted_keys,
                           c->misc + kMiscN, cnt);
        YX_HIP(hipGetLastError());
        uint32_t runs = 0;
        YX_HIP(hipMemcpy(&runs, cnt, 4, hipMemcpyDeviceToHost));
        *unique_keys = runs;
    }
    return YX_OK;
}

// The super-brick summary of the volume's current table, rebuilt on `st`
// before every walk (the table may have changed since the last call).
static int build_super(yx_compositor* c, CompArgs& a, hipStream_t st) {
    const size_t nsb = (size_t)1 << (3 * (a.lg - a.ls));
    if (c->super_bytes < nsb) {
        if (c->super) YX_HIP(hipFree(c->super));
        c->super = nullptr;
        c->super_bytes = 0;
        const hipError_t e = hipMalloc((void**)&c->super, nsb);
        if (e != hipSuccess) {
            c->super = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%zu) for the super-brick summary: %s", nsb, hipGetErrorString(e));
        }
        c->super_bytes = nsb;
    }
    const size_t threads = a.ls == 2 ? nsb * 64 : nsb;
    hipLaunchKernelGGL(comp_super_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, a.table,
                       a.num_slots, a.lg, a.ls, c->super);
    YX_HIP(hipGetLastError());
    a.super = c->super;
    return YX_OK;
}

extern "C" int yx_composite_forward(yx_compositor* c, const yx_volume_dev* vol, const float* views, int32_t n_views,
                                    int32_t w, int32_t h, float* rgb, float* trans, uint32_t* nsamp, void* stream) {
    clear_error();
    if (!c || !views) return fail(YX_ERR_INVALID_ARG, "null argument");
    int rc = check_vol(vol, n_views, w, h);
    if (rc != YX_OK) return rc;
    CompArgs a = make_args(vol, views, n_views, w, h);
    a.rgb = rgb;
    a.trans = trans;
    a.nsamp = nsamp;
    rc = build_super(c, a, (hipStream_t)stream);
    if (rc != YX_OK) return rc;
    const size_t waves = (size_t)n_views * ((w + 7) / 8) * ((h + 7) / 8);
    const dim3 grid((unsigned)((waves * 64 + 255) / 256)), block(256);
    if (a.g <= 1024)
        hipLaunchKernelGGL(comp_forward_kernel<true>, grid, block, 0, (hipStream_t)stream, a);
    else
        hipLaunchKernelGGL(comp_forward_kernel<false>, grid, block, 0, (hipStream_t)stream, a);
    YX_HIP(hipGetLastError());
    return YX_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The backward's workspace for `cap` records, n_waves ray tiles and
// num_slots brick slots (byte offsets, 256-aligned).
struct CompLayout {
    size_t misc, sums, offs, scan_ws, keys, vals, kA, iA, kB, iB, sort_ws, starts, bytes;
    size_t scan_bytes, sort_bytes;
};

static CompLayout comp_layout(uint64_t cap, size_t n_waves, uint32_t num_slots) {
    CompLayout L;
    L.scan_bytes = scan_workspace_bytes(n_waves);
    L.sort_bytes = sort_workspace_bytes(cap);
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += align256(bytes);
        return at;
    };
    L.misc = take(kMiscWords * 4);
    L.sums = take(n_waves * 4);
    L.offs = take(n_waves * 4);
    L.scan_ws = take(L.scan_bytes);
    L.keys = take(cap * 4);
    L.vals = take(cap * 16);
    L.kA = take(cap * 4);
    L.iA = take(cap * 4);
    L.kB = take(cap * 4);
    L.iB = take(cap * 4);
    L.sort_ws = take(L.sort_bytes);
    L.starts = take(((size_t)num_slots + 1) * 4);
    L.bytes = o;
    return L;
}

extern "C" int yx_composite_backward(yx_compositor* c, const yx_volume_dev* vol, const float* views,
                                     int32_t n_views, int32_t w, int32_t h, const float* rgb, const uint32_t* nsamp,
                                     const float* grad_rgb, float* grad_params, void* stream) {
    clear_error();
    if (!c || !views || !rgb || !nsamp || !grad_rgb || !grad_params)
        return fail(YX_ERR_INVALID_ARG, "null argument");
    int rc = check_vol(vol, n_views, w, h);
    if (rc != YX_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t rays = (size_t)n_views * w * h;
    const size_t nvox = (size_t)vol->num_slots * vol->brick * vol->brick * vol->brick;
    if (rays >= 0xFFFFFFFFull) return fail(YX_ERR_RANGE, "too many rays");
    // the last call's statistics point into the workspace, which may be
    // replaced below: forget them first
    c->sorted_keys = nullptr;
    c->counted = false;

    const size_t n_waves = (size_t)n_views * ((w + 7) / 8) * ((h + 7) / 8);
    CompArgs a = make_args(vol, views, n_views, w, h);
    a.nsamp_in = nsamp;
    a.n_waves = (uint32_t)n_waves;
    const dim3 wgrid((unsigned)((n_waves * 64 + 255) / 256)), wblock(256);

    // 0. capacity. A previous call whose records overflowed (read from its
    //    pinned copy of the device words; a copy still in flight reads as no
    //    overflow) grows it to that call's total. A call on a new geometry
    //    (table, views, sizes) is sized with one host wait, unless the caller
    //    reserved a capacity; the same geometry always has the same samples.
    if (c->h_misc && c->h_misc[kMiscOverflow]) {
        const uint64_t t = c->h_misc[kMiscTotal];
        c->cap = std::max<uint64_t>(c->cap, t + t / 4);
        c->h_misc[kMiscOverflow] = 0;
    }
    const uint64_t sig[8] = {(uint64_t)(uintptr_t)vol->table, (uint64_t)(uintptr_t)views, (uint64_t)n_views,
                             (uint64_t)w, (uint64_t)h, (uint64_t)vol->grid, (uint64_t)vol->brick,
                             (uint64_t)vol->num_slots};
    if (!c->reserved && (c->cap == 0 || std::memcmp(sig, c->sig, sizeof(sig)) != 0)) {
        const size_t sums_bytes = align256(n_waves * 4), red_bytes = scan_workspace_bytes(n_waves);
        const size_t bytes = sums_bytes + red_bytes + 256;
        if (c->d_probe) YX_HIP(hipFree(c->d_probe));
        c->d_probe = nullptr;
        hipError_t e = hipMalloc((void**)&c->d_probe, bytes);
        if (e != hipSuccess) {
            c->d_probe = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%zu) for sizing: %s", bytes, hipGetErrorString(e));
        }
        uint32_t* sums = c->d_probe;
        char* red_ws = (char*)c->d_probe + sums_bytes;
        uint32_t* tot = (uint32_t*)(red_ws + red_bytes);
        hipLaunchKernelGGL(comp_wave_sums_kernel, wgrid, wblock, 0, st, a, sums, (uint32_t)n_waves);
        rc = reduce_sum_u32(sums, n_waves, tot, red_ws, red_bytes, st);
        if (rc != YX_OK) return rc;
        uint32_t total = 0;
        YX_HIP(hipMemcpyAsync(&total, tot, 4, hipMemcpyDeviceToHost, st));
        YX_HIP(hipStreamSynchronize(st));
        c->cap = std::max<uint64_t>(c->cap, (uint64_t)total + total / 4 + 64);
        std::memcpy(c->sig, sig, sizeof(sig));
    }
    if (c->cap >= 0x7FFFFFFFull) return fail(YX_ERR_RANGE, "%llu records exceed 2^31", (unsigned long long)c->cap);
    const uint32_t cap = (uint32_t)std::max<uint64_t>(c->cap, 1);
    const CompLayout L = comp_layout(cap, n_waves, vol->num_slots);
    if (c->ws_bytes < L.bytes) {
        if (c->ws) YX_HIP(hipFree(c->ws));
        c->ws = nullptr;
        c->ws_bytes = 0;
        c->misc = nullptr;
        c->ran = false;
        hipError_t e = hipMalloc(&c->ws, L.bytes);
        if (e != hipSuccess) {
            c->ws = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%zu) workspace: %s", L.bytes, hipGetErrorString(e));
        }
        c->ws_bytes = L.bytes;
    }
    if (!c->h_misc) {
        hipError_t e = hipHostMalloc((void**)&c->h_misc, kMiscWords * 4, hipHostMallocDefault);
        if (e != hipSuccess) {
            c->h_misc = nullptr;
            return fail(YX_ERR_OOM, "hipHostMalloc(%d): %s", kMiscWords * 4, hipGetErrorString(e));
        }
        std::memset(c->h_misc, 0, kMiscWords * 4);
    }
    char* base = (char*)c->ws;
    uint32_t* misc = (uint32_t*)(base + L.misc);
    c->misc = misc;
    uint32_t* sums = (uint32_t*)(base + L.sums);
    uint32_t* offs = (uint32_t*)(base + L.offs);
    void* scan_ws = base + L.scan_ws;
    uint32_t* keys = (uint32_t*)(base + L.keys);
    float4* vals = (float4*)(base + L.vals);
    uint32_t* kbuf[2] = {(uint32_t*)(base + L.kA), (uint32_t*)(base + L.kB)};
    uint32_t* ibuf[2] = {(uint32_t*)(base + L.iA), (uint32_t*)(base + L.iB)};
    void* sort_ws = base + L.sort_ws;
    uint32_t* starts = (uint32_t*)(base + L.starts);
    uint32_t* n_dev = misc + kMiscN;

    // 1. per-wave sample counts (8x8 ray tiles), exclusive scan -> each wave's
    //    record offset; the total against the capacity
    hipLaunchKernelGGL(comp_wave_sums_kernel, wgrid, wblock, 0, st, a, sums, (uint32_t)n_waves);
    YX_HIP(hipGetLastError());
    rc = scan_excl_u32(sums, offs, n_waves, misc + kMiscTotal, scan_ws, L.scan_bytes, st);
    if (rc != YX_OK) return rc;
    hipLaunchKernelGGL(comp_check_kernel, dim3(1), dim3(64), 0, st, misc, cap);
    YX_HIP(hipGetLastError());

    // Bricks of up to 8^3 voxels are reduced per slot (records sorted by the
    // slot bits of the key only); larger bricks sort the whole key and sum
    // each key's run.
    const int vox_bits = 3 * __builtin_ctz((unsigned)vol->brick);
    const bool by_slot = vox_bits <= 9;
    if (!by_slot && nvox) {  // voxels without samples keep 0 (the slot reduction writes every voxel)
        hipLaunchKernelGGL(comp_zero_kernel, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, st,
                           (float4*)grad_params, nvox);
        YX_HIP(hipGetLastError());
    }
    int key_bits = 1;
    while (key_bits < 32 && ((uint64_t)1 << key_bits) < (uint64_t)nvox) ++key_bits;
    const int begin_bit = by_slot ? vox_bits : 0;
    const int end_bit = std::max(key_bits, begin_bit + 1);

    // 2. emit records (one wave = one 8x8 ray tile) at their offsets
    rc = build_super(c, a, st);
    if (rc != YX_OK) return rc;
    a.rgb_in = rgb;
    a.grad_rgb = grad_rgb;
    a.offsets = offs;
    a.total_records = cap;
    a.keys = keys;
    a.vals = vals;
    if (a.g <= 1024)
        hipLaunchKernelGGL(comp_backward_kernel<true>, wgrid, wblock, 0, st, a);
    else
        hipLaunchKernelGGL(comp_backward_kernel<false>, wgrid, wblock, 0, st, a);
    YX_HIP(hipGetLastError());

    // 3. record ids sorted by the slot bits (or the whole key), stable: LSD
    //    passes of 8-bit digits, each a digit histogram per block, a scan of
    //    the histograms and a scatter; the first pass's ids are the positions.
    // 4. sum each voxel's records in emission order
    const uint32_t* keys_sorted = nullptr;
    const uint32_t* ids_sorted = nullptr;
    rc = radix_sort<uint32_t>(keys, nullptr, kbuf, ibuf, cap, n_dev, begin_bit, end_bit, &keys_sorted, &ids_sorted,
                              sort_ws, L.sort_bytes, st);
    if (rc != YX_OK) return rc;
    if (by_slot) {
        const uint32_t ns = vol->num_slots;
        const uint32_t slot_mask = (uint32_t)(((uint64_t)1 << (end_bit - begin_bit)) - 1u);
        const size_t threads = (size_t)cap / 4 + 1;  // thread t covers records 4t .. 4t + 3, up to record n
        hipLaunchKernelGGL(comp_slot_starts_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st,
                           keys_sorted, n_dev, vox_bits, slot_mask, ns, starts);
        YX_HIP(hipGetLastError());
        if (ns) {
            const dim3 grid((unsigned)((ns + 3) / 4)), block(256);
            float4* g = (float4*)grad_params;
            uint32_t* distinct = misc + kMiscDistinct;
            const uint32_t* over = misc + kMiscOverflow;
            switch (vox_bits) {
                case 0:
                    hipLaunchKernelGGL(comp_reduce_slots_kernel<0>, grid, block, 0, st, keys_sorted, ids_sorted, vals,
                                       starts, ns, g, distinct, over);
                    break;
                case 3:
                    hipLaunchKernelGGL(comp_reduce_slots_kernel<3>, grid, block, 0, st, keys_sorted, ids_sorted, vals,
                                       starts, ns, g, distinct, over);
                    break;
                case 6:
                    hipLaunchKernelGGL(comp_reduce_slots_kernel<6>, grid, block, 0, st, keys_sorted, ids_sorted, vals,
                                       starts, ns, g, distinct, over);
                    break;
                default:
                    hipLaunchKernelGGL(comp_reduce_slots_kernel<9>, grid, block, 0, st, keys_sorted, ids_sorted, vals,
                                       starts, ns, g, distinct, over);
                    break;
            }
            YX_HIP(hipGetLastError());
        }
        c->counted = true;
    } else {
        hipLaunchKernelGGL(comp_reduce_runs_kernel, dim3((unsigned)(((size_t)cap + 255) / 256)), dim3(256), 0, st,
                           keys_sorted, ids_sorted, vals, n_dev, (float4*)grad_params, (uint32_t)nvox);
        YX_HIP(hipGetLastError());
        if (nvox)
            hipLaunchKernelGGL(comp_poison_kernel, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, st,
                               misc + kMiscOverflow, (float4*)grad_params, nvox);
        YX_HIP(hipGetLastError());
        c->sorted_keys = keys_sorted;
    }
    // 5. the device words to pinned host memory: the next call sees an
    //    overflow without waiting for this one
    YX_HIP(hipMemcpyAsync(c->h_misc, misc, kMiscWords * 4, hipMemcpyDeviceToHost, st));
    c->ran = true;
    return YX_OK;
}
