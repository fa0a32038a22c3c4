// Brick n-tree volumes on gfx950: GPU builder and DDA ray cast.
//
// Structure (ntree/ntree/nodes.h:13-100 Node Grid/Brick/Const, ch2ofs
// (z*G + y)*G + x at :99; ntree/ntree/gpu_nodes.h:6-21; 8^3 bricks as in
// BrickManager's brick pool, ntree/BrickManager.h:5-81): size N = 8 * 4^L * R
// voxels per axis; an R^3 root grid, L levels of 4^3 nodes, 8^3 u8 bricks.
// Device layout (HBM):
//   root   R^3 u32: index of the level-1 node (of the brick when L = 0) or
//          kNtEmpty; staged in LDS by the tracer.
//   nodes  16 B each {mask lo, mask hi, base, 0}: bit c = cell c non-empty,
//          child = base + popcount(mask below c); levels 1..L one after the
//          other, each in tree order, so a node's children are contiguous.
//   bricks 512 B each (voxel x + 8y + 64z), in tree order.
// Build (MarkBricks / PackBricks of reaction/voxtree/density.cu:131-221 as
// classify -> compact -> pack): one wave per brick evaluates its 512
// densities and votes; each node level is one wave per node whose 64 lanes
// are its 64 cells (the node's child mask is the wave's ballot); exclusive
// scans of the flags give every node and brick its index; the allocated
// bricks are evaluated again and written to their slots.
// Ray cast (castRay / walkBrick, reaction/trace/trace.cu:42-139): empty cells
// at every level are left in one step, brick voxels are crossed one by one
// (3D DDA) and composited front to back through a 256-entry RGBA8 transfer
// table (volvis.py:145-152), stopping once alpha > 0.99.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "internal.h"
#include "prims.h"
#include "trace_device.h"
#include "trace_persist.h"

namespace yx {

constexpr uint32_t kNtEmpty = 0xFFFFFFFFu;
constexpr int kNtMaxLevels = 4;  // 4096^3 with an 8^3 root grid uses 3
constexpr int kNtMaxRoot = 16;   // root grid side staged in LDS (16^3 x 4 B)

// ---- density: fBm of improved Perlin noise, float32 ------------------------

__device__ __forceinline__ float nt_fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ float nt_lerp(float t, float a, float b) { return a + t * (b - a); }

// The improved noise's gradient set (12 cube edges, 4 repeated) written as
// u +- v with u = h < 8 ? x : y and v = h < 4 ? y : (h == 12 || 14 ? x : z).
__device__ __forceinline__ float nt_grad(int h, float x, float y, float z) {
    h &= 15;
    const float u = h < 8 ? x : y;
    const float v = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
}

__device__ __forceinline__ float nt_noise(const int* P, float x, float y, float z) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    const float xf = x - fx, yf = y - fy, zf = z - fz;
    const float u = nt_fade(xf), v = nt_fade(yf), w = nt_fade(zf);
    const int a = P[X] + Y, aa = P[a] + Z, ab = P[a + 1] + Z;
    const int b = P[X + 1] + Y, ba = P[b] + Z, bb = P[b + 1] + Z;
    const float x1 = xf - 1.0f, y1 = yf - 1.0f, z1 = zf - 1.0f;
    const float l1 = nt_lerp(u, nt_grad(P[aa], xf, yf, zf), nt_grad(P[ba], x1, yf, zf));
    const float l2 = nt_lerp(u, nt_grad(P[ab], xf, y1, zf), nt_grad(P[bb], x1, y1, zf));
    const float l3 = nt_lerp(u, nt_grad(P[aa + 1], xf, yf, z1), nt_grad(P[ba + 1], x1, yf, z1));
    const float l4 = nt_lerp(u, nt_grad(P[ab + 1], xf, y1, z1), nt_grad(P[bb + 1], x1, y1, z1));
    return nt_lerp(w, nt_lerp(v, l1, l2), nt_lerp(v, l3, l4));
}

struct NtNoise {
    float inv, frequency, threshold, scale;
    int octaves;
};

__device__ __forceinline__ uint32_t nt_density(const NtNoise& q, const int* P, int x, int y, int z) {
    const float px = ((float)x + 0.5f) * q.inv * q.frequency;
    const float py = ((float)y + 0.5f) * q.inv * q.frequency;
    const float pz = ((float)z + 0.5f) * q.inv * q.frequency;
    float f = 0.0f, amp = 1.0f, fr = 1.0f;
    for (int o = 0; o < q.octaves; ++o) {
        f = f + amp * nt_noise(P, px * fr, py * fr, pz * fr);
        amp = amp * 0.5f;
        fr = fr * 2.0f;
    }
    if (f < q.threshold) return 0u;
    const float s = (f - q.threshold) * q.scale;
    return (uint32_t)(1 + (int)fminf(s, 254.0f));
}

// Brick coordinates of the brick with tree-order index i.
__device__ __forceinline__ void nt_brick_coords(uint64_t i, int L, int R, int& bx, int& by, int& bz) {
    bx = by = bz = 0;
    for (int l = L; l >= 1; --l) {
        const int dg = (int)(i & 63u);
        i >>= 6;
        const int s = 2 * (L - l);
        bx |= (dg & 3) << s;
        by |= ((dg >> 2) & 3) << s;
        bz |= (dg >> 4) << s;
    }
    const int r = (int)i;
    const int sh = 2 * L;
    bx |= (r % R) << sh;
    by |= ((r / R) % R) << sh;
    bz |= (r / (R * R)) << sh;
}

// Source of brick densities: the fBm (perm table in LDS) or a dense volume.
struct NtSource {
    NtNoise q;
    const int* perm;       // 512 entries (global); copied to LDS
    const uint8_t* dense;  // N^3, x fastest, or null for the fBm
    int size;
};

// lane k of the wave holds voxels (k & 7, k >> 3, z) for z = 0..7
__device__ __forceinline__ uint32_t nt_voxel(const NtSource& s, const int* P, int x, int y, int z) {
    if (s.dense) return s.dense[((size_t)z * s.size + y) * s.size + x];
    return nt_density(s.q, P, x, y, z);
}

__global__ __launch_bounds__(256) void nt_classify_kernel(NtSource s, int L, int R, uint64_t n_bricks,
                                                          uint8_t* flags) {
    __shared__ int P[512];
    for (int i = threadIdx.x; i < 512; i += 256) P[i] = s.perm ? s.perm[i] : 0;
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t wave = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    const uint64_t stride = ((uint64_t)gridDim.x * 256) >> 6;
    for (uint64_t i = wave; i < n_bricks; i += stride) {
        int bx, by, bz;
        nt_brick_coords(i, L, R, bx, by, bz);
        const int x = bx * 8 + (lane & 7), y = by * 8 + (lane >> 3);
        uint32_t any = 0;
        for (int z = 0; z < 8; ++z) any |= nt_voxel(s, P, x, y, bz * 8 + z);
        const uint64_t vote = __ballot(any != 0);
        if (lane == 0) flags[i] = vote != 0 ? 1 : 0;
    }
}

// One wave per node: lane c reads cell c's child flag; the ballot is the mask.
__global__ __launch_bounds__(256) void nt_level_kernel(const uint8_t* child_flags, uint64_t n_units,
                                                       uint8_t* flags, uint64_t* masks) {
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t wave = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    const uint64_t stride = ((uint64_t)gridDim.x * 256) >> 6;
    for (uint64_t u = wave; u < n_units; u += stride) {
        const uint64_t m = __ballot(child_flags[u * 64 + (uint64_t)lane] != 0);
        if (lane == 0) {
            masks[u] = m;
            flags[u] = m != 0 ? 1 : 0;
        }
    }
}

__global__ __launch_bounds__(256) void nt_write_nodes_kernel(const uint8_t* flags, const uint32_t* scan,
                                                             const uint64_t* masks, uint64_t n_units,
                                                             uint32_t offset, const uint32_t* child_scan,
                                                             uint32_t child_offset, uint4* nodes) {
    for (uint64_t u = (uint64_t)blockIdx.x * 256 + threadIdx.x; u < n_units; u += (uint64_t)gridDim.x * 256) {
        if (!flags[u]) continue;
        const uint64_t m = masks[u];
        const uint32_t base = child_offset + child_scan[u * 64];
        nodes[offset + scan[u]] = make_uint4((uint32_t)m, (uint32_t)(m >> 32), base, 0u);
    }
}

__global__ __launch_bounds__(256) void nt_write_root_kernel(const uint8_t* flags, const uint32_t* scan, uint32_t n,
                                                            uint32_t offset, uint32_t* root) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) root[r] = flags[r] ? offset + scan[r] : kNtEmpty;
}

__global__ __launch_bounds__(256) void nt_pack_kernel(NtSource s, int L, int R, uint64_t n_bricks,
                                                      const uint8_t* flags, const uint32_t* scan,
                                                      uint8_t* bricks) {
    __shared__ int P[512];
    for (int i = threadIdx.x; i < 512; i += 256) P[i] = s.perm ? s.perm[i] : 0;
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t wave = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    const uint64_t stride = ((uint64_t)gridDim.x * 256) >> 6;
    for (uint64_t i = wave; i < n_bricks; i += stride) {
        if (!flags[i]) continue;  // wave-uniform
        int bx, by, bz;
        nt_brick_coords(i, L, R, bx, by, bz);
        const int x = bx * 8 + (lane & 7), y = by * 8 + (lane >> 3);
        uint8_t* b = bricks + (size_t)scan[i] * 512u;
        for (int z = 0; z < 8; ++z) b[lane + 64 * z] = (uint8_t)nt_voxel(s, P, x, y, bz * 8 + z);
    }
}

// ---- ray cast -----------------------------------------------------------------

struct NtArgs {
    const uint32_t* root;
    const uint4* nodes;
    const uint8_t* bricks;
    const uint32_t* palette;  // 256 RGBA8
    float view[12];           // eye position, dir0, du, dv
    int size, R, w, h;
    uint32_t* rgba;
    float* t;
    int32_t* voxel;
    uint32_t* n_nodes;
    uint32_t* n_voxels;
    // persistent launch
    uint32_t* counters;  // [0]: next ray id, [1]: retired waves
    uint32_t total;      // rays in whole 8x8 tiles
    uint32_t n_waves;
    int refill_min;
};

__device__ __forceinline__ float nt_clamp_eps(float x) {
    const float eps = 1e-5f;
    return fabsf(x) < eps ? copysignf(eps, x) : x;
}

__device__ __forceinline__ float nt_saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ float pick3(int a, float x, float y, float z) { return a == 0 ? x : (a == 1 ? y : z); }
__device__ __forceinline__ int pick3(int a, int x, int y, int z) { return a == 0 ? x : (a == 1 ? y : z); }

// Other-axis coordinate after leaving a cell at time t: floor((o + t d) N),
// never behind the current voxel, kept inside the cell [lo, lo + cs).
__device__ __forceinline__ int nt_follow(float o, float d, float t, float nf, int cur, int sg, int lo, int cs) {
    int q = (int)floorf((o + t * d) * nf);
    q = sg > 0 ? max(q, cur) : min(q, cur);
    return min(max(q, lo), lo + cs - 1);
}

// State of one ray of the n-tree cast.
template <int L>
struct NtRay {
    // the origin is the camera position of the launch (a.view[0..2], the
    // same for every ray), read from the arguments where needed
    float dx, dy, dz, ix, iy, iz, tx, t;
    int vx, vy, vz, guard;
    float ar, ag, ab, aa, thit;
    int hx, hy, hz;
    uint32_t nn, nv;
    bool first, inb;
    uint32_t brick;  // index of the brick the ray is in (its 512 voxels at bricks + 512 * brick)
    // per level the node last read: its cell (kx | ky << 10 | kz << 20; the
    // coordinates stay below 2^9 for sizes up to 16384, so all ones means
    // none), child masks and child base. The deepest level keeps only its
    // cell (to count reads of a new node): its node is read again at every
    // lookup, which leaves the cast at 64 VGPRs and 8 waves per SIMD
    uint32_t nck[L > 0 ? L : 1];
    uint32_t ncm_lo[L > 0 ? L : 1], ncm_hi[L > 0 ? L : 1], ncb[L > 0 ? L : 1];
};

// Ray of pixel (x, y); returns false when it misses the unit cube (the
// outputs then stay those of an empty ray).
template <int L>
__device__ __forceinline__ bool nt_setup(const NtArgs& a, int x, int y, NtRay<L>& r) {
    const float* b = a.view + 3;
    float dx = b[0] + b[3] * (float)x + b[6] * (float)y;
    float dy = b[1] + b[4] * (float)x + b[7] * (float)y;
    float dz = b[2] + b[5] * (float)x + b[8] * (float)y;
    float n2 = 0.0f;
    n2 += dx * dx;
    n2 += dy * dy;
    n2 += dz * dz;
    const float nrm = sqrtf(n2);
    r.dx = nt_clamp_eps(dx / nrm);
    r.dy = nt_clamp_eps(dy / nrm);
    r.dz = nt_clamp_eps(dz / nrm);
    const float ox = a.view[0], oy = a.view[1], oz = a.view[2];
    r.ix = 1.0f / r.dx;
    r.iy = 1.0f / r.dy;
    r.iz = 1.0f / r.dz;
    r.nn = r.nv = 0;
    r.thit = -1.0f;
    r.hx = r.hy = r.hz = -1;
    r.ar = r.ag = r.ab = r.aa = 0.0f;
    r.first = true;
    r.inb = false;
    r.brick = 0;
    r.guard = 0;
#pragma unroll
    for (int l = 0; l < (L > 0 ? L : 1); ++l) r.nck[l] = 0xFFFFFFFFu;
    const float lox = (0.0f - ox) * r.ix, hix = (1.0f - ox) * r.ix;
    const float loy = (0.0f - oy) * r.iy, hiy = (1.0f - oy) * r.iy;
    const float loz = (0.0f - oz) * r.iz, hiz = (1.0f - oz) * r.iz;
    float te = fmaxf(fmaxf(fmaxf(-INFINITY, fminf(lox, hix)), fminf(loy, hiy)), fminf(loz, hiz));
    r.tx = fminf(fminf(fminf(INFINITY, fmaxf(lox, hix)), fmaxf(loy, hiy)), fmaxf(loz, hiz));
    te = fmaxf(te, 0.0f);
    if (r.tx < 0.0f || te > r.tx) return false;
    const int N = a.size;
    const float nf = (float)N;
    r.vx = min(max((int)floorf((ox + te * r.dx) * nf), 0), N - 1);
    r.vy = min(max((int)floorf((oy + te * r.dy) * nf), 0), N - 1);
    r.vz = min(max((int)floorf((oz + te * r.dz) * nf), 0), N - 1);
    r.t = te;
    return true;
}

// One iteration of the walk: a ray outside a brick looks its voxel up from
// the root (node records re-read only when the voxel has left that level's
// node) and either leaves an empty cell in one step or enters the brick; a
// ray inside a brick crosses up to four of its voxels (a voxel step moves
// only the exit axis; the four positions do not depend on the data, so their
// loads are issued together and composited in order). Returns true when the
// ray is finished.
template <int L>
__device__ __forceinline__ bool nt_iterate(const NtArgs& a, const uint32_t* s_root, const uint32_t* s_pal,
                                           NtRay<L>& r) {
    const int N = a.size, R = a.R;
    const float nf = (float)N, invN = 1.0f / nf;
    const float ox = a.view[0], oy = a.view[1], oz = a.view[2];
    const int sx = r.dx > 0.0f ? 1 : -1, sy = r.dy > 0.0f ? 1 : -1, sz = r.dz > 0.0f ? 1 : -1;
    const int ux = sx > 0 ? 1 : 0, uy = sy > 0 ? 1 : 0, uz = sz > 0 ? 1 : 0;
    constexpr int sr = 3 + 2 * L;  // log2 of the root cell side
    if (++r.guard > 4 * N) return true;
    if (!r.inb) {
        int lS = sr;  // log2 of the side of the cell reached
        uint32_t idx = s_root[((r.vz >> sr) * R + (r.vy >> sr)) * R + (r.vx >> sr)];
#pragma unroll
        for (int l = 1; l <= L; ++l) {
            if (idx != kNtEmpty) {
                const int sn = 3 + 2 * (L - l + 1);
                const uint32_t key = (uint32_t)(r.vx >> sn) | ((uint32_t)(r.vy >> sn) << 10) |
                                     ((uint32_t)(r.vz >> sn) << 20);
                uint32_t mlo, mhi, nbase;
                if (l == L) {  // deepest level: read again each time, no registers held
                    const uint4 nd = a.nodes[idx];
                    if (r.nck[l - 1] != key) {
                        ++r.nn;
                        r.nck[l - 1] = key;
                    }
                    mlo = nd.x;
                    mhi = nd.y;
                    nbase = nd.z;
                } else {
                    if (r.nck[l - 1] != key) {
                        const uint4 nd = a.nodes[idx];
                        ++r.nn;
                        r.ncm_lo[l - 1] = nd.x;
                        r.ncm_hi[l - 1] = nd.y;
                        r.ncb[l - 1] = nd.z;
                        r.nck[l - 1] = key;
                    }
                    mlo = r.ncm_lo[l - 1];
                    mhi = r.ncm_hi[l - 1];
                    nbase = r.ncb[l - 1];
                }
                lS -= 2;
                const int s = sn - 2;
                const int c = ((((r.vz >> s) & 3) * 4) + ((r.vy >> s) & 3)) * 4 + ((r.vx >> s) & 3);
                const bool set = c < 32 ? ((mlo >> c) & 1u) : ((mhi >> (c - 32)) & 1u);
                if (!set) {
                    idx = kNtEmpty;
                } else {
                    const uint32_t below = c < 32 ? __builtin_popcount(mlo & ((1u << c) - 1u))
                                                  : __builtin_popcount(mlo) +
                                                        __builtin_popcount(mhi & ((1u << (c - 32)) - 1u));
                    idx = nbase + below;
                }
            }
        }
        if (idx == kNtEmpty) {
            // leave the empty cell of side 2^lS in one step
            const int cx = r.vx >> lS, cy = r.vy >> lS, cz = r.vz >> lS;
            const float tX = ((float)((cx + ux) << lS) * invN - ox) * r.ix;
            const float tY = ((float)((cy + uy) << lS) * invN - oy) * r.iy;
            const float tZ = ((float)((cz + uz) << lS) * invN - oz) * r.iz;
            const int ax = (tX <= tY && tX <= tZ) ? 0 : (tY <= tZ ? 1 : 2);
            const float tn = pick3(ax, tX, tY, tZ);
            if (tn > r.tx) return true;
            const int cs = 1 << lS;
            const int nx = ax == 0 ? (sx > 0 ? (cx + 1) << lS : (cx << lS) - 1)
                                   : nt_follow(ox, r.dx, tn, nf, r.vx, sx, cx << lS, cs);
            const int ny = ax == 1 ? (sy > 0 ? (cy + 1) << lS : (cy << lS) - 1)
                                   : nt_follow(oy, r.dy, tn, nf, r.vy, sy, cy << lS, cs);
            const int nz = ax == 2 ? (sz > 0 ? (cz + 1) << lS : (cz << lS) - 1)
                                   : nt_follow(oz, r.dz, tn, nf, r.vz, sz, cz << lS, cs);
            const int na = pick3(ax, nx, ny, nz);
            if (na < 0 || na >= N) return true;
            r.t = tn;
            r.vx = nx;
            r.vy = ny;
            r.vz = nz;
            return false;
        }
        r.inb = true;
        r.brick = idx;
    }
    // every valid position of the batch is in the current brick (the batch
    // stops at the first step that leaves it): keep its offset in the brick,
    // x + 8 y + 64 z, and the brick's corner once
    const int basex = r.vx & ~7, basey = r.vy & ~7, basez = r.vz & ~7;
    int off[4];
    float pt[4];
    bool valid[4];
    bool more = true, end = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        valid[j] = more;
        off[j] = (r.vx & 7) + 8 * (r.vy & 7) + 64 * (r.vz & 7);
        pt[j] = r.t;
        if (more) {
            const float tX = ((float)(r.vx + ux) * invN - ox) * r.ix;
            const float tY = ((float)(r.vy + uy) * invN - oy) * r.iy;
            const float tZ = ((float)(r.vz + uz) * invN - oz) * r.iz;
            const int ax = (tX <= tY && tX <= tZ) ? 0 : (tY <= tZ ? 1 : 2);
            const float tn = pick3(ax, tX, tY, tZ);
            const int oa = pick3(ax, r.vx, r.vy, r.vz);
            const int na = oa + pick3(ax, sx, sy, sz);
            if (tn > r.tx || na < 0 || na >= N) {
                end = true;
                more = false;
            } else {
                r.t = tn;
                r.vx = ax == 0 ? na : r.vx;
                r.vy = ax == 1 ? na : r.vy;
                r.vz = ax == 2 ? na : r.vz;
                // only the exit axis moved (by one): the ray left the brick iff
                // that coordinate crossed a multiple of 8
                if ((na >> 3) != (oa >> 3)) more = false;
            }
        }
    }
    uint32_t dd[4];
    const uint8_t* brick = a.bricks + (size_t)r.brick * 512u;
#pragma unroll
    for (int j = 0; j < 4; ++j) dd[j] = valid[j] ? brick[off[j]] : 0u;
    bool stop = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (valid[j] && !stop) {
            ++r.nv;
            if (dd[j]) {
                const uint32_t p = s_pal[dd[j]];
                const float k = 1.0f / 255.0f;
                float cr = (float)(p & 255u) * k, cg = (float)((p >> 8) & 255u) * k;
                float cb = (float)((p >> 16) & 255u) * k;
                const float ca = (float)(p >> 24) * k;
                cr = cr * ca;
                cg = cg * ca;
                cb = cb * ca;
                const float s1 = 1.0f - r.aa;
                r.ar = r.ar + cr * s1;
                r.ag = r.ag + cg * s1;
                r.ab = r.ab + cb * s1;
                r.aa = r.aa + ca * s1;
                if (r.first) {
                    r.first = false;
                    r.thit = pt[j];
                    r.hx = basex + (off[j] & 7);
                    r.hy = basey + ((off[j] >> 3) & 7);
                    r.hz = basez + (off[j] >> 6);
                }
                if (r.aa > 0.99f) {
                    stop = true;
                    end = true;
                }
            }
        }
    }
    if (!more) r.inb = false;  // left the brick (or the ray ended)
    return end;
}

// COUNT: the caller asked for the node / voxel read counts (without them the
// counters are dead and take no registers)
template <int L, bool COUNT>
__device__ __forceinline__ void nt_write(const NtArgs& a, size_t pix, const NtRay<L>& r) {
    if (a.rgba)
        a.rgba[pix] = (uint32_t)(nt_saturate(r.ar) * 255.0f) | ((uint32_t)(nt_saturate(r.ag) * 255.0f) << 8) |
                      ((uint32_t)(nt_saturate(r.ab) * 255.0f) << 16) |
                      ((uint32_t)(nt_saturate(r.aa) * 255.0f) << 24);
    if (a.t) a.t[pix] = r.thit;
    if (a.voxel) {
        a.voxel[3 * pix] = r.hx;
        a.voxel[3 * pix + 1] = r.hy;
        a.voxel[3 * pix + 2] = r.hz;
    }
    if constexpr (COUNT) {
        if (a.n_nodes) a.n_nodes[pix] = r.nn;
        if (a.n_voxels) a.n_voxels[pix] = r.nv;
    }
}

__device__ __forceinline__ void nt_stage(const NtArgs& a, uint32_t* s_root, uint32_t* s_pal) {
    const int nroot = a.R * a.R * a.R;
    for (int i = threadIdx.x; i < nroot; i += 256) s_root[i] = a.root[i];
    s_pal[threadIdx.x] = a.palette[threadIdx.x];
    __syncthreads();
}

// One lane per pixel; wave w covers 8x8 pixel tile w (row-major tiles).
template <int L, bool COUNT>
__global__ __launch_bounds__(256) void nt_trace_kernel(NtArgs a) {
    __shared__ uint32_t s_root[kNtMaxRoot * kNtMaxRoot * kNtMaxRoot];
    __shared__ uint32_t s_pal[256];
    nt_stage(a, s_root, s_pal);
    const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6;
    const int lane = (int)(threadIdx.x & 63);
    const int tiles_x = (a.w + 7) >> 3;
    const int x = (int)(wave % (uint32_t)tiles_x) * 8 + (lane & 7);
    const int y = (int)(wave / (uint32_t)tiles_x) * 8 + (lane >> 3);
    if (y >= a.h || x >= a.w) return;
    NtRay<L> r;
    if (nt_setup(a, x, y, r))
        while (!nt_iterate(a, s_root, s_pal, r)) {
        }
    nt_write<L, COUNT>(a, (size_t)y * a.w + x, r);
}

// Persistent waves over an atomic ray queue (ray ids in 8x8 pixel tiles).
// A lane whose ray finished parks until at least refill_min lanes of its
// wave wait (or none is walking); then the wave writes their outputs, takes
// that many ray ids with one atomic and sets the new rays up together, so
// short rays do not leave their lanes idle until the wave's longest ray ends.
// The last wave to retire resets the queue for the next launch.
template <int L>
__global__ __launch_bounds__(256) void nt_trace_persistent(NtArgs a) {
    __shared__ uint32_t s_root[kNtMaxRoot * kNtMaxRoot * kNtMaxRoot];
    __shared__ uint32_t s_pal[256];
    nt_stage(a, s_root, s_pal);
    const int lane = (int)(threadIdx.x & 63);
    const int tiles_x = (a.w + 7) >> 3;
    enum : int { kIdle = 0, kWalk = 1, kDone = 2, kRetired = 3 };
    int status = kIdle;
    bool drained = false;  // wave-uniform
    size_t pix = 0;
    NtRay<L> r;
    for (;;) {
        const uint64_t walking = __ballot(status == kWalk);
        const uint64_t waiting = __ballot(status == kIdle || status == kDone);
        if (waiting != 0 && (walking == 0 || __popcll(waiting) >= a.refill_min)) {
            if (status == kDone) {
                nt_write<L, true>(a, pix, r);
                status = kIdle;
            }
            const uint64_t need = __ballot(status == kIdle);
            if (need != 0) {
                if (drained) {
                    if (status == kIdle) status = kRetired;
                } else {
                    const uint32_t cnt = (uint32_t)__popcll(need);
                    const int leader = __ffsll((unsigned long long)need) - 1;
                    uint32_t got = 0;
                    if (lane == leader) got = atomicAdd(&a.counters[0], cnt);
                    got = __shfl(got, leader);
                    if (got + cnt >= a.total) drained = true;
                    if (status == kIdle) {
                        const uint32_t my = got + lanes_below(need);
                        if (my >= a.total) {
                            status = kRetired;
                        } else {
                            const uint32_t tile = my >> 6, within = my & 63u;
                            const int x = (int)(tile % (uint32_t)tiles_x) * 8 + (int)(within & 7u);
                            const int y = (int)(tile / (uint32_t)tiles_x) * 8 + (int)(within >> 3);
                            if (x < a.w && y < a.h) {
                                pix = (size_t)y * a.w + x;
                                status = nt_setup(a, x, y, r) ? kWalk : kDone;
                            }
                        }
                    }
                }
            }
        }
        if (__ballot(status != kRetired) == 0) break;
        if (status == kWalk && nt_iterate(a, s_root, s_pal, r)) status = kDone;
    }
    if (lane == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&a.counters[1], 1u);
        if (done + 1 == a.n_waves) {
            atomicExch(&a.counters[0], 0u);
            atomicExch(&a.counters[1], 0u);
            __threadfence();
        }
    }
}

}  // namespace yx

using namespace yx;

#define YX_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(YX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

struct yx_ntree {
    int size = 0, levels = 0, R = 0;
    uint32_t* d_root = nullptr;
    uint4* d_nodes = nullptr;
    uint8_t* d_bricks = nullptr;
    uint64_t n_nodes = 0, n_bricks = 0;
    uint32_t* d_counters = nullptr;  // ray queue of the persistent cast
    // one lane per pixel by default: measured 1.19 ms per 1080p frame of
    // brick_4096 against 1.37-2.5 ms for the persistent waves (refill 60-8);
    // a wave on one 8x8 tile keeps its rays' bricks and code paths together
    int persistent = 0;
    int refill_min = 16;
    int num_cus = 0;
};

static int nt_shape(int size, int levels, int* R) {
    auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
    if (!pow2(size) || size < 8 || size > 16384)
        return fail(YX_ERR_INVALID_ARG, "size %d: a power of two in [8, 16384]", size);
    if (levels < 0 || levels > kNtMaxLevels) return fail(YX_ERR_INVALID_ARG, "levels %d: 0..%d", levels, kNtMaxLevels);
    const int nb = size / 8;
    if ((nb >> (2 * levels)) < 1 || (nb >> (2 * levels)) << (2 * levels) != nb)
        return fail(YX_ERR_INVALID_ARG, "size %d has no 8 * 4^%d * R form", size, levels);
    *R = nb >> (2 * levels);
    if (*R > kNtMaxRoot) return fail(YX_ERR_INVALID_ARG, "root grid %d^3 exceeds %d^3: use more levels", *R, kNtMaxRoot);
    return YX_OK;
}

// RAII set of device buffers for one build.
struct DevBufs {
    std::vector<void*> ptrs;
    ~DevBufs() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** p, size_t n) {
        *p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) {
            *p = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%zu) in the n-tree build", n * sizeof(T));
        }
        ptrs.push_back(*p);
        return YX_OK;
    }
};

static int nt_scan(const uint8_t* flags, uint32_t* out, uint64_t n, DevBufs& tmp, hipStream_t st) {
    if (n > 0xFFFFFFFFull) return fail(YX_ERR_RANGE, "more than 2^32 units on one level");
    const size_t bytes = scan_workspace_bytes(n);
    void* ws = nullptr;
    int rc = tmp.alloc((uint8_t**)&ws, bytes);
    if (rc != YX_OK) return rc;
    return scan_excl_u8(flags, out, n, nullptr, ws, bytes, st);
}

// count of set flags = scan[n-1] + flags[n-1]
static int nt_count(const uint8_t* flags, const uint32_t* scan, uint64_t n, hipStream_t st, uint64_t* count) {
    uint32_t s = 0;
    uint8_t f = 0;
    YX_HIP(hipMemcpyAsync(&s, scan + (n - 1), 4, hipMemcpyDeviceToHost, st));
    YX_HIP(hipMemcpyAsync(&f, flags + (n - 1), 1, hipMemcpyDeviceToHost, st));
    YX_HIP(hipStreamSynchronize(st));
    *count = (uint64_t)s + f;
    return YX_OK;
}

static int nt_build(const NtSource& src, int size, int levels, yx_ntree** out, hipStream_t st) {
    int R = 0;
    int rc = nt_shape(size, levels, &R);
    if (rc != YX_OK) return rc;
    int dev = 0, cus = 0;
    YX_HIP(hipGetDevice(&dev));
    YX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int L = levels;
    const uint64_t nb = (uint64_t)size / 8;
    const uint64_t n_bricks_total = nb * nb * nb;
    DevBufs tmp;
    // per level u = 1..L (nodes) and L+1 (bricks): flags, scans, masks
    std::vector<uint64_t> units((size_t)L + 2, 0);
    std::vector<uint8_t*> flags((size_t)L + 2, nullptr);
    std::vector<uint32_t*> scans((size_t)L + 2, nullptr);
    std::vector<uint64_t*> masks((size_t)L + 2, nullptr);
    units[(size_t)L + 1] = n_bricks_total;
    for (int l = L; l >= 1; --l) units[(size_t)l] = units[(size_t)l + 1] / 64;
    for (int l = 1; l <= L + 1; ++l) {
        if ((rc = tmp.alloc(&flags[(size_t)l], units[(size_t)l])) != YX_OK) return rc;
        if ((rc = tmp.alloc(&scans[(size_t)l], units[(size_t)l])) != YX_OK) return rc;
        if (l <= L && (rc = tmp.alloc(&masks[(size_t)l], units[(size_t)l])) != YX_OK) return rc;
    }
    const dim3 blk(256);
    const uint32_t persistent = (uint32_t)cus * 8u;
    auto grid_for = [&](uint64_t waves) {
        return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(persistent, (waves + 3) / 4)));
    };
    hipLaunchKernelGGL(nt_classify_kernel, grid_for(n_bricks_total), blk, 0, st, src, L, R, n_bricks_total,
                       flags[(size_t)L + 1]);
    YX_HIP(hipGetLastError());
    for (int l = L; l >= 1; --l) {
        hipLaunchKernelGGL(nt_level_kernel, grid_for(units[(size_t)l]), blk, 0, st, flags[(size_t)l + 1],
                           units[(size_t)l], flags[(size_t)l], masks[(size_t)l]);
        YX_HIP(hipGetLastError());
    }
    std::vector<uint64_t> count((size_t)L + 2, 0), offset((size_t)L + 2, 0);
    for (int l = 1; l <= L + 1; ++l) {
        if ((rc = nt_scan(flags[(size_t)l], scans[(size_t)l], units[(size_t)l], tmp, st)) != YX_OK) return rc;
        if ((rc = nt_count(flags[(size_t)l], scans[(size_t)l], units[(size_t)l], st, &count[(size_t)l])) != YX_OK)
            return rc;
    }
    uint64_t total_nodes = 0;
    for (int l = 1; l <= L; ++l) {
        offset[(size_t)l] = total_nodes;
        total_nodes += count[(size_t)l];
    }
    const uint64_t total_bricks = count[(size_t)L + 1];
    if (total_nodes >= kNtEmpty || total_bricks >= kNtEmpty) return fail(YX_ERR_RANGE, "n-tree too large");
    yx_ntree* t = new (std::nothrow) yx_ntree();
    if (!t) return fail(YX_ERR_OOM, "out of host memory");
    t->size = size;
    t->levels = L;
    t->R = R;
    t->n_nodes = total_nodes;
    t->n_bricks = total_bricks;
    t->num_cus = cus;
    auto release = [&](int code) {
        yx_ntree_free(t);
        return code;
    };
    if (hipMalloc((void**)&t->d_counters, 2 * sizeof(uint32_t)) != hipSuccess ||
        hipMemsetAsync(t->d_counters, 0, 2 * sizeof(uint32_t), st) != hipSuccess)
        return release(fail(YX_ERR_OOM, "hipMalloc for the n-tree ray queue"));
    if (hipMalloc((void**)&t->d_root, (size_t)R * R * R * 4) != hipSuccess ||
        hipMalloc((void**)&t->d_nodes, std::max<uint64_t>(1, total_nodes) * sizeof(uint4)) != hipSuccess ||
        hipMalloc((void**)&t->d_bricks, std::max<uint64_t>(1, total_bricks) * 512u) != hipSuccess)
        return release(fail(YX_ERR_OOM, "hipMalloc for the n-tree (%llu nodes, %llu bricks)",
                            (unsigned long long)total_nodes, (unsigned long long)total_bricks));
    for (int l = 1; l <= L; ++l) {
        const uint32_t child_off = l < L ? (uint32_t)offset[(size_t)l + 1] : 0u;
        hipLaunchKernelGGL(nt_write_nodes_kernel, grid_for(units[(size_t)l] / 64 + 1), blk, 0, st, flags[(size_t)l],
                           scans[(size_t)l], masks[(size_t)l], units[(size_t)l], (uint32_t)offset[(size_t)l],
                           scans[(size_t)l + 1], child_off, t->d_nodes);
        if (hipGetLastError() != hipSuccess) return release(fail(YX_ERR_HIP, "n-tree node write launch"));
    }
    const uint32_t nroot = (uint32_t)(R * R * R);
    hipLaunchKernelGGL(nt_write_root_kernel, dim3((nroot + 255) / 256), blk, 0, st, flags[1], scans[1], nroot,
                       L >= 1 ? (uint32_t)offset[1] : 0u, t->d_root);
    hipLaunchKernelGGL(nt_pack_kernel, grid_for(n_bricks_total), blk, 0, st, src, L, R, n_bricks_total,
                       flags[(size_t)L + 1], scans[(size_t)L + 1], t->d_bricks);
    if (hipGetLastError() != hipSuccess) return release(fail(YX_ERR_HIP, "n-tree pack launch"));
    if (hipStreamSynchronize(st) != hipSuccess) return release(fail(YX_ERR_HIP, "n-tree build"));
    *out = t;
    return YX_OK;
}

extern "C" int yx_ntree_build_fbm(const yx_ntree_desc* d, yx_ntree** out, void* stream) {
    clear_error();
    if (!d || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (d->octaves < 1 || d->octaves > 16) return fail(YX_ERR_INVALID_ARG, "octaves %d: 1..16", d->octaves);
    if (!(d->scale >= 0.0f) || d->frequency != d->frequency || d->threshold != d->threshold)
        return fail(YX_ERR_INVALID_ARG, "frequency, threshold and scale must be numbers, scale >= 0");
    int R = 0;
    const int shape = nt_shape(d->size, d->levels, &R);
    if (shape != YX_OK) return shape;
    const hipStream_t st = (hipStream_t)stream;
    Perlin perlin(d->seed);
    int* d_perm = nullptr;
    YX_HIP(hipMalloc((void**)&d_perm, 512 * sizeof(int)));
    hipError_t e = hipMemcpy(d_perm, perlin.perm, 512 * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d_perm);
        return fail(YX_ERR_HIP, "perm upload: %s", hipGetErrorString(e));
    }
    NtSource s;
    s.q.inv = 1.0f / (float)d->size;
    s.q.frequency = d->frequency;
    s.q.threshold = d->threshold;
    s.q.scale = d->scale;
    s.q.octaves = d->octaves;
    s.perm = d_perm;
    s.dense = nullptr;
    s.size = d->size;
    const int rc = nt_build(s, d->size, d->levels, out, st);
    (void)hipFree(d_perm);
    return rc;
}

extern "C" int yx_ntree_from_dense(const uint8_t* dense, int32_t size, int32_t levels, yx_ntree** out) {
    clear_error();
    if (!dense || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int R = 0;
    int rc = nt_shape(size, levels, &R);
    if (rc != YX_OK) return rc;
    const size_t n = (size_t)size * size * size;
    uint8_t* d_dense = nullptr;
    YX_HIP(hipMalloc((void**)&d_dense, n));
    hipError_t e = hipMemcpy(d_dense, dense, n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d_dense);
        return fail(YX_ERR_HIP, "volume upload: %s", hipGetErrorString(e));
    }
    NtSource s;
    std::memset(&s, 0, sizeof(s));
    s.dense = d_dense;
    s.size = size;
    rc = nt_build(s, size, levels, out, nullptr);
    (void)hipFree(d_dense);
    return rc;
}

extern "C" void yx_ntree_free(yx_ntree* t) {
    if (!t) return;
    if (t->d_root) (void)hipFree(t->d_root);
    if (t->d_nodes) (void)hipFree(t->d_nodes);
    if (t->d_bricks) (void)hipFree(t->d_bricks);
    if (t->d_counters) (void)hipFree(t->d_counters);
    delete t;
}

extern "C" int yx_ntree_set_launch(yx_ntree* t, int32_t persistent, int32_t refill_min) {
    clear_error();
    if (!t) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (refill_min < 1 || refill_min > 64) return fail(YX_ERR_INVALID_ARG, "refill_min must be 1..64");
    t->persistent = persistent != 0;
    t->refill_min = refill_min;
    return YX_OK;
}

extern "C" int yx_ntree_get_info(const yx_ntree* t, yx_ntree_info* info) {
    clear_error();
    if (!t || !info) return fail(YX_ERR_INVALID_ARG, "null argument");
    info->size = t->size;
    info->levels = t->levels;
    info->root_side = t->R;
    info->num_nodes = t->n_nodes;
    info->num_bricks = t->n_bricks;
    info->bytes = (uint64_t)t->R * t->R * t->R * 4u + t->n_nodes * 16u + t->n_bricks * 512u;
    return YX_OK;
}

extern "C" int yx_ntree_download(const yx_ntree* t, uint32_t* root, uint32_t* nodes, uint8_t* bricks) {
    clear_error();
    if (!t) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (root) YX_HIP(hipMemcpy(root, t->d_root, (size_t)t->R * t->R * t->R * 4u, hipMemcpyDeviceToHost));
    if (nodes && t->n_nodes) YX_HIP(hipMemcpy(nodes, t->d_nodes, t->n_nodes * 16u, hipMemcpyDeviceToHost));
    if (bricks && t->n_bricks) YX_HIP(hipMemcpy(bricks, t->d_bricks, t->n_bricks * 512u, hipMemcpyDeviceToHost));
    return YX_OK;
}

extern "C" int yx_ntree_render(const yx_ntree* t, const float view[12], int32_t w, int32_t h,
                               const uint32_t* palette, const yx_ntree_outputs* out, void* stream) {
    clear_error();
    if (!t || !view || !palette || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (w <= 0 || h <= 0 || (int64_t)w * h > ((int64_t)1 << 28))
        return fail(YX_ERR_INVALID_ARG, "resolution %dx%d", w, h);
    NtArgs a;
    std::memset(&a, 0, sizeof(a));
    a.root = t->d_root;
    a.nodes = t->d_nodes;
    a.bricks = t->d_bricks;
    a.palette = palette;
    std::memcpy(a.view, view, sizeof(a.view));
    a.size = t->size;
    a.R = t->R;
    a.w = w;
    a.h = h;
    a.rgba = out->rgba;
    a.t = out->t;
    a.voxel = out->voxel;
    a.n_nodes = out->nodes;
    a.n_voxels = out->voxels;
    const uint32_t waves = (uint32_t)(((w + 7) / 8) * ((h + 7) / 8));
    const dim3 blk(256);
    const hipStream_t st = (hipStream_t)stream;
    if (!t->persistent) {
        const dim3 grid((waves + 3) / 4);
        const bool count = a.n_nodes || a.n_voxels;
#define YX_NT_CAST(L)                                                                \
    do {                                                                             \
        if (count)                                                                   \
            hipLaunchKernelGGL((nt_trace_kernel<L, true>), grid, blk, 0, st, a);     \
        else                                                                         \
            hipLaunchKernelGGL((nt_trace_kernel<L, false>), grid, blk, 0, st, a);    \
    } while (0)
        switch (t->levels) {
            case 0: YX_NT_CAST(0); break;
            case 1: YX_NT_CAST(1); break;
            case 2: YX_NT_CAST(2); break;
            case 3: YX_NT_CAST(3); break;
            default: YX_NT_CAST(4); break;
        }
#undef YX_NT_CAST
        YX_HIP(hipGetLastError());
        return YX_OK;
    }
    a.counters = t->d_counters;
    a.total = waves * 64u;
    a.refill_min = t->refill_min;
    const void* fn = nullptr;
    void (*kern)(NtArgs) = nullptr;
    switch (t->levels) {
        case 0: kern = nt_trace_persistent<0>; break;
        case 1: kern = nt_trace_persistent<1>; break;
        case 2: kern = nt_trace_persistent<2>; break;
        case 3: kern = nt_trace_persistent<3>; break;
        default: kern = nt_trace_persistent<4>; break;
    }
    fn = (const void*)kern;
    int occ = 0;
    YX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 256, 0));
    const uint32_t grid = (uint32_t)(std::max(1, occ) * t->num_cus);
    a.n_waves = grid * 4u;
    hipLaunchKernelGGL(kern, dim3(grid), blk, 0, st, a);
    YX_HIP(hipGetLastError());
    return YX_OK;
}
