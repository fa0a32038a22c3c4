// Device-side SVO traversal for gfx950 (HIP). Included by the trace kernels.
//
// The traversal is the iterative form of the reference's recursive RecTrace
// (cell/ppu_renderer.cpp:18-41) with FindFirstChild / GoNext as in the SPE
// twins (cell/spu/trace_spu.cpp:48-93), SetupTrace as in
// cell/spu/trace_spu.c_:79-99 and AdjustDir as ntree/ntree_trace.cu:38-50.
// Every float operation is the same, in the same order, as the CPU restatement
// the tests compare against, so hit ids and voxel coordinates are bit-exact;
// the library is built with -ffp-contract=off and correctly rounded fp32
// division and square root.
//
// Node records (8 bytes, uint2) are the unified breadth-first layout:
//   node: x = first child record, y = child_mask | leaf_mask << 8
//   leaf: x = RGBA8 payload,      y = canonical parent id
// A present child c lives at x + popcount((child_mask|leaf_mask) & ((1<<c)-1)).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace yx {

struct Ray3 {
    float x, y, z;
};

__device__ __forceinline__ float dmax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float dmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float max3(float x, float y, float z) { return dmax(x, dmax(y, z)); }
__device__ __forceinline__ float min3(float x, float y, float z) { return dmin(x, dmin(y, z)); }

__device__ __forceinline__ float adjust_coord(float x) {
    const float eps = 1e-6f;
    if (fabsf(x) < eps) x = (x < 0) ? -eps : eps;
    return x;
}

// normalized(dir0 + du*x + dv*y), cell/ppu_renderer.cpp:56; `len` receives
// |dir0 + du*x + dv*y| (the ray's view-depth scale: z = t / len)
__device__ __forceinline__ Ray3 pixel_dir_n(const float* b, int x, int y, float& len) {
    float xf = (float)x, yf = (float)y;
    float dx = b[0] + b[3] * xf, dy = b[1] + b[4] * xf, dz = b[2] + b[5] * xf;
    dx = dx + b[6] * yf;
    dy = dy + b[7] * yf;
    dz = dz + b[8] * yf;
    float n2 = 0.0f;
    n2 += dx * dx;
    n2 += dy * dy;
    n2 += dz * dz;
    float n = __builtin_sqrtf(n2);
    len = n;
    Ray3 r;
    r.x = dx / n;
    r.y = dy / n;
    r.z = dz / n;
    return r;
}

__device__ __forceinline__ Ray3 pixel_dir(const float* b, int x, int y) {
    float len;
    return pixel_dir_n(b, x, y, len);
}

// Exact per-ray traversal state of one level.
struct Cell {
    float t1x, t1y, t1z, t2x, t2y, t2z;
};

__device__ __forceinline__ bool setup_trace(const float p[3], const Ray3& d, Cell& c, uint32_t& flags) {
    flags = 0;
    // plane x = 0 at a, plane x = 1 at b; entry is the nearer one
    const float ax = (-p[0]) / d.x, bx = (1.0f - p[0]) / d.x;
    const float ay = (-p[1]) / d.y, by = (1.0f - p[1]) / d.y;
    const float az = (-p[2]) / d.z, bz = (1.0f - p[2]) / d.z;
    c.t1x = (d.x > 0) ? ax : bx;
    c.t2x = (d.x > 0) ? bx : ax;
    c.t1y = (d.y > 0) ? ay : by;
    c.t2y = (d.y > 0) ? by : ay;
    c.t1z = (d.z > 0) ? az : bz;
    c.t2z = (d.z > 0) ? bz : az;
    if (d.x < 0) flags |= 1u;
    if (d.y < 0) flags |= 2u;
    if (d.z < 0) flags |= 4u;
    return max3(c.t1x, c.t1y, c.t1z) < min3(c.t2x, c.t2y, c.t2z);
}

// FindFirstChild, cell/spu/trace_spu.cpp:48-68
__device__ __forceinline__ int find_first_child(Cell& c) {
    float mx = 0.5f * (c.t1x + c.t2x);
    float my = 0.5f * (c.t1y + c.t2y);
    float mz = 0.5f * (c.t1z + c.t2z);
    float te = max3(c.t1x, c.t1y, c.t1z);
    bool bx = te > mx, by = te > my, bz = te > mz;
    c.t1x = bx ? mx : c.t1x;  c.t2x = bx ? c.t2x : mx;
    c.t1y = by ? my : c.t1y;  c.t2y = by ? c.t2y : my;
    c.t1z = bz ? mz : c.t1z;  c.t2z = bz ? c.t2z : mz;
    return (bx ? 1 : 0) | (by ? 2 : 0) | (bz ? 4 : 0);
}

// Exit axis of a cell: argmin(t2) with the tie order of GoNext,
// cell/spu/trace_spu.cpp:75-78.
__device__ __forceinline__ int exit_axis(const Cell& c) {
    if (c.t2x > c.t2y) return (c.t2y < c.t2z) ? 1 : 2;
    return (c.t2x < c.t2z) ? 0 : 2;
}

// GoNext, cell/spu/trace_spu.cpp:70-93: leave through the exit plane into the
// sibling, or fail when the current child already lies on that side. Written
// without branches on the axis (selects), so lanes of a wave that leave
// through different planes run one instruction stream; the float operations
// are the reference's: dt = t2 - t1, t1 = t2, t2 += dt on the exit axis.
__device__ __forceinline__ bool go_next(int& ch, Cell& c) {
    const bool gt = c.t2x > c.t2y;
    const bool ex = !gt && (c.t2x < c.t2z);
    const bool ey = gt && (c.t2y < c.t2z);
    const int mask = ex ? 1 : (ey ? 2 : 4);
    if (ch & mask) return false;
    ch ^= mask;
    const float t1 = ex ? c.t1x : (ey ? c.t1y : c.t1z);
    const float t2 = ex ? c.t2x : (ey ? c.t2y : c.t2z);
    const float dt = t2 - t1;
    const float n2 = t2 + dt;
    const bool ez = !ex && !ey;
    c.t1x = ex ? t2 : c.t1x;
    c.t2x = ex ? n2 : c.t2x;
    c.t1y = ey ? t2 : c.t1y;
    c.t2y = ey ? n2 : c.t2y;
    c.t1z = ez ? t2 : c.t1z;
    c.t2z = ez ? n2 : c.t2z;
    return true;
}

// Entry-face axis of the hit cell: argmax(t1) with the tie order of
// bittree/trace.cu:100-106.
__device__ __forceinline__ int entry_axis(const Cell& c) {
    if (c.t1x < c.t1y) return (c.t1y > c.t1z) ? 1 : 2;
    return (c.t1x > c.t1z) ? 0 : 2;
}

// Lambert over the RGBA8 payload; DEFINED HERE (the reference shader source is
// missing): k = 0.1 + 0.9 * max(0, n.L) * visible, n = entry-face normal;
// channel = (uint8)(c * k + 0.5), alpha 255.
__device__ __forceinline__ uint32_t shade_rgba(uint32_t leaf, float k) {
    uint32_t out = 0xff000000u;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = (float)((leaf >> (8 * c)) & 0xffu) * k + 0.5f;
        out |= ((uint32_t)v & 0xffu) << (8 * c);
    }
    return out;
}

// ---- the reference's leaf payload: VoxData = colour16 + normal16 ----------
// (reaction/report/main.tex:38,65; unpacked by UnpackVoxData / UnpackColor /
// UnpackNormal, demo/SVORenderer.cpp:181-185, whose bodies are missing). The
// bit layout is DEFINED HERE:
//   bits 0-15:  colour, RGB565 (red in bits 11-15), widened to 8 bits by bit
//               replication, alpha 255;
//   bits 16-31: normal, quantized as in Baptista, "Higher Accuracy Quantized
//               Normals" (the thesis's citation): bits 15, 14, 13 of the half
//               are the signs of x, y, z; bits 7-12 (u) and 0-6 (v) a point of
//               the triangle u + v <= 126 of (|x|, |y|) / (|x| + |y| + |z|) *
//               126, points with u >= 64 folded to (127 - u, 127 - v); the
//               normal is (u, v, 126 - u - v) with the signs, normalised.
// Every float operation is written out in order so the oracle's restatement
// gives the same bits.
__device__ __forceinline__ uint32_t rgba_from_565(uint32_t c) {
    const uint32_t r5 = (c >> 11) & 31u, g6 = (c >> 5) & 63u, b5 = c & 31u;
    const uint32_t r = (r5 << 3) | (r5 >> 2), g = (g6 << 2) | (g6 >> 4), b = (b5 << 3) | (b5 >> 2);
    return r | (g << 8) | (b << 16) | 0xff000000u;
}

__device__ __forceinline__ void normal_from_16(uint32_t q, float& nx, float& ny, float& nz) {
    int u = (int)((q >> 7) & 63u), v = (int)(q & 127u);
    if (u + v > 126) {
        u = 127 - u;
        v = 127 - v;
    }
    const float x = (float)u, y = (float)v, z = (float)(126 - u - v);
    float n2 = 0.0f;
    n2 += x * x;
    n2 += y * y;
    n2 += z * z;
    const float len = __builtin_sqrtf(n2);
    nx = (q & 0x8000u) ? -(x / len) : x / len;
    ny = (q & 0x4000u) ? -(y / len) : y / len;
    nz = (q & 0x2000u) ? -(z / len) : z / len;
}

// Lambert term max(0, n.L) of a colour16 + normal16 leaf toward unit L.
__device__ __forceinline__ float lambert_c16n16(uint32_t leaf, float lx, float ly, float lz) {
    float nx, ny, nz;
    normal_from_16(leaf >> 16, nx, ny, nz);
    float d = 0.0f;
    d += nx * lx;
    d += ny * ly;
    d += nz * lz;
    return dmax(d, 0.0f);
}

struct HitInfo {
    uint32_t leaf_rec;  // record index of the hit leaf
    int child;          // real child slot
    int level;          // leaf level (root children = 1)
    int vx, vy, vz;
    Cell cell;          // hit cell (t1 = entry times)
};

// Per-lane traversal stack. Level l holds the parent state pushed when
// descending from level l to l+1.
template <int MAXD>
struct PrivStack {
    float t[MAXD][6];
    uint2 node[MAXD];
    int ch[MAXD];
    __device__ __forceinline__ void push(int l, const Cell& c, uint2 n, int h) {
        t[l][0] = c.t1x; t[l][1] = c.t1y; t[l][2] = c.t1z;
        t[l][3] = c.t2x; t[l][4] = c.t2y; t[l][5] = c.t2z;
        node[l] = n;
        ch[l] = h;
    }
    __device__ __forceinline__ void pop(int l, Cell& c, uint2& n, int& h) const {
        c.t1x = t[l][0]; c.t1y = t[l][1]; c.t1z = t[l][2];
        c.t2x = t[l][3]; c.t2y = t[l][4]; c.t2z = t[l][5];
        n = node[l];
        h = ch[l];
    }
};

// One ray from p along d (already AdjustDir'ed). occlusion: leaf cells behind
// the origin are skipped (shadow rays). Returns true on a hit.
template <int MAXD>
__device__ bool trace_ray(const uint2* __restrict__ recs, const float p[3], const Ray3& d,
                          bool occlusion, HitInfo& hit, uint32_t& visits) {
    Cell c;
    uint32_t flags;
    if (!setup_trace(p, d, c, flags)) return false;
    if (min3(c.t2x, c.t2y, c.t2z) <= 0) return false;
    PrivStack<MAXD> stk;
    ++visits;
    uint2 node = recs[0];
    int level = 0;
    int px = 0, py = 0, pz = 0;
    int ch = find_first_child(c);
    for (;;) {
        const int real = ch ^ (int)flags;
        const uint32_t cm = node.y & 0xffu, lm = (node.y >> 8) & 0xffu;
        const uint32_t below = (cm | lm) & ((1u << real) - 1u);
        if ((lm >> real) & 1u) {
            if (!occlusion || min3(c.t2x, c.t2y, c.t2z) > 0) {
                hit.leaf_rec = node.x + (uint32_t)__builtin_popcount(below);
                hit.child = real;
                hit.level = level + 1;
                hit.vx = 2 * px + (real & 1);
                hit.vy = 2 * py + ((real >> 1) & 1);
                hit.vz = 2 * pz + ((real >> 2) & 1);
                hit.cell = c;
                return true;
            }
        } else if (((cm >> real) & 1u) && min3(c.t2x, c.t2y, c.t2z) > 0) {
            ++visits;
            stk.push(level, c, node, ch);
            node = recs[node.x + (uint32_t)__builtin_popcount(below)];
            px = 2 * px + (real & 1);
            py = 2 * py + ((real >> 1) & 1);
            pz = 2 * pz + ((real >> 2) & 1);
            ++level;
            ch = find_first_child(c);
            continue;
        }
        while (!go_next(ch, c)) {
            if (level == 0) return false;
            --level;
            stk.pop(level, c, node, ch);
            px >>= 1;
            py >>= 1;
            pz >>= 1;
        }
    }
}

// Workgroup order for one-tile-per-wave launches. The hardware hands
// workgroup b to XCD b mod 8, so consecutive workgroups land on different
// XCDs, each with its own L2. Workgroup b instead takes block b / 8 of XCD
// (b mod 8)'s share, one contiguous eighth of the grid, so each XCD walks one
// run of neighbouring tiles and its L2 holds what those rays share. A
// bijection on the first 8 * floor(nb / 8) blocks; the rest keep their index.
// Callers whose outputs depend only on the tile get the same results in any
// order.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t full = nb & ~7u;
    if (b >= full) return b;
    return (b & 7u) * (full >> 3) + (b >> 3);
}

}  // namespace yx
