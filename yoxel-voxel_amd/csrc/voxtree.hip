// Brick classification of a bit volume on the GPU: the voxtree pipeline of
// reaction/voxtree/voxtree.py:80-95 over the kernels of
// reaction/voxtree/density.cu (CalcDensity :32-60, MarkBricks :131-176,
// PackBricks :210-221), with the column-start scan that the reference does on
// the host (cumsum_ex, voxtree.py:12-18) done on the GPU.
//
// Formats (B = bit volume side, V = B / 2, G = V / 4 bricks per axis):
// - bit volume: uint32 words [B / 128 slices][B rows (y)][B (x)][4 layers];
//   bit b of layer l of slice s is voxel z = 128 s + 32 l + b;
// - density: 4-bit counts (0..8) of the 2x2x2 bit blocks, uint32 words
//   [V (y)][V (x)][V / 8]; word w of a column holds z = 8 w .. 8 w + 7, one
//   count per nibble from the low nibble up;
// - brick state: uint32 [G (y)][G (x)][G (z)]. A brick covers the 5^3 counts
//   from 4 * its index (one count of overlap on each upper face; counts beyond
//   the volume are 0). EMPTY (0x40000000) when all are 0, SOLID (0x80000000)
//   when all are 8, else mixed. The low bits hold the number of mixed bricks
//   below it in its column (also for uniform bricks, which OR in their bit);
// - column count / start: uint32 [G (y)][G (x)], mixed bricks per column and
//   their exclusive prefix sum over columns (yx_voxtree_classify appends the
//   total as start[G * G]);
// - packed: the mixed bricks as x | y << 8 | z << 16, by column, then by z.
//
// The reference's MarkBricks reads a neighbour's word from shared memory in a
// loop without a barrier between iterations; here the neighbour word comes from
// a lane shuffle inside the 64-wide wavefront, one wavefront per brick column.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"

#define VT_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(YX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

namespace yx {
namespace {

constexpr uint32_t kSolid = 0x80000000u;
constexpr uint32_t kEmpty = 0x40000000u;
constexpr uint32_t kUniform = kSolid | kEmpty;

// pairwise sums of adjacent bits: 16 two-bit fields
__device__ __forceinline__ uint32_t vt_pairs(uint32_t v) { return ((v >> 1) & 0x55555555u) + (v & 0x55555555u); }

// interleave the nibbles of the even and odd field sums: nibble 2m from even
// nibble m, nibble 2m + 1 from odd nibble m (m = 0..3)
__device__ __forceinline__ uint32_t vt_interleave(uint32_t even, uint32_t odd) {
    uint32_t r = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        r |= ((even >> (4 * m)) & 0xFu) << (8 * m);
        r |= ((odd >> (4 * m)) & 0xFu) << (8 * m + 4);
    }
    return r;
}

// One thread per (slice, x, y) of the half-resolution grid, slice fastest: it
// reads the 4 layers of bit columns (2x, 2y), (2x+1, 2y), (2x, 2y+1),
// (2x+1, 2y+1) as two 32-byte runs and writes the 4 layers' 8 density words
// as one 32-byte run.
__global__ __launch_bounds__(256) void vt_density_kernel(const uint4* __restrict__ bits, int B, int S,
                                                         uint4* __restrict__ dens) {
    const int V = B >> 1;
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (size_t)S * V * V) return;
    const int s = (int)(t % (size_t)S);
    const int x = (int)((t / (size_t)S) % (size_t)V);
    const int y = (int)(t / ((size_t)S * V));
    // word index ((s * B + row) * B + col) * 4 + layer, in uint4 units: / 4
    const size_t r0 = ((size_t)s * B + 2 * y) * B + 2 * x;
    const size_t r1 = r0 + (size_t)B;
    const uint4 a0 = bits[r0], a1 = bits[r0 + 1], b0 = bits[r1], b1 = bits[r1 + 1];
    const uint32_t w00[4] = {a0.x, a0.y, a0.z, a0.w}, w01[4] = {a1.x, a1.y, a1.z, a1.w};
    const uint32_t w10[4] = {b0.x, b0.y, b0.z, b0.w}, w11[4] = {b1.x, b1.y, b1.z, b1.w};
    uint32_t out[8];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const uint32_t v0 = vt_pairs(w00[l]), v1 = vt_pairs(w01[l]), v2 = vt_pairs(w10[l]), v3 = vt_pairs(w11[l]);
        const uint32_t even = (v0 & 0x33333333u) + (v1 & 0x33333333u) + (v2 & 0x33333333u) + (v3 & 0x33333333u);
        const uint32_t odd = ((v0 >> 2) & 0x33333333u) + ((v1 >> 2) & 0x33333333u) + ((v2 >> 2) & 0x33333333u) +
                             ((v3 >> 2) & 0x33333333u);
        out[2 * l] = vt_interleave(even, odd);
        out[2 * l + 1] = vt_interleave(even >> 16, odd >> 16);
    }
    // density words of column (x, y): (y * V + x) * (V / 8), slice s starts at 8 s
    uint4* d = dens + (((size_t)y * V + x) * (size_t)S + s) * 2;
    d[0] = make_uint4(out[0], out[1], out[2], out[3]);
    d[1] = make_uint4(out[4], out[5], out[6], out[7]);
}

// state of 5 consecutive counts (the low 5 nibbles of v)
__device__ __forceinline__ uint32_t vt_state5(uint32_t v) {
    uint32_t st = 0;
    if ((v & 0x88888u) == 0x88888u) st = kSolid;
    if ((v & 0xFFFFFu) == 0) st = kEmpty;
    return st;
}

// One wavefront per brick column; lane l holds density word l (z = 8l..8l+7),
// which covers bricks 2l (z 8l..8l+4) and 2l+1 (z 8l+4..8l+8, the next word's
// first nibble). W = V / 8 <= 64 words per column.
__global__ __launch_bounds__(256) void vt_mark_kernel(const uint32_t* __restrict__ dens, int V,
                                                      uint2* __restrict__ state, uint32_t* __restrict__ col_count) {
    const int G = V >> 2, W = V >> 3;
    const int lane = threadIdx.x & 63;
    const size_t col = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (col >= (size_t)G * G) return;  // whole wavefronts leave together
    const int bx = (int)(col & (size_t)(G - 1)), by = (int)(col / (size_t)G);
    const bool live = lane < W;
    // all 25 loads first, so they are in flight together
    uint32_t v[25];
#pragma unroll
    for (int q = 0; q < 25; ++q) {
        const int y = 4 * by + q / 5, x = 4 * bx + q % 5;
        v[q] = (live && x < V && y < V) ? dens[((size_t)y * V + x) * W + lane] : 0u;
    }
    uint32_t s1 = kUniform, s2 = kUniform;
#pragma unroll
    for (int q = 0; q < 25; ++q) {
        uint32_t next = __shfl_down(v[q], 1, 64);
        if (lane + 1 >= W) next = 0;  // past the top of the column
        s1 &= vt_state5(v[q]);
        s2 &= vt_state5((next << 16) | (v[q] >> 16));
    }
    const bool m1 = live && s1 == 0, m2 = live && s2 == 0;
    const unsigned long long b1 = __ballot(m1), b2 = __ballot(m2);
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t p1 = (uint32_t)(__popcll(b1 & below) + __popcll(b2 & below));
    const uint32_t p2 = p1 + (m1 ? 1u : 0u);
    if (live) state[col * (size_t)W + lane] = make_uint2(p1 | s1, p2 | s2);
    if (lane == 0) col_count[col] = (uint32_t)(__popcll(b1) + __popcll(b2));
}

// Exclusive prefix sum of the column counts (4 n4 of them) in one workgroup
// of 1024 threads, 4096 counts per step: each thread loads 4 as one 16-byte
// word, wavefronts scan with lane shuffles, the 16 wavefront totals are
// scanned in LDS, and a running carry joins the steps. Writes the grand total
// to *total.
__global__ __launch_bounds__(1024) void vt_scan_kernel(const uint4* __restrict__ cnt, uint32_t n4,
                                                       uint4* __restrict__ start, uint32_t* __restrict__ total) {
    __shared__ uint32_t wsum[16];
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n4; base += 1024) {
        const uint32_t i = base + t;
        const uint4 c = i < n4 ? cnt[i] : make_uint4(0u, 0u, 0u, 0u);
        const uint32_t s = c.x + c.y + c.z + c.w;
        uint32_t inc = s;  // inclusive scan over the wavefront
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = __shfl_up(inc, off, 64);
            if (lane >= (uint32_t)off) inc += o;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        if (w == 0) {
            uint32_t v = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                const uint32_t o = __shfl_up(v, off, 64);
                if (lane >= (uint32_t)off) v += o;
            }
            if (lane < 16) wsum[lane] = v;
        }
        __syncthreads();
        const uint32_t ex = carry + (w ? wsum[w - 1] : 0u) + inc - s;
        if (i < n4) start[i] = make_uint4(ex, ex + c.x, ex + c.x + c.y, ex + c.x + c.y + c.z);
        carry += wsum[15];
        __syncthreads();  // wsum is rewritten by the next step
    }
    if (t == 0) *total = carry;
}

// One thread per 4 bricks of a column (one 16-byte load): mixed bricks write
// their packed position. G is a power of two, lg its log2.
__global__ __launch_bounds__(256) void vt_pack_kernel(const uint4* __restrict__ state,
                                                      const uint32_t* __restrict__ col_start, int lg,
                                                      uint32_t* __restrict__ packed, uint32_t cap) {
    const uint32_t G = 1u << lg;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;  // G^3 / 4 <= 2^19 quads
    if (i >= (G * G * G) >> 2) return;
    const uint4 q = state[i];
    const uint32_t st[4] = {q.x, q.y, q.z, q.w};
    const uint32_t col = (i * 4) >> lg, z0 = (i * 4) & (G - 1);
    const uint32_t x = col & (G - 1), y = col >> lg;
    const uint32_t base = col_start[col];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (st[k] & kUniform) continue;
        const uint32_t at = base + st[k];
        if (at < cap) packed[at] = x | (y << 8) | ((z0 + k) << 16);
    }
}

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// the kernels move 16-byte words
int check_aligned(const void* p, const char* what) {
    if ((reinterpret_cast<uintptr_t>(p) & 15u) != 0) return fail(YX_ERR_INVALID_ARG, "%s must be 16-byte aligned", what);
    return YX_OK;
}

int check_bits(int B) {
    if (!pow2(B) || B < 128 || B > 1024)
        return fail(YX_ERR_INVALID_ARG, "bit volume side %d: a power of two in 128..1024", B);
    return YX_OK;
}

int check_vol(int V) {
    if (!pow2(V) || V < 64 || V > 512)
        return fail(YX_ERR_INVALID_ARG, "density volume side %d: a power of two in 64..512", V);
    return YX_OK;
}

int launch_density(const uint32_t* bits, int B, uint32_t* dens, hipStream_t st) {
    const int V = B / 2, S = B / 128;
    const size_t n = (size_t)S * V * V;
    hipLaunchKernelGGL(vt_density_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<const uint4*>(bits), B, S, reinterpret_cast<uint4*>(dens));
    VT_HIP(hipGetLastError());
    return YX_OK;
}

int launch_mark(const uint32_t* dens, int V, uint32_t* state, uint32_t* col_count, hipStream_t st) {
    const size_t cols = (size_t)(V / 4) * (V / 4);
    hipLaunchKernelGGL(vt_mark_kernel, dim3((unsigned)((cols + 3) / 4)), dim3(256), 0, st, dens, V,
                       reinterpret_cast<uint2*>(state), col_count);
    VT_HIP(hipGetLastError());
    return YX_OK;
}

int launch_pack(const uint32_t* state, const uint32_t* col_start, int V, uint32_t* packed, uint32_t cap,
                hipStream_t st) {
    const int G = V / 4;
    int lg = 0;
    while ((1 << lg) < G) ++lg;
    const size_t n = (size_t)G * G * G / 4;
    hipLaunchKernelGGL(vt_pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<const uint4*>(state), col_start, lg, packed, cap);
    VT_HIP(hipGetLastError());
    return YX_OK;
}

}  // namespace
}  // namespace yx

using namespace yx;

extern "C" int yx_voxtree_calc_density(const uint32_t* bits, int32_t bit_size, uint32_t* density, void* stream) {
    clear_error();
    if (!bits || !density) return fail(YX_ERR_INVALID_ARG, "null argument");
    int rc = check_bits(bit_size);
    if (rc == YX_OK) rc = check_aligned(bits, "bits");
    if (rc == YX_OK) rc = check_aligned(density, "density");
    if (rc != YX_OK) return rc;
    return launch_density(bits, bit_size, density, (hipStream_t)stream);
}

extern "C" int yx_voxtree_mark_bricks(const uint32_t* density, int32_t vol_size, uint32_t* brick_state,
                                      uint32_t* column_count, void* stream) {
    clear_error();
    if (!density || !brick_state || !column_count) return fail(YX_ERR_INVALID_ARG, "null argument");
    int rc = check_vol(vol_size);
    if (rc == YX_OK) rc = check_aligned(brick_state, "brick_state");
    if (rc != YX_OK) return rc;
    return launch_mark(density, vol_size, brick_state, column_count, (hipStream_t)stream);
}

extern "C" int yx_voxtree_pack_bricks(const uint32_t* brick_state, const uint32_t* column_start, int32_t vol_size,
                                      uint32_t* packed, uint32_t capacity, void* stream) {
    clear_error();
    if (!brick_state || !column_start || (!packed && capacity)) return fail(YX_ERR_INVALID_ARG, "null argument");
    int rc = check_vol(vol_size);
    if (rc == YX_OK) rc = check_aligned(brick_state, "brick_state");
    if (rc != YX_OK) return rc;
    return launch_pack(brick_state, column_start, vol_size, packed, capacity, (hipStream_t)stream);
}

extern "C" int yx_voxtree_classify(const uint32_t* bits, int32_t bit_size, uint32_t* density, uint32_t* brick_state,
                                   uint32_t* column_count, uint32_t* column_start, uint32_t* packed,
                                   uint32_t capacity, uint32_t* num_packed, void* stream) {
    clear_error();
    if (!bits || !density || !brick_state || !column_count || !column_start || !num_packed ||
        (!packed && capacity))
        return fail(YX_ERR_INVALID_ARG, "null argument");
    int rc = check_bits(bit_size);
    if (rc == YX_OK) rc = check_aligned(bits, "bits");
    if (rc == YX_OK) rc = check_aligned(density, "density");
    if (rc == YX_OK) rc = check_aligned(brick_state, "brick_state");
    if (rc == YX_OK) rc = check_aligned(column_count, "column_count");
    if (rc == YX_OK) rc = check_aligned(column_start, "column_start");
    if (rc != YX_OK) return rc;
    const hipStream_t st = (hipStream_t)stream;
    const int V = bit_size / 2, G = V / 4;
    const uint32_t cols = (uint32_t)G * (uint32_t)G;
    if ((rc = launch_density(bits, bit_size, density, st)) != YX_OK) return rc;
    if ((rc = launch_mark(density, V, brick_state, column_count, st)) != YX_OK) return rc;
    // column_start[cols] receives the grand total (cols = G^2 is a multiple of 4)
    hipLaunchKernelGGL(vt_scan_kernel, dim3(1), dim3(1024), 0, st, reinterpret_cast<const uint4*>(column_count),
                       cols / 4, reinterpret_cast<uint4*>(column_start), column_start + cols);
    VT_HIP(hipGetLastError());
    uint32_t total = 0;
    VT_HIP(hipMemcpyAsync(&total, column_start + cols, 4, hipMemcpyDeviceToHost, st));
    VT_HIP(hipStreamSynchronize(st));
    *num_packed = total;
    if (total > capacity) return fail(YX_ERR_RANGE, "%u mixed bricks, room for %u", total, capacity);
    return launch_pack(brick_state, column_start, V, packed, capacity, st);
}
