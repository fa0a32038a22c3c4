// Post passes of the shading pipeline (SVORenderer::Render,
// demo/SVORenderer.cpp:95-149): depth-limited blur of the z buffer (Run_BlurZ,
// five passes) and Phong shading with normals taken from the z buffer
// (Run_ShadeSimple). The reference's kernels are missing; the definitions are
// the ones stated in include/yoxel.h beside yx_render_pipeline.
//
// Both passes are image-space stencils bound by memory traffic and by the
// per-pixel arithmetic, not by MFMA: a blur pass reads each z once into an LDS
// tile with a 2-pixel apron and writes each z once; the shade pass reads z (with
// its 4 neighbours, served from L1/L2) and the colour, and writes RGBA8.
// Every float operation follows the stated definition in order, with no
// contraction, so results are bit-identical to a CPU evaluation of it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "yoxel.h"

namespace yx {

constexpr int kBlurK = 5;  // window side (BlurZKernSize)
constexpr int kBlurR = kBlurK / 2;
constexpr int kBlurPasses = 5;
constexpr int kBlurTX = 64, kBlurTY = 16;  // 256 threads, 4 output rows each
constexpr int kBlurRows = 4;
constexpr int kBlurThreads = kBlurTX * kBlurTY / kBlurRows;
constexpr int kBlurLX = kBlurTX + 2 * kBlurR, kBlurLY = kBlurTY + 2 * kBlurR;

struct BlurArgs {
    const float* zin;
    float* zout;
    int w, h;
    float zlimit;  // pixels at least this deep are copied
    float edge;    // window pixels whose depth differs by this much or more are left out
    float kern[kBlurK * kBlurK];
};

// A workgroup stages its 64x16 tile with a 2-pixel apron (68x20) in LDS; a
// thread keeps the 8x5 window column it shares between its 4 vertically
// adjacent outputs in registers. Each output sums its own 5x5 window in row
// order, as the definition does.
__global__ __launch_bounds__(kBlurThreads) void blur_z_kernel(BlurArgs a) {
    __shared__ float tile[kBlurLY * kBlurLX];
    const int x0 = (int)blockIdx.x * kBlurTX, y0 = (int)blockIdx.y * kBlurTY;
    // apron pixels outside the image read as +inf: they fail the depth test,
    // exactly as if they were left out of the window
    for (int i = (int)threadIdx.x; i < kBlurLY * kBlurLX; i += kBlurThreads) {
        const int ly = i / kBlurLX, lx = i - ly * kBlurLX;
        const int gx = x0 + lx - kBlurR, gy = y0 + ly - kBlurR;
        float v = __builtin_inff();
        if (gx >= 0 && gx < a.w && gy >= 0 && gy < a.h) v = a.zin[(size_t)gy * (size_t)a.w + (size_t)gx];
        tile[i] = v;
    }
    __syncthreads();
    const int tx = (int)threadIdx.x % kBlurTX, ty = ((int)threadIdx.x / kBlurTX) * kBlurRows;
    const int x = x0 + tx;
    if (x >= a.w) return;
    float win[kBlurRows + kBlurK - 1][kBlurK];
#pragma unroll
    for (int r = 0; r < kBlurRows + kBlurK - 1; ++r)
#pragma unroll
        for (int c = 0; c < kBlurK; ++c) win[r][c] = tile[(ty + r) * kBlurLX + tx + c];
#pragma unroll
    for (int j = 0; j < kBlurRows; ++j) {
        const int y = y0 + ty + j;
        if (y >= a.h) break;
        const float zp = win[j + kBlurR][kBlurR];
        float out = zp;
        if (zp < a.zlimit) {
            float num = 0.0f, den = 0.0f;
#pragma unroll
            for (int dy = 0; dy < kBlurK; ++dy)
#pragma unroll
                for (int dx = 0; dx < kBlurK; ++dx) {
                    const float zq = win[j + dy][dx];
                    if (__builtin_fabsf(zq - zp) < a.edge) {
                        const float wq = a.kern[dy * kBlurK + dx];
                        num += wq * zq;
                        den += wq;
                    }
                }
            out = num / den;
        }
        a.zout[(size_t)y * (size_t)a.w + (size_t)x] = out;
    }
}

constexpr int kShadeTX = 64, kShadeTY = 4;

struct ShadeArgs {
    const float* z;
    const uint32_t* albedo;
    uint32_t* rgba;
    float* zout;  // copy of z (the caller's z output), may be null
    int w, h;
    float basis[9];  // dir0, du, dv
    float eye[3];
    float frame[9];  // right, down, forward
    float pix;       // pixel size at unit depth
    float ambient;
    int show_normals;
    yx_light_params lights[YX_MAX_LIGHTS];
};

__device__ __forceinline__ float sh_dot(const float* a, const float* b) {
    float r = 0.0f;
    r += a[0] * b[0];
    r += a[1] * b[1];
    r += a[2] * b[2];
    return r;
}

__device__ __forceinline__ void sh_normalize(float* v) {
    const float n = __builtin_sqrtf(sh_dot(v, v));
    v[0] = v[0] / n;
    v[1] = v[1] / n;
    v[2] = v[2] / n;
}

// one-sided depth difference with the smaller magnitude (ztools absMin);
// neighbours outside the image or missed are left out
__device__ __forceinline__ float sh_slope(const float* z, float zc, bool has_fwd, size_t pf, bool has_back,
                                          size_t pb) {
    const float inf = __builtin_inff();
    const float zf = has_fwd ? z[pf] : inf;
    const float zb = has_back ? z[pb] : inf;
    const bool va = zf < inf, vb = zb < inf;
    const float da = zf - zc, db = zc - zb;
    if (va && vb) return (__builtin_fabsf(da) < __builtin_fabsf(db)) ? da : db;
    if (va) return da;
    if (vb) return db;
    return 0.0f;
}

__global__ __launch_bounds__(kShadeTX* kShadeTY) void shade_simple_kernel(ShadeArgs a) {
    const int x = (int)blockIdx.x * kShadeTX + (int)threadIdx.x % kShadeTX;
    const int y = (int)blockIdx.y * kShadeTY + (int)threadIdx.x / kShadeTX;
    if (x >= a.w || y >= a.h) return;
    const size_t p = (size_t)y * (size_t)a.w + (size_t)x;
    const float zc = a.z[p];
    if (a.zout) a.zout[p] = zc;
    if (!(zc < __builtin_inff())) {
        if (a.rgba) a.rgba[p] = 0u;
        return;
    }
    const float dzx = sh_slope(a.z, zc, x + 1 < a.w, p + 1, x > 0, p - 1);
    const float dzy = sh_slope(a.z, zc, y + 1 < a.h, p + (size_t)a.w, y > 0, p - (size_t)a.w);
    // view-space normal from z, pointing away from the eye (ztools.calcNormals)
    float n[3];
    n[0] = ((-a.pix) * dzx) * zc;
    n[1] = ((-a.pix) * dzy) * zc;
    n[2] = ((a.pix * a.pix) * zc) * zc;
    sh_normalize(n);
    float N[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) N[i] = -((n[0] * a.frame[i] + n[1] * a.frame[3 + i]) + n[2] * a.frame[6 + i]);
    sh_normalize(N);
    uint32_t res = 0xff000000u;
    if (a.show_normals) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float v = (N[c] * 0.5f + 0.5f) * 255.0f + 0.5f;
            res |= ((uint32_t)v & 0xffu) << (8 * c);
        }
        if (a.rgba) a.rgba[p] = res;
        return;
    }
    const float xf = (float)x, yf = (float)y;
    float q[3], P[3], V[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        q[i] = a.basis[i] + a.basis[3 + i] * xf;
        q[i] = q[i] + a.basis[6 + i] * yf;
        P[i] = a.eye[i] + zc * q[i];
        V[i] = q[i];
    }
    sh_normalize(V);
#pragma unroll
    for (int i = 0; i < 3; ++i) V[i] = -V[i];
    float dif[3] = {0.0f, 0.0f, 0.0f}, spe[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < YX_MAX_LIGHTS; ++m) {
        const yx_light_params& lt = a.lights[m];
        if (!lt.enabled) continue;
        float L[3] = {lt.pos[0] - P[0], lt.pos[1] - P[1], lt.pos[2] - P[2]};
        const float dist = __builtin_sqrtf(sh_dot(L, L));
        if (!(dist > 0.0f)) continue;
#pragma unroll
        for (int i = 0; i < 3; ++i) L[i] = L[i] / dist;
        const float ndl = sh_dot(N, L);
        if (!(ndl > 0.0f)) continue;
        const float att =
            1.0f / ((lt.attenuation[0] + lt.attenuation[1] * dist) + lt.attenuation[2] * (dist * dist));
#pragma unroll
        for (int c = 0; c < 3; ++c) dif[c] += (att * lt.diffuse[c]) * ndl;
        float R[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) R[i] = (2.0f * ndl) * N[i] - L[i];
        const float rv = sh_dot(R, V);
        if (rv > 0.0f) {
            const float r2 = rv * rv, r4 = r2 * r2, r8 = r4 * r4;
            const float s = r8 * r2;  // specular exponent 10
#pragma unroll
            for (int c = 0; c < 3; ++c) spe[c] += (att * lt.specular[c]) * s;
        }
    }
    const uint32_t al = a.albedo[p];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = (float)((al >> (8 * c)) & 0xffu) * (a.ambient + dif[c]) + 255.0f * spe[c];
        v = (255.0f < v) ? 255.0f : v;
        res |= ((uint32_t)(v + 0.5f) & 0xffu) << (8 * c);
    }
    if (a.rgba) a.rgba[p] = res;
}

}  // namespace yx
