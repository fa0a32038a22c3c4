/*
 * ssna_oracle.c - CPU restatement of the post passes of the reference's GPU
 * pipeline (SVORenderer::Render, demo/SVORenderer.cpp:95-149): five
 * depth-limited blurs of the z buffer (Run_BlurZ) and Phong shading with
 * normals taken from the z buffer (Run_ShadeSimple).
 *
 * TEST INFRASTRUCTURE. The parity checker of the HIP post kernels and a CPU
 * baseline; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it. The product library never links it.
 *
 * The CUDA kernels (cpp/trace.cu) and their header (BlurZKernSize, LightParams,
 * MaxLightsNum) are missing from the reference, so the passes are DEFINED HERE
 * from what survives:
 *   blur kernel table       SVORenderer::InitBlur, demo/SVORenderer.cpp:55-79
 *                           (size 5: DEFINED HERE)
 *   five passes, zlimit     demo/SVORenderer.cpp:131-145
 *   edge threshold          demo/dumps/ztools.py:4-15,87 findEdges(z, 4/2048)
 *   normals from z          demo/dumps/ztools.py:23-42 calcNormals (one-sided
 *                           differences, the smaller in magnitude)
 *   Phong terms             reaction/report/voxel.tex:97-98; specularExp 10,
 *                           ambient 0.1 (demo/SVORenderer.cpp:112-113)
 *   light parameters        demo/Demo.cpp:141-146 (pos, diffuse, specular,
 *                           attenuation coefficients)
 *   detailCoef / pixelAng   demo/SVORenderer.cpp:104-105
 *
 * Floating point: -ffp-contract=off, no fast-math; every operation is rounded
 * as written, in the order written.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

#define SSNA_K 5      /* BlurZKernSize, DEFINED HERE */
#define SSNA_PASSES 5 /* demo/SVORenderer.cpp:136 */

typedef struct {
    int32_t enabled;
    float pos[3];
    float diffuse[3];
    float specular[3];
    float attenuation[3]; /* constant, linear, quadratic */
} o_light;

static float s_min(float a, float b) { return (b < a) ? b : a; }

static float s_dot(const float* a, const float* b) {
    float r = 0.0f;
    r += a[0] * b[0];
    r += a[1] * b[1];
    r += a[2] * b[2];
    return r;
}

static void s_normalize(float* v) {
    float n = sqrtf(s_dot(v, v));
    v[0] = v[0] / n;
    v[1] = v[1] / n;
    v[2] = v[2] / n;
}

/* SVORenderer::InitBlur (demo/SVORenderer.cpp:55-79), float as written there */
void oracle_blur_kernel(float* kern) {
    float h = (float)(SSNA_K / 2);
    float scale = 2.0f;
    float sum = 0.0f;
    for (int y = 0; y < SSNA_K; ++y)
        for (int x = 0; x < SSNA_K; ++x) {
            float tx = scale * ((float)x - h) / h;
            float ty = scale * ((float)y - h) / h;
            tx = tx * tx;
            ty = ty * ty;
            float v = expf(-(tx + ty));
            kern[y * SSNA_K + x] = v;
            sum += v;
        }
    for (int i = 0; i < SSNA_K * SSNA_K; ++i) kern[i] /= sum;
}

/* Frame constants of one view (demo/SVORenderer.cpp:104-105,131-139):
 * out[0] detail = detailCoef * grad2rad(fov / 2) / width
 * out[1] pixel angle = grad2rad(fov) / width
 * out[2] pixel size at unit depth = 2 tan(grad2rad(fov / 2)) / width
 * out[3] voxel size = 2^-depth; out[4] edge threshold = 4 voxel sizes
 * out[5..9] zlimit of the five passes, blurSize 3, 6, 9, 12, 15 */
void oracle_pipeline_params(float fov, int w, int depth, float coef, float* out) {
    const float deg2rad = (float)(3.14159265358979323846 / (double)180.0f);
    out[0] = (coef * ((fov / 2.0f) * deg2rad)) / (float)w;
    out[1] = (fov * deg2rad) / (float)w;
    float da = tanf((fov / 2.0f) * deg2rad) / (float)w;
    out[2] = 2.0f * da;
    float vox = ldexpf(1.0f, -depth);
    out[3] = vox;
    out[4] = 4.0f * vox;
    float blur = 3.0f;
    for (int i = 0; i < SSNA_PASSES; ++i) {
        out[5 + i] = (5.0f * vox) / (out[1] * blur);
        blur += 3.0f;
    }
}

/* One BlurZ pass. A hit pixel nearer than zlimit becomes the kernel-weighted
 * mean of the pixels of its 5x5 window (inside the image) whose depth differs
 * from its own by less than `edge`, renormalised by their weights; every other
 * pixel (farther, or a miss: z = +inf) is copied. Rows y0..y1. */
void oracle_blur_z(int w, int h, const float* zin, float* zout, float zlimit, float edge,
                   const float* kern, int y0, int y1) {
    const int r = SSNA_K / 2;
    for (int y = y0; y < y1 && y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t p = (size_t)y * (size_t)w + (size_t)x;
            float zp = zin[p];
            if (!(zp < zlimit)) {
                zout[p] = zp;
                continue;
            }
            float num = 0.0f, den = 0.0f;
            for (int dy = -r; dy <= r; ++dy) {
                int yy = y + dy;
                if (yy < 0 || yy >= h) continue;
                for (int dx = -r; dx <= r; ++dx) {
                    int xx = x + dx;
                    if (xx < 0 || xx >= w) continue;
                    float zq = zin[(size_t)yy * (size_t)w + (size_t)xx];
                    if (fabsf(zq - zp) < edge) {
                        float wq = kern[(dy + r) * SSNA_K + (dx + r)];
                        num += wq * zq;
                        den += wq;
                    }
                }
            }
#if ORACLE_PLANT == 10
            zout[p] = num; /* planted: no renormalisation */
#else
            zout[p] = num / den;
#endif
        }
}

/* Camera frame: right, down (= -up'), forward (RendererBase::InitRayDir's
 * vectors, cell/renderer_base.h:50-61). */
void oracle_view_frame(const float* dir, const float* up, float* out) {
    float f[3] = {dir[0], dir[1], dir[2]};
    s_normalize(f);
    float r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    s_normalize(r);
    float u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    for (int i = 0; i < 3; ++i) {
        out[i] = r[i];
        out[3 + i] = -u[i];
        out[6 + i] = f[i];
    }
}

/* one-sided depth difference with the smaller magnitude (ztools.absMin) */
static float s_slope(const float* z, size_t p, float zc, int has_fwd, size_t pf, int has_back, size_t pb) {
    float a = 0.0f, b = 0.0f;
    int va = has_fwd && z[pf] < INFINITY;
    int vb = has_back && z[pb] < INFINITY;
    if (va) a = z[pf] - zc;
    if (vb) b = zc - z[pb];
    (void)p;
    if (va && vb) return (fabsf(a) < fabsf(b)) ? a : b;
    if (va) return a;
    if (vb) return b;
    return 0.0f;
}

/* Run_ShadeSimple over rows y0..y1. albedo RGBA8 (r low byte); z the (blurred)
 * view depth, +inf on a miss; basis = dir0, du, dv; frame = right, down,
 * forward; pix = pixel size at unit depth. Misses are (0, 0, 0, 0). */
void oracle_shade_simple(int w, int h, const float* z, const uint32_t* albedo, const float* basis,
                         const float* eye, const float* frame, float pix, const o_light* lights,
                         int n_lights, float ambient, int show_normals, uint32_t* out, int y0, int y1) {
    for (int y = y0; y < y1 && y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t p = (size_t)y * (size_t)w + (size_t)x;
            float zc = z[p];
            if (!(zc < INFINITY)) {
                out[p] = 0u;
                continue;
            }
            float dzx = s_slope(z, p, zc, x + 1 < w, p + 1, x > 0, p - 1);
            float dzy = s_slope(z, p, zc, y + 1 < h, p + (size_t)w, y > 0, p - (size_t)w);
            /* ztools.calcNormals: view-space normal, pointing away from the eye */
            float n[3];
            n[0] = ((-pix) * dzx) * zc;
            n[1] = ((-pix) * dzy) * zc;
            n[2] = ((pix * pix) * zc) * zc;
            s_normalize(n);
            /* world space, facing the eye */
            float N[3];
            for (int i = 0; i < 3; ++i) {
#if ORACLE_PLANT == 11
                N[i] = (n[0] * frame[i] + n[1] * frame[3 + i]) + n[2] * frame[6 + i]; /* planted: not flipped */
#else
                N[i] = -((n[0] * frame[i] + n[1] * frame[3 + i]) + n[2] * frame[6 + i]);
#endif
            }
            s_normalize(N);
            uint32_t res = 0xff000000u;
            if (show_normals) {
                for (int c = 0; c < 3; ++c) {
                    float v = (N[c] * 0.5f + 0.5f) * 255.0f + 0.5f;
                    res |= ((uint32_t)v & 0xffu) << (8 * c);
                }
                out[p] = res;
                continue;
            }
            float xf = (float)x, yf = (float)y;
            float q[3];
            for (int i = 0; i < 3; ++i) q[i] = basis[i] + basis[3 + i] * xf;
            for (int i = 0; i < 3; ++i) q[i] = q[i] + basis[6 + i] * yf;
            float P[3], V[3] = {q[0], q[1], q[2]};
            for (int i = 0; i < 3; ++i) P[i] = eye[i] + zc * q[i];
            s_normalize(V);
            for (int i = 0; i < 3; ++i) V[i] = -V[i];
            float dif[3] = {0.0f, 0.0f, 0.0f}, spe[3] = {0.0f, 0.0f, 0.0f};
            for (int m = 0; m < n_lights; ++m) {
                const o_light* lt = &lights[m];
                if (!lt->enabled) continue;
                float L[3] = {lt->pos[0] - P[0], lt->pos[1] - P[1], lt->pos[2] - P[2]};
                float dist = sqrtf(s_dot(L, L));
                if (!(dist > 0.0f)) continue;
                for (int i = 0; i < 3; ++i) L[i] = L[i] / dist;
                float ndl = s_dot(N, L);
                if (!(ndl > 0.0f)) continue;
                float att = 1.0f / ((lt->attenuation[0] + lt->attenuation[1] * dist) +
                                    lt->attenuation[2] * (dist * dist));
                for (int c = 0; c < 3; ++c) dif[c] += (att * lt->diffuse[c]) * ndl;
                float R[3];
                for (int i = 0; i < 3; ++i) R[i] = (2.0f * ndl) * N[i] - L[i];
                float rv = s_dot(R, V);
                if (rv > 0.0f) {
                    float r2 = rv * rv, r4 = r2 * r2, r8 = r4 * r4;
                    float s = r8 * r2; /* specularExp 10 */
                    for (int c = 0; c < 3; ++c) spe[c] += (att * lt->specular[c]) * s;
                }
            }
            uint32_t a = albedo[p];
            for (int c = 0; c < 3; ++c) {
                float v = (float)((a >> (8 * c)) & 0xffu) * (ambient + dif[c]) + 255.0f * spe[c];
                v = s_min(v, 255.0f);
                res |= ((uint32_t)(v + 0.5f) & 0xffu) << (8 * c);
            }
            out[p] = res;
        }
}
