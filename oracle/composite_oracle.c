/*
 * composite_oracle.c - FP64 CPU reference of the differentiable front-to-back
 * compositing over brick voxels, and of its gradient.
 *
 * TEST INFRASTRUCTURE: the checker of the HIP compositing kernels; only tests/
 * and bench.py's cpu_baseline leg load it.
 *
 * Forward model (paths inside znah/yoxel-voxel):
 *   ray direction      RendererBase::InitRayDir, cell/renderer_base.h:50-61,
 *                      normalized(dir0 + du*x + dv*y), cell/ppu_renderer.cpp:56
 *   eps clamp          PrepareRay, reaction/trace/trace.cu:21,31-33 (1e-5, sign kept)
 *   unit-cube slabs    intersect, reaction/trace/trace.cu:66-80, with 1/dir;
 *                      t_enter clamped at 0, :104-106
 *   sample lattice     walkBrick, reaction/trace/trace.cu:44-48: t = k*dt,
 *                      ceil(t_enter/dt) <= k, t < t_exit; dt = one voxel
 *   compositing        accum += col * (1 - accum.w), reaction/trace/trace.cu:58;
 *                      recurrence of reaction/report/voxel.tex:59-87
 * The discrete part (which samples exist and which voxel each one reads) is
 * computed in fp32 exactly as the GPU computes it - it is the exact discrete
 * traversal both sides share - and without any empty-brick skipping; every
 * continuous quantity (softplus, alpha, T, colour, gradients) is computed in
 * double. No early exit.
 *
 * Pins (tests/test_composite_oracle.py): a 1-D ray against numpy.cumprod in
 * fp64, the closed-form sample count of axis rays, analytic gradient against
 * central finite differences in fp64, and a planted sign mistake
 * (ORACLE_PLANT == 6) that the FD pin must reject.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

void oracle_pixel_dir(const float* basis, int x, int y, float* out);

typedef struct {
    const int32_t* table;
    const float* params; /* 4 floats per voxel slot: raw, r, g, b */
    int n, b, g, lb;
    uint32_t num_slots;
    const float* views;
    int n_views, w, h;
} vol_t;

typedef struct {
    float o[3], d[3], r[3];
    int k, kend;
} cray_t;

static float clamp_eps(float x) {
    const float eps = 1e-5f;
    return fabsf(x) < eps ? copysignf(eps, x) : x;
}

static int ray_setup(const vol_t* v, uint32_t ray, cray_t* c) {
    int per_view = v->w * v->h;
    int view = (int)(ray / (uint32_t)per_view);
    int rem = (int)(ray % (uint32_t)per_view);
    int y = rem / v->w, x = rem % v->w;
    const float* vw = v->views + 12 * view;
    float d[3];
    oracle_pixel_dir(vw + 3, x, y, d);
    float lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        c->o[i] = vw[i];
        c->d[i] = clamp_eps(d[i]);
        c->r[i] = 1.0f / c->d[i];
        lo[i] = c->r[i] * (0.0f - c->o[i]);
        hi[i] = c->r[i] * (1.0f - c->o[i]);
    }
    float te = fmaxf(fminf(lo[0], hi[0]), fmaxf(fminf(lo[1], hi[1]), fminf(lo[2], hi[2])));
    float tx = fminf(fmaxf(lo[0], hi[0]), fminf(fmaxf(lo[1], hi[1]), fmaxf(lo[2], hi[2])));
    te = fmaxf(te, 0.0f);
    if (tx < 0.0f || te > tx) return 0;
    float nf = (float)v->n;
    c->k = (int)ceilf(te * nf);
    c->kend = (int)ceilf(tx * nf);
    return c->k < c->kend;
}

/* voxel slot index of sample k, or -1 when its brick is not allocated */
static int64_t sample_voxel(const vol_t* v, const cray_t* c, int k) {
    float nf = (float)v->n, inv_n = 1.0f / nf;
    float t = (float)k * inv_n;
    int vx[3];
    for (int i = 0; i < 3; ++i) {
        float p = c->o[i] + t * c->d[i];
        int q = (int)floorf(p * nf);
        vx[i] = q < 0 ? 0 : (q > v->n - 1 ? v->n - 1 : q);
    }
    int bx = vx[0] >> v->lb, by = vx[1] >> v->lb, bz = vx[2] >> v->lb;
    int32_t slot = v->table[((int64_t)bz * v->g + by) * v->g + bx];
    if (slot < 0 || (uint32_t)slot >= v->num_slots) return -1;
    int m = v->b - 1;
    int64_t local = (vx[0] & m) + ((int64_t)(vx[1] & m) << v->lb) + ((int64_t)(vx[2] & m) << (2 * v->lb));
    return ((int64_t)slot << (3 * v->lb)) + local;
}

static double softplus_d(double x) { return x > 30.0 ? x : log1p(exp(x)); }

/* parameter value in double, with an optional perturbation of one entry */
static double param(const vol_t* v, int64_t idx, int comp, int64_t p_idx, int p_comp, double delta) {
    double x = (double)v->params[4 * idx + comp];
    if (idx == p_idx && comp == p_comp) x += delta;
    return x;
}

static void forward_ray(const vol_t* v, const cray_t* c, int64_t p_idx, int p_comp, double delta, double* rgb,
                        double* trans, uint32_t* ns) {
    double C[3] = {0, 0, 0}, T = 1.0;
    uint32_t n = 0;
    for (int k = c->k; k < c->kend; ++k) {
        int64_t idx = sample_voxel(v, c, k);
        if (idx < 0) continue;
        double sig = softplus_d(param(v, idx, 0, p_idx, p_comp, delta));
        double e = exp(-sig), al = -expm1(-sig), w = T * al;
        for (int j = 0; j < 3; ++j) C[j] += w * param(v, idx, 1 + j, p_idx, p_comp, delta);
        T *= e;
        ++n;
    }
    for (int j = 0; j < 3; ++j) rgb[j] = C[j];
    *trans = T;
    *ns = n;
}

static vol_t mkvol(const int32_t* table, const float* params, int grid, int brick, uint32_t num_slots,
                   const float* views, int n_views, int w, int h) {
    vol_t v;
    v.table = table;
    v.params = params;
    v.n = grid;
    v.b = brick;
    v.g = grid / brick;
    v.lb = 0;
    while ((1 << v.lb) < brick) ++v.lb;
    v.num_slots = num_slots;
    v.views = views;
    v.n_views = n_views;
    v.w = w;
    v.h = h;
    return v;
}

/* Forward over rays [r0, r1): rgb (3 per ray), trans, nsamp (arrays indexed
 * by ray - r0). */
void comp_oracle_forward(const int32_t* table, const float* params, int grid, int brick, uint32_t num_slots,
                         const float* views, int n_views, int w, int h, uint32_t r0, uint32_t r1, double* rgb,
                         double* trans, uint32_t* nsamp) {
    vol_t v = mkvol(table, params, grid, brick, num_slots, views, n_views, w, h);
    for (uint32_t ray = r0; ray < r1; ++ray) {
        cray_t c;
        size_t o = ray - r0;
        if (!ray_setup(&v, ray, &c)) {
            rgb[3 * o] = rgb[3 * o + 1] = rgb[3 * o + 2] = 0.0;
            trans[o] = 1.0;
            nsamp[o] = 0;
            continue;
        }
        forward_ray(&v, &c, -1, 0, 0.0, rgb + 3 * o, trans + o, nsamp + o);
    }
}

/* Loss sum_rays |C - target|^2 over rays [r0, r1) and its gradient with
 * respect to the selected voxels (sel_map[voxel slot index] = row in grad or
 * -1; grad has 4 doubles per row: d raw, d r, d g, d b). Also lists, per
 * selected voxel, up to max_rays ray ids that sample it. */
void comp_oracle_grad(const int32_t* table, const float* params, int grid, int brick, uint32_t num_slots,
                      const float* views, int n_views, int w, int h, uint32_t r0, uint32_t r1,
                      const float* target, const int32_t* sel_map, double* grad, double* loss,
                      uint32_t* ray_lists, uint32_t* ray_counts, int max_rays) {
    vol_t v = mkvol(table, params, grid, brick, num_slots, views, n_views, w, h);
    double L = 0.0;
    for (uint32_t ray = r0; ray < r1; ++ray) {
        cray_t c;
        double C[3] = {0, 0, 0}, T;
        uint32_t ns = 0;
        int ok = ray_setup(&v, ray, &c);
        if (ok) forward_ray(&v, &c, -1, 0, 0.0, C, &T, &ns);
        double g[3];
        for (int j = 0; j < 3; ++j) {
            double diff = C[j] - (double)target[3 * (size_t)ray + j];
            L += diff * diff;
            g[j] = 2.0 * diff;
        }
        if (!ok) continue;
        double P[3] = {0, 0, 0};
        T = 1.0;
        for (int k = c.k; k < c.kend; ++k) {
            int64_t idx = sample_voxel(&v, &c, k);
            if (idx < 0) continue;
            double raw = (double)params[4 * idx];
            double sig = softplus_d(raw), e = exp(-sig), al = -expm1(-sig), wgt = T * al;
            double col[3] = {params[4 * idx + 1], params[4 * idx + 2], params[4 * idx + 3]};
            for (int j = 0; j < 3; ++j) P[j] += wgt * col[j];
            double tn = T * e, dsig = 0.0;
            for (int j = 0; j < 3; ++j) dsig += g[j] * (tn * col[j] - (C[j] - P[j]));
#if ORACLE_PLANT == 6
            dsig = -dsig; /* planted: adjoint sign flipped */
#endif
            int32_t row = sel_map[idx];
            if (row >= 0) {
                double sgm = 1.0 / (1.0 + exp(-raw));
                grad[4 * row] += dsig * sgm;
                for (int j = 0; j < 3; ++j) grad[4 * row + 1 + j] += wgt * g[j];
                if (ray_lists && ray_counts[row] < (uint32_t)max_rays) {
                    uint32_t m = ray_counts[row];
                    if (m == 0 || ray_lists[(size_t)row * max_rays + m - 1] != ray)
                        ray_lists[(size_t)row * max_rays + ray_counts[row]++] = ray;
                }
            }
            T = tn;
        }
    }
    *loss += L;
}

/* Loss over the listed rays with params[p_idx][p_comp] shifted by delta (in
 * double), for central finite differences. */
double comp_oracle_loss(const int32_t* table, const float* params, int grid, int brick, uint32_t num_slots,
                        const float* views, int n_views, int w, int h, const uint32_t* rays, int n_rays,
                        const float* target, int64_t p_idx, int p_comp, double delta) {
    vol_t v = mkvol(table, params, grid, brick, num_slots, views, n_views, w, h);
    double L = 0.0;
    for (int i = 0; i < n_rays; ++i) {
        cray_t c;
        double C[3] = {0, 0, 0}, T;
        uint32_t ns;
        if (ray_setup(&v, rays[i], &c)) forward_ray(&v, &c, p_idx, p_comp, delta, C, &T, &ns);
        for (int j = 0; j < 3; ++j) {
            double diff = C[j] - (double)target[3 * (size_t)rays[i] + j];
            L += diff * diff;
        }
    }
    return L;
}

/* The voxel slot indices sampled by one ray, in order (for the 1-D pins). */
int comp_oracle_ray_voxels(const int32_t* table, int grid, int brick, uint32_t num_slots, const float* views,
                           int n_views, int w, int h, uint32_t ray, int64_t* out, int max_out) {
    vol_t v = mkvol(table, NULL, grid, brick, num_slots, views, n_views, w, h);
    cray_t c;
    if (!ray_setup(&v, ray, &c)) return 0;
    int n = 0;
    for (int k = c.k; k < c.kend && n < max_out; ++k) {
        int64_t idx = sample_voxel(&v, &c, k);
        if (idx >= 0) out[n++] = idx;
    }
    return n;
}
