/*
 * svo_oracle.c - CPU restatement of the reference's sparse-voxel-octree ray cast.
 *
 * TEST INFRASTRUCTURE. This file is the parity checker for the HIP kernels and
 * the CPU baseline of the benchmark. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it; the product library never links it.
 *
 * It follows, line by line (paths inside znah/yoxel-voxel):
 *   RendererBase::InitRayDir      cell/renderer_base.h:50-61
 *   PPURendererBase::RecTrace     cell/ppu_renderer.cpp:18-41
 *   PPURendererBase::RenderRect   cell/ppu_renderer.cpp:43-70
 *   FindFirstChild / GoNext       cell/spu/trace_spu.cpp:48-93 (SPE twins; the
 *                                 shared trace_utils.h is absent from the tree)
 *   SetupTrace / dirFlags         cell/spu/trace_spu.c_:66-105
 *   AdjustDir                     ntree/ntree_trace.cu:38-50 (adjustCoord, eps 1e-6)
 *   VoxNode / EmptyNode           reaction/report/main.tex:36-56
 *   cg::normalized, operator ^    nest/include/geometry/primitives/point.h:418-425,
 *                                 :435-442, :462-500
 *   cg::grad2rad (float)          nest/include/geometry/xmath.h:57
 * Pieces the reference does not define (its shader and helper headers are
 * missing) are this project's definitions and are marked DEFINED HERE:
 * IsNull for every id with the high bit set, the Lambert shade, the entry-face
 * normal (tie order of bittree/trace.cu:100-106 argmax), shadow rays.
 *
 * Floating point: compile with -O2 -ffp-contract=off and without fast-math so
 * every product and sum is rounded exactly as written.
 *
 * Pins (tests/test_oracle_pins.py): brute-force first-hit search on dense grids
 * of depth 1..5, closed-form rays, hand-derived tie cases, invariants, and
 * planted-mistake builds (ORACLE_PLANT) that the pins must reject.
 * Shading pins (tests/test_oracle_shade_pins.py): axis rays and each of the six
 * faces against the closed-form Lambert term, the entry face against a float64
 * brute force of the hit box, and shadow-ray occlusion against an exhaustive
 * float64 search from the stepped origin; plants 12-14 (light sign flipped,
 * origin stepped into the hit voxel, ray cast away from the light) fail them.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

#define O_EMPTY_NODE 0x80000000u
#define O_MISS 0xFFFFFFFFu

typedef struct {
    float x, y, z;
} f3;

/* ---- cg math (nest/include/geometry) ------------------------------------ */

static f3 f3_make(float x, float y, float z) {
    f3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}

/* point * point scalar product: res = 0; res += a[n]*b[n] (point.h:418-425) */
static float f3_dot(f3 a, f3 b) {
    float res = 0.0f;
    res += a.x * b.x;
    res += a.y * b.y;
    res += a.z * b.z;
    return res;
}

/* normalized(): p /= sqrt(p*p), component-wise division (point.h:462-500) */
static f3 f3_normalized(f3 p) {
    float n = sqrtf(f3_dot(p, p));
    return f3_make(p.x / n, p.y / n, p.z / n);
}

/* operator ^ (point.h:435-442) */
static f3 f3_cross(f3 a, f3 b) {
    return f3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* std::max / std::min argument order, as fmaxf(x, fmaxf(y, z)) in
 * ntree/ntree_trace.cu:56-57 maxCoord/minCoord */
static float o_max(float a, float b) { return (a < b) ? b : a; }
static float o_min(float a, float b) { return (b < a) ? b : a; }
static float max_coord(const float* p) { return o_max(p[0], o_max(p[1], p[2])); }
static float min_coord(const float* p) { return o_min(p[0], o_min(p[1], p[2])); }

/* ---- InitRayDir (cell/renderer_base.h:50-61) ----------------------------- */
/* out = dir0[3], du[3], dv[3]. Integer factors convert to float first, as the
 * cg point operators do (MAX_TYPE(int, float) == float). */
void oracle_init_ray_dir(const float* dir, const float* up, float fov, int w, int h,
                         float* out) {
    f3 vfwd = f3_normalized(f3_make(dir[0], dir[1], dir[2]));
    f3 vright = f3_normalized(f3_cross(vfwd, f3_make(up[0], up[1], up[2])));
    f3 vup = f3_cross(vright, vfwd);
    const float deg2rad = (float)(3.14159265358979323846 / (double)180.0f);
    float da = tanf((fov / 2.0f) * deg2rad) / (float)w;
    f3 du = f3_make((vright.x * 2.0f) * da, (vright.y * 2.0f) * da, (vright.z * 2.0f) * da);
    f3 dv = f3_make((vup.x * -2.0f) * da, (vup.y * -2.0f) * da, (vup.z * -2.0f) * da);
    float wf = (float)w, hf = (float)h;
    f3 d0 = f3_make((vfwd.x - (du.x * wf) / 2.0f) - (dv.x * hf) / 2.0f,
                    (vfwd.y - (du.y * wf) / 2.0f) - (dv.y * hf) / 2.0f,
                    (vfwd.z - (du.z * wf) / 2.0f) - (dv.z * hf) / 2.0f);
    out[0] = d0.x; out[1] = d0.y; out[2] = d0.z;
    out[3] = du.x; out[4] = du.y; out[5] = du.z;
    out[6] = dv.x; out[7] = dv.y; out[8] = dv.z;
}

/* normalized(dir0 + du*x + dv*y) (cell/ppu_renderer.cpp:56) */
static f3 pixel_dir(const float* basis, int x, int y) {
    float xf = (float)x, yf = (float)y;
    f3 d = f3_make(basis[0] + basis[3] * xf, basis[1] + basis[4] * xf, basis[2] + basis[5] * xf);
    d = f3_make(d.x + basis[6] * yf, d.y + basis[7] * yf, d.z + basis[8] * yf);
    return f3_normalized(d);
}

/* AdjustDir: keep 1/dir finite (adjustCoord, ntree/ntree_trace.cu:38-50) */
static float adjust_coord(float x) {
    const float eps = 1e-6f;
    if (fabsf(x) < eps) x = (x < 0) ? -eps : eps;
    return x;
}

/* SetupTrace: slab entry/exit of the unit cube per axis and the octant mirror
 * mask (cell/spu/trace_spu.c_:79-99). */
static int setup_trace(const float* p, const float* d, float* t1, float* t2, unsigned* dir_flags) {
    unsigned flags = 0;
    for (int i = 0; i < 3; ++i) {
#if ORACLE_PLANT == 5
        /* planted: both branches use the positive-direction assignment */
        t1[i] = (-p[i]) / d[i];
        t2[i] = (1.0f - p[i]) / d[i];
        if (d[i] < 0) { float tt = t1[i]; t1[i] = t2[i]; t2[i] = tt; }
        if (d[i] < 0) flags |= 1u << ((i + 1) % 3);
#else
        if (d[i] > 0) {
            t1[i] = (-p[i]) / d[i];
            t2[i] = (1.0f - p[i]) / d[i];
        } else {
            t2[i] = (-p[i]) / d[i];
            t1[i] = (1.0f - p[i]) / d[i];
        }
        if (d[i] < 0) flags |= 1u << i;
#endif
    }
    *dir_flags = flags;
    return max_coord(t1) < min_coord(t2);
}

/* FindFirstChild (cell/spu/trace_spu.cpp:48-68) */
static int find_first_child(float* t1, float* t2) {
    float tm[3];
    for (int i = 0; i < 3; ++i) {
#if ORACLE_PLANT == 4
        tm[i] = t1[i] + 0.5f * t2[i]; /* planted: wrong midpoint */
#else
        tm[i] = 0.5f * (t1[i] + t2[i]);
#endif
    }
    float t_enter = max_coord(t1);
    int child = 0;
    for (int i = 0; i < 3; ++i) {
#if ORACLE_PLANT == 2
        int cmp = t_enter >= tm[i]; /* planted: non-strict comparison */
#else
        int cmp = t_enter > tm[i];
#endif
        if (cmp) {
            t1[i] = tm[i];
            child |= 1 << i;
        } else {
            t2[i] = tm[i];
        }
    }
    return child;
}

/* GoNext (cell/spu/trace_spu.cpp:70-93): argmin with the twin's tie order */
static int go_next(int* child, float* t1, float* t2) {
    int exit_plane;
#if ORACLE_PLANT == 1
    /* planted: x/y tie order reversed */
    if (t2[0] >= t2[1])
        exit_plane = (t2[1] < t2[2]) ? 1 : 2;
    else
        exit_plane = (t2[0] < t2[2]) ? 0 : 2;
#else
    if (t2[0] > t2[1])
        exit_plane = (t2[1] < t2[2]) ? 1 : 2;
    else
        exit_plane = (t2[0] < t2[2]) ? 0 : 2;
#endif
    int mask = 1 << exit_plane;
    if ((*child & mask) != 0) return 0;
    *child ^= mask;
    float dt = t2[exit_plane] - t1[exit_plane];
    t1[exit_plane] = t2[exit_plane];
    t2[exit_plane] += dt;
    return 1;
}

/* ---- RecTrace (cell/ppu_renderer.cpp:18-41) ------------------------------ */

typedef struct {
    const uint32_t* nodes; /* 10 u32 per VoxNode */
    uint32_t num_nodes;
    unsigned dir_flags;
    int occlusion;   /* DEFINED HERE: shadow rays skip leaf cells behind the origin */
    float detail;    /* LOD stop coefficient (RenderParams::detailCoef); 0 = exact */
    uint32_t visits; /* node records read: RecTrace calls past the null/behind test */
    int bad_id;      /* a child id >= num_nodes was followed */
} trace_ctx;

typedef struct {
    uint32_t node;
    int child; /* -1: the ray stopped on the node itself (LOD) */
    float t;
    int level; /* level of the hit cell: 1 for a child of the root */
    int vx, vy, vz;
    float t1[3]; /* entry times of the hit cell (entry-face normal) */
} trace_hit;

/* IsNull: every id with the high bit set (EmptyNode, FullNode), DEFINED HERE
 * from the special-value rule of reaction/report/main.tex:52-55. */
static int is_null(uint32_t id) { return (id & 0x80000000u) != 0; }
static int leaf_flag(uint32_t flags, int i) { return (int)((flags >> i) & 1u); }

/* a leaf or a non-null child id in some slot */
static int has_child(const uint32_t* node) {
    for (int i = 0; i < 8; ++i)
        if (leaf_flag(node[0], i) || !is_null(node[2 + i])) return 1;
    return 0;
}

static int rec_trace(trace_ctx* c, uint32_t node_id, const float* t1_in, const float* t2_in,
                     int level, int px, int py, int pz, trace_hit* res) {
    float t1[3] = {t1_in[0], t1_in[1], t1_in[2]};
    float t2[3] = {t2_in[0], t2_in[1], t2_in[2]};
    if (is_null(node_id) || min_coord(t2) <= 0) return 0;
    if (node_id >= c->num_nodes) {
        c->bad_id = 1;
        return 0;
    }
    c->visits++;
    const uint32_t* node = c->nodes + (size_t)10 * node_id;
    /* LOD stop, DEFINED HERE after RenderParams::detailCoef (demo/SVORenderer.cpp:
     * 104) and RayData::endNodeChild < 0 (:171-179): a node below the root whose
     * cell edge 2^-level is smaller than its entry distance times `detail`, and
     * which has a non-null child, is the hit itself (its colour is node data). */
#if ORACLE_PLANT == 9
    float edge = ldexpf(1.0f, -(level + 1)); /* planted: half the cell edge */
#else
    float edge = ldexpf(1.0f, -level);
#endif
    if (level > 0 && has_child(node) && edge < max_coord(t1) * c->detail) {
        res->node = node_id;
        res->child = -1;
        res->t = max_coord(t1);
        res->level = level;
        res->vx = px;
        res->vy = py;
        res->vz = pz;
        memcpy(res->t1, t1, sizeof(t1));
        return 1;
    }
    int ch = find_first_child(t1, t2);
    for (;;) {
#if ORACLE_PLANT == 3
        int real = ch; /* planted: octant mirror dropped */
#else
        int real = ch ^ (int)c->dir_flags;
#endif
        int cx = 2 * px + (real & 1), cy = 2 * py + ((real >> 1) & 1), cz = 2 * pz + ((real >> 2) & 1);
        if (leaf_flag(node[0], real)) {
            if (!c->occlusion || min_coord(t2) > 0) {
                res->node = node_id;
                res->child = real;
                res->t = max_coord(t1);
                res->level = level + 1;
                res->vx = cx;
                res->vy = cy;
                res->vz = cz;
                memcpy(res->t1, t1, sizeof(t1));
                return 1;
            }
        } else if (rec_trace(c, node[2 + real], t1, t2, level + 1, cx, cy, cz, res)) {
            return 1;
        }
        if (!go_next(&ch, t1, t2)) return 0;
    }
}

/* Entry-face axis: argmax of the hit cell's entry times with the tie order of
 * bittree/trace.cu:100-106. DEFINED HERE. */
static int entry_axis(const float* t1) {
    if (t1[0] < t1[1])
        return (t1[1] > t1[2]) ? 1 : 2;
    else
        return (t1[0] > t1[2]) ? 0 : 2;
}

/* Trace one ray from `pos` along the (already normalized) `dir`. Returns 1 on a
 * hit. mode: 0 primary (reference RecTrace), 1 occlusion (shadow rays). */
static int trace_ray(const uint32_t* nodes, uint32_t num_nodes, uint32_t root, const float* pos,
                     const float* dir_in, int occlusion, float detail, trace_hit* hit,
                     uint32_t* visits, int* bad_id) {
    float d[3] = {adjust_coord(dir_in[0]), adjust_coord(dir_in[1]), adjust_coord(dir_in[2])};
    float t1[3], t2[3];
    unsigned flags;
    if (!setup_trace(pos, d, t1, t2, &flags)) return 0;
    trace_ctx c;
    c.nodes = nodes;
    c.num_nodes = num_nodes;
    c.dir_flags = flags;
    c.occlusion = occlusion;
    c.detail = detail;
    c.visits = 0;
    c.bad_id = 0;
    int r = rec_trace(&c, root, t1, t2, 0, 0, 0, 0, hit);
    *visits += c.visits;
    if (c.bad_id) *bad_id = 1;
    return r;
}

/* ---- shading: Lambert over RGBA8 leaves, DEFINED HERE --------------------- */
/* light_mode 0: light at the eye (RendererBase::SetViewPos puts the shader's
 * light at the viewer, cell/renderer_base.h:30-35); 1: directional light;
 * 2: directional light with one shadow ray per lit hit. */
typedef struct {
    uint32_t hit_id;
    int32_t vox[3];
    float t;
    uint32_t rgba;
    uint8_t shadow;
    uint32_t visits;
} pixel_out;

/* The reference's VoxData as colour16 + normal16 (reaction/report/main.tex:
 * 38,65; the unpacking bodies of demo/SVORenderer.cpp:181-185 are missing).
 * DEFINED HERE: bits 0-15 RGB565 (red in 11-15), widened by bit replication;
 * bits 16-31 a normal after Baptista, "Higher Accuracy Quantized Normals":
 * bits 15/14/13 the signs of x/y/z, bits 7-12 u and 0-6 v on the triangle
 * u + v <= 126, codes with u + v > 126 folded from (127 - u, 127 - v); the
 * normal is (u, v, 126 - u - v) with the signs, normalised. */
static uint32_t rgba_from_565(uint32_t c) {
    uint32_t r5 = (c >> 11) & 31u, g6 = (c >> 5) & 63u, b5 = c & 31u;
    uint32_t r = (r5 << 3) | (r5 >> 2), g = (g6 << 2) | (g6 >> 4), b = (b5 << 3) | (b5 >> 2);
    return r | (g << 8) | (b << 16) | 0xff000000u;
}

static f3 normal_from_16(uint32_t q) {
    int u = (int)((q >> 7) & 63u), v = (int)(q & 127u);
    if (u + v > 126) {
        u = 127 - u;
        v = 127 - v;
    }
#if ORACLE_PLANT == 15
    f3 p = f3_make((float)v, (float)u, (float)(126 - u - v)); /* planted: u and v swapped */
#else
    f3 p = f3_make((float)u, (float)v, (float)(126 - u - v));
#endif
    float len = sqrtf(f3_dot(p, p));
    f3 n = f3_make(p.x / len, p.y / len, p.z / len);
    if (q & 0x8000u) n.x = -n.x;
    if (q & 0x4000u) n.y = -n.y;
    if (q & 0x2000u) n.z = -n.z;
    return n;
}

static void shade_pixel(const uint32_t* nodes, uint32_t num_nodes, uint32_t root, const float* pos,
                        f3 dirv, int light_mode, const float* light, int payload, pixel_out* o, int* bad_id) {
    float dir[3] = {dirv.x, dirv.y, dirv.z};
    o->hit_id = O_MISS;
    o->vox[0] = o->vox[1] = o->vox[2] = -1;
    o->t = -1.0f;
    o->rgba = 0; /* Color32(0, 0, 0, 0), cell/ppu_renderer.cpp:54 */
    o->shadow = 2;
    o->visits = 0;
    /* RenderRect applies AdjustDir to dir itself (cell/ppu_renderer.cpp:57) */
    dir[0] = adjust_coord(dir[0]);
    dir[1] = adjust_coord(dir[1]);
    dir[2] = adjust_coord(dir[2]);
    trace_hit h;
    if (!trace_ray(nodes, num_nodes, root, pos, dir, 0, 0.0f, &h, &o->visits, bad_id)) return;
    o->hit_id = h.node * 8u + (uint32_t)h.child;
    o->vox[0] = h.vx;
    o->vox[1] = h.vy;
    o->vox[2] = h.vz;
    o->t = h.t;
    int e = entry_axis(h.t1);
    float lam;
    if (light_mode == 0) {
        lam = fabsf(dir[e]);
    } else {
        /* n = outward normal of the entry face: +e_e when the ray moves toward
         * -e (dir[e] < 0), else -e_e; n.L = +-light[e] */
#if ORACLE_PLANT == 12
        lam = (dir[e] < 0) ? -light[e] : light[e]; /* planted: light sign flipped */
#else
        lam = (dir[e] < 0) ? light[e] : -light[e];
#endif
        lam = o_max(lam, 0.0f);
    }
    uint32_t leaf = nodes[(size_t)10 * h.node + 2 + (uint32_t)h.child];
    uint32_t color = leaf;
    if (payload == 1) { /* the leaf's own colour and normal */
        f3 n = normal_from_16(leaf >> 16);
        f3 L = light_mode == 0 ? f3_make(-dir[0], -dir[1], -dir[2]) : f3_make(light[0], light[1], light[2]);
        lam = o_max(f3_dot(n, L), 0.0f);
        color = rgba_from_565(leaf & 0xffffu);
    }
    int lit = 1;
    if (light_mode == 2 && lam > 0.0f) {
        /* shadow ray: from the hit point, stepped half a cell back across the
         * entry face, toward the light; occlusion traversal */
        float p[3];
        for (int i = 0; i < 3; ++i) p[i] = pos[i] + dir[i] * h.t;
        float half = ldexpf(1.0f, -(h.level + 1));
#if ORACLE_PLANT == 13
        if (dir[e] < 0) /* planted: origin stepped into the hit voxel */
            p[e] = p[e] - half;
        else
            p[e] = p[e] + half;
#else
        if (dir[e] < 0)
            p[e] = p[e] + half;
        else
            p[e] = p[e] - half;
#endif
        trace_hit sh;
#if ORACLE_PLANT == 14
        const float away[3] = {-light[0], -light[1], -light[2]}; /* planted: ray cast away from the light */
        lit = !trace_ray(nodes, num_nodes, root, p, away, 1, 0.0f, &sh, &o->visits, bad_id);
#else
        lit = !trace_ray(nodes, num_nodes, root, p, light, 1, 0.0f, &sh, &o->visits, bad_id);
#endif
        o->shadow = lit ? 0 : 1;
    }
    float k = lit ? 0.1f + 0.9f * lam : 0.1f;
    uint32_t out = 0;
    for (int ch = 0; ch < 3; ++ch) {
        float v = (float)((color >> (8 * ch)) & 0xffu) * k + 0.5f;
        out |= ((uint32_t)v & 0xffu) << (8 * ch);
    }
    out |= 0xffu << 24;
    o->rgba = out;
}

/* ---- RenderRect (cell/ppu_renderer.cpp:43-70) over rows [y0, y1) --------- */
/* Outputs are full-frame arrays (row-major); any may be NULL. Returns the number
 * of rays traced (primary + shadow), or -1 if a child id was out of range. */
long long oracle_render_rows_fmt(const uint32_t* nodes, uint32_t num_nodes, uint32_t root,
                                 const float* pos, const float* basis, int w, int h, int y0, int y1,
                                 int light_mode, const float* light, int payload, uint32_t* hit_id,
                                 int32_t* voxel, float* t, uint32_t* rgba, uint8_t* shadow,
                                 uint32_t* visits) {
    long long rays = 0;
    int bad = 0;
    if (y1 > h) y1 = h;
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t offs = (size_t)y * (size_t)w + (size_t)x;
            pixel_out o;
            shade_pixel(nodes, num_nodes, root, pos, pixel_dir(basis, x, y), light_mode, light, payload, &o,
                        &bad);
            rays += 1 + (o.shadow < 2 ? 1 : 0);
            if (hit_id) hit_id[offs] = o.hit_id;
            if (voxel) {
                voxel[3 * offs + 0] = o.vox[0];
                voxel[3 * offs + 1] = o.vox[1];
                voxel[3 * offs + 2] = o.vox[2];
            }
            if (t) t[offs] = o.t;
            if (rgba) rgba[offs] = o.rgba;
            if (shadow) shadow[offs] = o.shadow;
            if (visits) visits[offs] = o.visits;
        }
    }
    return bad ? -1 : rays;
}

/* RGBA8 payloads (the default format). */
long long oracle_render_rows(const uint32_t* nodes, uint32_t num_nodes, uint32_t root,
                             const float* pos, const float* basis, int w, int h, int y0, int y1,
                             int light_mode, const float* light, uint32_t* hit_id, int32_t* voxel,
                             float* t, uint32_t* rgba, uint8_t* shadow, uint32_t* visits) {
    return oracle_render_rows_fmt(nodes, num_nodes, root, pos, basis, w, h, y0, y1, light_mode, light, 0,
                                  hit_id, voxel, t, rgba, shadow, visits);
}

/* One colour16 + normal16 payload decoded: RGBA8 and the unit normal. */
void oracle_decode_c16n16(uint32_t leaf, uint32_t* rgba, float* normal) {
    f3 n = normal_from_16(leaf >> 16);
    *rgba = rgba_from_565(leaf & 0xffffu);
    normal[0] = n.x;
    normal[1] = n.y;
    normal[2] = n.z;
}

/* Single rays with explicit origins and directions (for the pins). Directions
 * are used as given (AdjustDir is applied). out_hit[i] = node*8+child or miss,
 * out_vox 3 per ray, out_t, out_level, out_t1 3 per ray. Returns -1 on a bad id. */
int oracle_trace_rays(const uint32_t* nodes, uint32_t num_nodes, uint32_t root, int n,
                      const float* pos, const float* dir, int occlusion, uint32_t* out_hit,
                      int32_t* out_vox, float* out_t, int32_t* out_level, float* out_t1,
                      uint32_t* out_visits) {
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        trace_hit hh;
        uint32_t v = 0;
        int r = trace_ray(nodes, num_nodes, root, pos + 3 * i, dir + 3 * i, occlusion, 0.0f, &hh, &v, &bad);
        if (out_visits) out_visits[i] = v;
        if (r) {
            out_hit[i] = hh.node * 8u + (uint32_t)hh.child;
            out_vox[3 * i] = hh.vx;
            out_vox[3 * i + 1] = hh.vy;
            out_vox[3 * i + 2] = hh.vz;
            out_t[i] = hh.t;
            out_level[i] = hh.level;
            for (int k = 0; k < 3; ++k) out_t1[3 * i + k] = hh.t1[k];
        } else {
            out_hit[i] = O_MISS;
            out_vox[3 * i] = out_vox[3 * i + 1] = out_vox[3 * i + 2] = -1;
            out_t[i] = -1.0f;
            out_level[i] = 0;
            for (int k = 0; k < 3; ++k) out_t1[3 * i + k] = 0.0f;
        }
    }
    return bad ? -1 : 0;
}

/* Shaded single rays with explicit origins and directions (for the pins on
 * the Lambert term and the shadow rays): exactly the per-pixel work of
 * RenderRect (shade_pixel) with the pixel direction replaced by dir[i].
 * Outputs per ray: hit id (node*8+child or miss), RGBA8, shadow flag (0 lit,
 * 1 occluded, 2 no shadow ray), t, hit level and node visits (primary +
 * shadow). Returns -1 on a bad child id. */
int oracle_shade_rays(const uint32_t* nodes, uint32_t num_nodes, uint32_t root, int n, const float* pos,
                      const float* dir, int light_mode, const float* light, int payload, uint32_t* out_hit,
                      uint32_t* out_rgba, uint8_t* out_shadow, float* out_t, uint32_t* out_visits) {
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        pixel_out o;
        shade_pixel(nodes, num_nodes, root, pos + 3 * i,
                    f3_make(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]), light_mode, light, payload, &o, &bad);
        out_hit[i] = o.hit_id;
        out_rgba[i] = o.rgba;
        out_shadow[i] = o.shadow;
        out_t[i] = o.t;
        if (out_visits) out_visits[i] = o.visits;
    }
    return bad ? -1 : 0;
}

/* Pixel directions of a camera basis (for pins on InitRayDir). */
void oracle_pixel_dirs(const float* basis, int w, int h, float* out) {
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            f3 d = pixel_dir(basis, x, y);
            size_t o = 3 * ((size_t)y * (size_t)w + (size_t)x);
            out[o] = d.x;
            out[o + 1] = d.y;
            out[o + 2] = d.z;
        }
}

int oracle_plant(void) { return ORACLE_PLANT; }

/* One pixel direction (shared with the compositing restatement). */
void oracle_pixel_dir(const float* basis, int x, int y, float* out) {
    f3 d = pixel_dir(basis, x, y);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

/* ---- LOD colours: VoxNode::data as the subtree mean ------------------------ */
/* reaction/report/main.tex:61: a node's data is the arithmetic mean of the
 * colours of every voxel in its subtree. DEFINED HERE for RGBA8 leaves: per
 * channel (sum + count/2) / count in integers, 0 for a subtree without leaves.
 * out[id] for every node reachable from the root (others untouched). Returns
 * -1 on an id out of range or a node reached twice, else 0. */
int oracle_lod_colors(const uint32_t* nodes, uint32_t num_nodes, uint32_t root, uint32_t* out,
                      uint64_t* scratch /* 5 per node */, uint32_t* stack /* 2 per node */) {
    if (is_null(root)) return 0;
    if (root >= num_nodes) return -1;
    for (uint32_t i = 0; i < num_nodes; ++i) scratch[5 * (size_t)i + 4] = UINT64_MAX; /* unvisited */
    /* iterative post order: stack entries (id, next slot) */
    size_t sp = 0;
    stack[0] = root;
    stack[1] = 0;
    scratch[5 * (size_t)root + 4] = 0;
    for (int k = 0; k < 4; ++k) scratch[5 * (size_t)root + k] = 0;
    sp = 1;
    while (sp > 0) {
        uint32_t id = stack[2 * (sp - 1)];
        uint32_t slot = stack[2 * (sp - 1) + 1];
        const uint32_t* n = nodes + (size_t)10 * id;
        uint64_t* acc = scratch + 5 * (size_t)id;
        if (slot < 8) {
            stack[2 * (sp - 1) + 1] = slot + 1;
            uint32_t ch = n[2 + slot];
            if (leaf_flag(n[0], (int)slot)) {
                for (int k = 0; k < 4; ++k) acc[k] += (ch >> (8 * k)) & 0xffu;
                acc[4] += 1;
            } else if (!is_null(ch)) {
                if (ch >= num_nodes) return -1;
                if (scratch[5 * (size_t)ch + 4] != UINT64_MAX) return -1;
                for (int k = 0; k < 5; ++k) scratch[5 * (size_t)ch + k] = 0;
                stack[2 * sp] = ch;
                stack[2 * sp + 1] = 0;
                ++sp;
            }
            continue;
        }
        /* done with id: its mean, then add its sums into the parent */
        uint32_t mean = 0;
        if (acc[4] > 0)
            for (int k = 0; k < 4; ++k) mean |= (uint32_t)((acc[k] + acc[4] / 2) / acc[4]) << (8 * k);
        out[id] = mean;
        --sp;
        if (sp > 0) {
            uint64_t* pa = scratch + 5 * (size_t)stack[2 * (sp - 1)];
            for (int k = 0; k < 5; ++k) pa[k] += acc[k];
        }
    }
    return 0;
}

/* ---- LOD trace of the GPU pipeline (Run_Trace, demo/SVORenderer.cpp:125) ---- */
/* Primary rays only. Per pixel: t = max(t1) of the hit cell (-1 on a miss);
 * z = t / |dir0 + du x + dv y|, the view-space depth (+inf on a miss); albedo =
 * the leaf's RGBA8, or the node's LOD colour on an LOD stop (0 on a miss);
 * level of the hit cell (0 on a miss); visits. Returns the number of rays or -1
 * on a bad child id. */
long long oracle_lod_render_rows(const uint32_t* nodes, uint32_t num_nodes, uint32_t root,
                                 const uint32_t* lod_colors, const float* pos, const float* basis,
                                 int w, int h, int y0, int y1, float detail, float* t, float* z,
                                 uint32_t* albedo, uint8_t* level, uint32_t* visits) {
    long long rays = 0;
    int bad = 0;
    if (y1 > h) y1 = h;
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t offs = (size_t)y * (size_t)w + (size_t)x;
            float xf = (float)x, yf = (float)y;
            f3 q = f3_make(basis[0] + basis[3] * xf, basis[1] + basis[4] * xf, basis[2] + basis[5] * xf);
            q = f3_make(q.x + basis[6] * yf, q.y + basis[7] * yf, q.z + basis[8] * yf);
            float qn = sqrtf(f3_dot(q, q));
            float dir[3] = {adjust_coord(q.x / qn), adjust_coord(q.y / qn), adjust_coord(q.z / qn)};
            trace_hit hh;
            uint32_t v = 0;
            int r = trace_ray(nodes, num_nodes, root, pos, dir, 0, detail, &hh, &v, &bad);
            ++rays;
            if (visits) visits[offs] = v;
            if (r) {
                const uint32_t* n = nodes + (size_t)10 * hh.node;
                if (t) t[offs] = hh.t;
                if (z) z[offs] = hh.t / qn;
                if (albedo) albedo[offs] = hh.child < 0 ? lod_colors[hh.node] : n[2 + hh.child];
                if (level) level[offs] = (uint8_t)hh.level;
            } else {
                if (t) t[offs] = -1.0f;
                if (z) z[offs] = INFINITY;
                if (albedo) albedo[offs] = 0;
                if (level) level[offs] = 0;
            }
        }
    }
    return bad ? -1 : rays;
}

/* Single rays with explicit origins and directions and an LOD coefficient
 * (for the pins): hit node id (O_MISS on a miss), child slot (-1 on an LOD
 * stop), t and level. */
int oracle_lod_trace_rays(const uint32_t* nodes, uint32_t num_nodes, uint32_t root, int n,
                          const float* pos, const float* dir, float detail, uint32_t* out_node,
                          int32_t* out_child, float* out_t, int32_t* out_level) {
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        trace_hit hh;
        uint32_t v = 0;
        if (trace_ray(nodes, num_nodes, root, pos + 3 * i, dir + 3 * i, 0, detail, &hh, &v, &bad)) {
            out_node[i] = hh.node;
            out_child[i] = hh.child;
            out_t[i] = hh.t;
            out_level[i] = hh.level;
        } else {
            out_node[i] = O_MISS;
            out_child[i] = -2;
            out_t[i] = -1.0f;
            out_level[i] = 0;
        }
    }
    return bad ? -1 : 0;
}
