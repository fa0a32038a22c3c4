/* CPU restatement of the brick n-tree volume and its ray cast - TEST
 * INFRASTRUCTURE ONLY (the parity oracle for the HIP n-tree builder and
 * tracer; never linked into the product).
 *
 * Volume (the n-tree of ntree/ntree/nodes.h:13-100, ntree/ntree/ntree.h:11-63,
 * ntree/ntree/gpu_nodes.h:6-21, with 8^3 voxel bricks and a root grid):
 *   size N = 8 * 4^L * R voxels per axis: an R^3 root grid, L levels of 4^3
 *   nodes (cell offset (z*4 + y)*4 + x, the ch2ofs convention of
 *   ntree/ntree/nodes.h:99), 8^3 u8 bricks (voxel offset x + 8y + 64z).
 *   density(x, y, z) = 0 where f < threshold, else 1 + min(254, (int)((f -
 *   threshold) * scale)), f = sum_o 0.5^o noise3(p * frequency * 2^o), p =
 *   (voxel + 0.5) / N, noise3 = improved Perlin noise (Perlin, SIGGRAPH 2002)
 *   over a SplitMix64-seeded permutation, all in float32 without contraction.
 *   Bricks with a non-zero voxel are allocated; a node cell is set when its
 *   subtree holds a brick; nodes and bricks are stored level by level in tree
 *   order (root cell, then cell digits from the top), so each node's children
 *   are contiguous: child = base + popcount(mask below the cell).
 *
 * Ray cast (castRay / walkBrick / intersect of reaction/trace/trace.cu:42-139,
 * with the empty-space skipping of the n-tree and one sample per voxel the
 * ray crosses, i.e. a 3D DDA inside bricks):
 *   ray: pixel-corner direction of RendererBase::InitRayDir
 *        (cell/renderer_base.h:50-61), normalised, components below 1e-5
 *        clamped with their sign (PrepareRay, trace.cu:31-33); unit-cube slab
 *        test, t_enter clamped at 0 (trace.cu:99-106).
 *   walk: from the entry voxel (floor(p * N), clamped), look the voxel up from
 *        the root: an empty root cell / node cell of side S is left in one
 *        step through its exit plane; a brick voxel is sampled and left
 *        through its own exit plane. Exit plane time t_a = ((c_a + (d_a > 0))
 *        * S / N - o_a) / d_a per axis, axis = argmin with ties to x, then y;
 *        the coordinate on the exit axis steps by one voxel, the others are
 *        floor((o + t d) * N) kept inside the cell left and never moved
 *        backwards.
 *   sample (walkBrick, trace.cu:51-60, with the transfer table of
 *        volvis.py:145-152): col = palette[density] as RGBA8 / 255, rgb *= a,
 *        accum += col * (1 - accum.a), stop once accum.a > 0.99.
 *   outputs: RGBA8 by float2uint (reaction/trace/utils.h:37-44), t and voxel
 *        of the first non-empty voxel (-1 without one), node records and
 *        voxels read (for byte counts).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define NT_EMPTY 0xFFFFFFFFu
#define NT_MAX_LEVELS 8

typedef struct {
    int32_t size, levels, octaves;
    float frequency, threshold, scale;
    uint64_t seed;
} nt_desc;

/* ---- noise --------------------------------------------------------------- */

static uint64_t sm_next(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void nt_perm(uint64_t seed, int32_t perm[512]) {
    int p[256];
    uint64_t s = seed;
    for (int i = 0; i < 256; ++i) p[i] = i;
    for (int i = 255; i > 0; --i) {
        int j = (int)(sm_next(&s) % (uint64_t)(i + 1));
        int t = p[i];
        p[i] = p[j];
        p[j] = t;
    }
    for (int i = 0; i < 512; ++i) perm[i] = p[i & 255];
}

static float fadef(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
static float lerpf(float t, float a, float b) { return a + t * (b - a); }

static float grad3f(int h, float x, float y, float z) {
    switch (h & 15) {
        case 0: return x + y;
        case 1: return -x + y;
        case 2: return x - y;
        case 3: return -x - y;
        case 4: return x + z;
        case 5: return -x + z;
        case 6: return x - z;
        case 7: return -x - z;
        case 8: return y + z;
        case 9: return -y + z;
        case 10: return y - z;
        case 11: return -y - z;
        case 12: return x + y;
        case 13: return -y + z;
        case 14: return -x + y;
        default: return -y - z;
    }
}

static float noise3f(const int32_t* P, float x, float y, float z) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    const float xf = x - fx, yf = y - fy, zf = z - fz;
    const float u = fadef(xf), v = fadef(yf), w = fadef(zf);
    const int a = P[X] + Y, aa = P[a] + Z, ab = P[a + 1] + Z;
    const int b = P[X + 1] + Y, ba = P[b] + Z, bb = P[b + 1] + Z;
    const float x1 = xf - 1.0f, y1 = yf - 1.0f, z1 = zf - 1.0f;
    const float l1 = lerpf(u, grad3f(P[aa], xf, yf, zf), grad3f(P[ba], x1, yf, zf));
    const float l2 = lerpf(u, grad3f(P[ab], xf, y1, zf), grad3f(P[bb], x1, y1, zf));
    const float l3 = lerpf(u, grad3f(P[aa + 1], xf, yf, z1), grad3f(P[ba + 1], x1, yf, z1));
    const float l4 = lerpf(u, grad3f(P[ab + 1], xf, y1, z1), grad3f(P[bb + 1], x1, y1, z1));
    return lerpf(w, lerpf(v, l1, l2), lerpf(v, l3, l4));
}

float nt_fbm(const nt_desc* d, const int32_t* P, int x, int y, int z) {
    const float inv = 1.0f / (float)d->size;
    const float px = ((float)x + 0.5f) * inv * d->frequency;
    const float py = ((float)y + 0.5f) * inv * d->frequency;
    const float pz = ((float)z + 0.5f) * inv * d->frequency;
    float f = 0.0f, amp = 1.0f, fr = 1.0f;
    for (int o = 0; o < d->octaves; ++o) {
        f = f + amp * noise3f(P, px * fr, py * fr, pz * fr);
        amp = amp * 0.5f;
        fr = fr * 2.0f;
    }
    return f;
}

static uint8_t density_of(const nt_desc* d, float f) {
    if (f < d->threshold) return 0;
    const float q = (f - d->threshold) * d->scale;
    return (uint8_t)(1 + (int)fminf(q, 254.0f));
}

/* densities of the box [x0, x0+nx) x [y0, y0+ny) x [z0, z0+nz), x fastest */
void nt_density_box(const nt_desc* d, int x0, int y0, int z0, int nx, int ny, int nz, uint8_t* out) {
    int32_t P[512];
    nt_perm(d->seed, P);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                out[((size_t)z * ny + y) * nx + x] = density_of(d, nt_fbm(d, P, x0 + x, y0 + y, z0 + z));
}

/* fraction of non-zero voxels at n pseudo-random voxel positions */
double nt_occupancy_sample(const nt_desc* d, int64_t n, uint64_t seed) {
    int32_t P[512];
    nt_perm(d->seed, P);
    uint64_t s = seed;
    int64_t hit = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int x = (int)(sm_next(&s) % (uint64_t)d->size);
        const int y = (int)(sm_next(&s) % (uint64_t)d->size);
        const int z = (int)(sm_next(&s) % (uint64_t)d->size);
        hit += density_of(d, nt_fbm(d, P, x, y, z)) != 0;
    }
    return (double)hit / (double)n;
}

/* ---- tree from a dense volume ---------------------------------------------
 * dense: N^3 densities, x fastest. Writes root[R^3] (level-1 node index, or
 * brick index when L = 0, or NT_EMPTY), nodes (4 u32 each: mask lo, mask hi,
 * base, 0), bricks (512 B each). With caps too small only the counts are
 * returned (*n_nodes, *n_bricks); returns 0, or -1 on a bad shape. */
static int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

int nt_build_dense(const uint8_t* dense, int size, int levels, uint32_t* root, uint32_t* nodes, uint8_t* bricks,
                   uint64_t cap_nodes, uint64_t cap_bricks, uint64_t* n_nodes, uint64_t* n_bricks) {
    if (size < 8 || (size & (size - 1)) || levels < 0) return -1;
    const int nb = size / 8;  /* bricks per axis */
    const int lb = ilog2(nb);
    if (2 * levels > lb) return -1;
    const int R = nb >> (2 * levels);
    /* brick occupancy in tree order */
    const uint64_t nbr = (uint64_t)nb * nb * nb;
    uint8_t* occ = (uint8_t*)calloc(nbr, 1);
    uint64_t* order = (uint64_t*)malloc(nbr * sizeof(uint64_t));
    if (!occ || !order) {
        free(occ);
        free(order);
        return -1;
    }
    for (int bz = 0; bz < nb; ++bz)
        for (int by = 0; by < nb; ++by)
            for (int bx = 0; bx < nb; ++bx) {
                int any = 0;
                for (int z = 0; z < 8 && !any; ++z)
                    for (int y = 0; y < 8 && !any; ++y)
                        for (int x = 0; x < 8 && !any; ++x)
                            any = dense[(((size_t)(bz * 8 + z) * size) + (by * 8 + y)) * size + (bx * 8 + x)] != 0;
                const int sh = 2 * levels;
                uint64_t idx = (uint64_t)((((bz >> sh) * R) + (by >> sh)) * R + (bx >> sh));
                for (int l = 1; l <= levels; ++l) {
                    const int s = 2 * (levels - l);
                    const int dx = (bx >> s) & 3, dy = (by >> s) & 3, dz = (bz >> s) & 3;
                    idx = idx * 64 + (uint64_t)((dz * 4 + dy) * 4 + dx);
                }
                occ[idx] = (uint8_t)any;
                order[idx] = ((uint64_t)bz << 42) | ((uint64_t)by << 21) | (uint64_t)bx;
            }
    /* per level, occupancy of each unit in tree order: level L+1 = bricks */
    uint8_t** lev = (uint8_t**)calloc((size_t)levels + 2, sizeof(uint8_t*));
    uint64_t* cnt = (uint64_t*)calloc((size_t)levels + 2, sizeof(uint64_t));
    lev[levels + 1] = occ;
    cnt[levels + 1] = nbr;
    for (int l = levels; l >= 1; --l) {
        cnt[l] = cnt[l + 1] / 64;
        lev[l] = (uint8_t*)calloc(cnt[l], 1);
        for (uint64_t i = 0; i < cnt[l]; ++i) {
            int any = 0;
            for (int c = 0; c < 64; ++c) any |= lev[l + 1][i * 64 + c];
            lev[l][i] = (uint8_t)any;
        }
    }
    /* indices: nodes of level 1 first, then 2, ...; bricks in tree order */
    uint64_t total_nodes = 0, total_bricks = 0;
    uint64_t* first = (uint64_t*)calloc((size_t)levels + 2, sizeof(uint64_t));
    for (int l = 1; l <= levels; ++l) {
        first[l] = total_nodes;
        for (uint64_t i = 0; i < cnt[l]; ++i) total_nodes += lev[l][i];
    }
    for (uint64_t i = 0; i < nbr; ++i) total_bricks += occ[i];
    *n_nodes = total_nodes;
    *n_bricks = total_bricks;
    int rc = 0;
    if (cap_nodes >= total_nodes && cap_bricks >= total_bricks && root) {
        /* running index of each level's units in tree order */
        uint64_t** idx = (uint64_t**)calloc((size_t)levels + 2, sizeof(uint64_t*));
        for (int l = 1; l <= levels + 1; ++l) {
            idx[l] = (uint64_t*)malloc(cnt[l] * sizeof(uint64_t));
            uint64_t k = l <= levels ? first[l] : 0;
            for (uint64_t i = 0; i < cnt[l]; ++i) idx[l][i] = lev[l][i] ? k++ : NT_EMPTY;
        }
        const uint64_t nroot = (uint64_t)R * R * R;
        for (uint64_t r = 0; r < nroot; ++r) root[r] = lev[1][r] ? (uint32_t)idx[1][r] : NT_EMPTY;
        for (int l = 1; l <= levels; ++l)
            for (uint64_t i = 0; i < cnt[l]; ++i) {
                if (!lev[l][i]) continue;
                uint64_t mask = 0, base = NT_EMPTY;
                for (int c = 0; c < 64; ++c)
                    if (lev[l + 1][i * 64 + c]) {
                        if (base == NT_EMPTY) base = idx[l + 1][i * 64 + c];
                        mask |= 1ull << c;
                    }
                uint32_t* n = nodes + 4 * idx[l][i];
                n[0] = (uint32_t)mask;
                n[1] = (uint32_t)(mask >> 32);
                n[2] = (uint32_t)base;
                n[3] = 0;
            }
        for (uint64_t i = 0; i < nbr; ++i) {
            if (!occ[i]) continue;
            const uint64_t o = order[i];
            const int bx = (int)(o & 0x1FFFFF), by = (int)((o >> 21) & 0x1FFFFF), bz = (int)(o >> 42);
            uint8_t* b = bricks + 512 * idx[levels + 1][i];
            for (int z = 0; z < 8; ++z)
                for (int y = 0; y < 8; ++y)
                    for (int x = 0; x < 8; ++x)
                        b[x + 8 * y + 64 * z] =
                            dense[(((size_t)(bz * 8 + z) * size) + (by * 8 + y)) * size + (bx * 8 + x)];
        }
        for (int l = 1; l <= levels + 1; ++l) free(idx[l]);
        free(idx);
    }
    for (int l = 1; l <= levels; ++l) free(lev[l]);
    free(lev);
    free(cnt);
    free(first);
    free(occ);
    free(order);
    return rc;
}

/* ---- ray cast ------------------------------------------------------------ */

typedef struct {
    const uint32_t* root;
    const uint32_t* nodes;
    const uint8_t* bricks;
    int size, levels;
} nt_tree;

static float clamp_eps(float x) {
    const float eps = 1e-5f;
    return fabsf(x) < eps ? copysignf(eps, x) : x;
}

static float saturatef(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

/* one ray; returns outputs through the pointers */
static void trace_one(const nt_tree* T, const float* view, const uint32_t* pal, int x, int y, uint32_t* rgba,
                      float* thit, int32_t* hv, uint32_t* nnodes, uint32_t* nvox) {
    const float* b = view + 3;
    float dx = b[0] + b[3] * (float)x + b[6] * (float)y;
    float dy = b[1] + b[4] * (float)x + b[7] * (float)y;
    float dz = b[2] + b[5] * (float)x + b[8] * (float)y;
    float n2 = 0.0f;
    n2 += dx * dx;
    n2 += dy * dy;
    n2 += dz * dz;
    const float nrm = sqrtf(n2);
    dx = clamp_eps(dx / nrm);
    dy = clamp_eps(dy / nrm);
    dz = clamp_eps(dz / nrm);
    const float o[3] = {view[0], view[1], view[2]};
    const float d[3] = {dx, dy, dz};
    const float inv[3] = {1.0f / dx, 1.0f / dy, 1.0f / dz};
    *rgba = 0;
    *thit = -1.0f;
    hv[0] = hv[1] = hv[2] = -1;
    *nnodes = 0;
    *nvox = 0;
    float te = -INFINITY, tx = INFINITY;
    for (int a = 0; a < 3; ++a) {
        const float lo = (0.0f - o[a]) * inv[a], hi = (1.0f - o[a]) * inv[a];
        te = fmaxf(te, fminf(lo, hi));
        tx = fminf(tx, fmaxf(lo, hi));
    }
    te = fmaxf(te, 0.0f);
    if (tx < 0.0f || te > tx) return;
    const int N = T->size, L = T->levels;
    const float nf = (float)N, invN = 1.0f / nf;
    const int lr = ilog2(N / 8) - 2 * L; /* log2 of the root grid side */
    const int R = 1 << lr;
    int v[3];
    for (int a = 0; a < 3; ++a) {
        const float p = o[a] + te * d[a];
        int c = (int)floorf(p * nf);
        v[a] = c < 0 ? 0 : (c > N - 1 ? N - 1 : c);
    }
    float t = te;
    float ar = 0.0f, ag = 0.0f, ab = 0.0f, aa = 0.0f;
    int first = 1;
    const int sgn[3] = {d[0] > 0.0f ? 1 : -1, d[1] > 0.0f ? 1 : -1, d[2] > 0.0f ? 1 : -1};
    int bk_valid = 0, bk[3] = {0, 0, 0};
    uint32_t bk_idx = 0;
    int nc_valid[NT_MAX_LEVELS + 1] = {0}, nc_key[NT_MAX_LEVELS + 1][3];
    uint64_t nc_mask[NT_MAX_LEVELS + 1];
    uint32_t nc_base[NT_MAX_LEVELS + 1];
    if (L > NT_MAX_LEVELS) return;
    for (int guard = 0; guard < 4 * N; ++guard) {
        /* look the voxel up from the root: empty cell side S, or brick idx.
         * The brick last found is reused while the voxel stays in it, and the
         * node record of each level is re-read only when the voxel has left
         * that node's cube (node reads are counted at those re-reads). */
        int S = N / R; /* root cell side in voxels */
        uint32_t idx;
        int l = 0;
        if (bk_valid && (v[0] >> 3) == bk[0] && (v[1] >> 3) == bk[1] && (v[2] >> 3) == bk[2]) {
            idx = bk_idx;
            S = 8;
        } else {
            bk_valid = 0;
            idx = T->root[((v[2] >> (3 + 2 * L)) * R + (v[1] >> (3 + 2 * L))) * R + (v[0] >> (3 + 2 * L))];
            while (idx != NT_EMPTY && l < L) {
                ++l;
                const int sn = 3 + 2 * (L - l + 1); /* log2 side of this node's cube */
                const int key[3] = {v[0] >> sn, v[1] >> sn, v[2] >> sn};
                if (!(nc_valid[l] && nc_key[l][0] == key[0] && nc_key[l][1] == key[1] && nc_key[l][2] == key[2])) {
                    const uint32_t* nd = T->nodes + 4 * (size_t)idx;
                    ++*nnodes;
                    nc_mask[l] = (uint64_t)nd[0] | ((uint64_t)nd[1] << 32);
                    nc_base[l] = nd[2];
                    nc_key[l][0] = key[0];
                    nc_key[l][1] = key[1];
                    nc_key[l][2] = key[2];
                    nc_valid[l] = 1;
                }
                S >>= 2;
                const int s = sn - 2;
                const int c = ((((v[2] >> s) & 3) * 4) + ((v[1] >> s) & 3)) * 4 + ((v[0] >> s) & 3);
                const uint64_t mask = nc_mask[l];
                if (!((mask >> c) & 1)) {
                    idx = NT_EMPTY;
                    break;
                }
                const uint64_t below = c ? (mask & ((1ull << c) - 1)) : 0;
                idx = nc_base[l] + (uint32_t)__builtin_popcountll(below);
            }
            if (idx != NT_EMPTY) {
                bk_valid = 1;
                bk[0] = v[0] >> 3;
                bk[1] = v[1] >> 3;
                bk[2] = v[2] >> 3;
                bk_idx = idx;
            }
        }
        int cs; /* side of the cell to leave */
        if (idx == NT_EMPTY) {
            cs = S;
#if ORACLE_PLANT == 7 /* planted: an empty node cell skips its parent's whole cell */
            if (l > 0) cs = S * 4;
#endif
        } else {
            /* brick voxel: sample it */
            cs = 1;
            const uint8_t dd = T->bricks[512 * (size_t)idx + (size_t)((v[0] & 7) + 8 * (v[1] & 7) + 64 * (v[2] & 7))];
            ++*nvox;
            if (dd) {
                const uint32_t p = pal[dd];
                const float k = 1.0f / 255.0f;
                float cr = (float)(p & 255u) * k, cg = (float)((p >> 8) & 255u) * k;
                float cb = (float)((p >> 16) & 255u) * k, ca = (float)(p >> 24) * k;
#if ORACLE_PLANT != 8 /* planted: colour not weighted by its opacity */
                cr = cr * ca;
                cg = cg * ca;
                cb = cb * ca;
#endif
                const float s = 1.0f - aa;
                ar = ar + cr * s;
                ag = ag + cg * s;
                ab = ab + cb * s;
                aa = aa + ca * s;
                if (first) {
                    first = 0;
                    *thit = t;
                    hv[0] = v[0];
                    hv[1] = v[1];
                    hv[2] = v[2];
                }
                if (aa > 0.99f) break;
            }
        }
        /* leave the cell of side cs containing v */
        int c[3];
        float ta[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = v[a] / cs;
            const int plane = (c[a] + (sgn[a] > 0 ? 1 : 0)) * cs;
            ta[a] = ((float)plane * invN - o[a]) * inv[a];
        }
        const int ax = (ta[0] <= ta[1] && ta[0] <= ta[2]) ? 0 : (ta[1] <= ta[2] ? 1 : 2);
        t = ta[ax];
        if (t > tx) break;
        int nv[3];
        for (int a = 0; a < 3; ++a) {
            if (a == ax) {
                nv[a] = sgn[a] > 0 ? (c[a] + 1) * cs : c[a] * cs - 1;
            } else {
                int q = (int)floorf((o[a] + t * d[a]) * nf);
                const int lo = c[a] * cs, hi = lo + cs - 1;
                if (sgn[a] > 0) {
                    q = q < v[a] ? v[a] : q;
                } else {
                    q = q > v[a] ? v[a] : q;
                }
                nv[a] = q < lo ? lo : (q > hi ? hi : q);
            }
        }
        if (nv[ax] < 0 || nv[ax] >= N) break;
        v[0] = nv[0];
        v[1] = nv[1];
        v[2] = nv[2];
    }
    *rgba = (uint32_t)(saturatef(ar) * 255.0f) | ((uint32_t)(saturatef(ag) * 255.0f) << 8) |
            ((uint32_t)(saturatef(ab) * 255.0f) << 16) | ((uint32_t)(saturatef(aa) * 255.0f) << 24);
}

typedef struct {
    const nt_tree* T;
    const float* view;
    const uint32_t* pal;
    int w, h, y0, y1, ystep;
    uint32_t* rgba;
    float* t;
    int32_t* voxel;
    uint32_t* nodes;
    uint32_t* voxels;
} job_t;

static void* run_rows(void* p) {
    job_t* j = (job_t*)p;
    for (int y = j->y0; y < j->y1; y += j->ystep)
        for (int x = 0; x < j->w; ++x) {
            const size_t i = (size_t)y * j->w + x;
            trace_one(j->T, j->view, j->pal, x, y, &j->rgba[i], &j->t[i], &j->voxel[3 * i], &j->nodes[i],
                      &j->voxels[i]);
        }
    return NULL;
}

/* Render rows y0, y0 + ystep, ... < h of a w x h view (outputs indexed by the
 * full image) with `threads` threads over interleaved row sets. */
void nt_render(const uint32_t* root, const uint32_t* nodes, const uint8_t* bricks, int size, int levels,
               const float view[12], const uint32_t palette[256], int w, int h, int ystep, int threads,
               uint32_t* rgba, float* t, int32_t* voxel, uint32_t* nnodes, uint32_t* nvox) {
    nt_tree T = {root, nodes, bricks, size, levels};
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    job_t jobs[64];
    for (int k = 0; k < threads; ++k) {
        jobs[k] = (job_t){&T, view, palette, w, h, k * ystep, h, threads * ystep, rgba, t, voxel, nnodes, nvox};
        pthread_create(&th[k], NULL, run_rows, &jobs[k]);
    }
    for (int k = 0; k < threads; ++k) pthread_join(th[k], NULL);
}
