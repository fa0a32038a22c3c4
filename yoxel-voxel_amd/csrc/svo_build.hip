// GPU builder of the synthetic sphere + heightfield octrees (the scene side of
// the hot path; DynamicSVO::BuildRange with SphereSource / IsoSource,
// ore/src/main.cpp:37-129, and the GPU brick classification / compaction of
// reaction/voxtree/density.cu:131-221 as the pattern).
//
// It produces the canonical VoxNode tree of yx_scene_build_synthetic byte for
// byte: the same voxel definition evaluated with the same float64 operations
// (sphere columns, fBm heightfield), the same surface rule, the same
// breadth-first node order (level by level, nodes in Morton order).
//   1. one thread per (x, y) column: the column's inside intervals - every
//      sphere whose bounding box covers the column, the heightfield's
//      [0, top] - sorted by start and merged;
//   2. one thread per column: surface voxels = inside minus the voxels whose
//      z neighbours and four column neighbours are inside; counted, scanned,
//      then written as Morton codes;
//   3. radix sort of the codes; each tree level = the distinct parent codes of
//      the level below (select-unique on code >> 3);
//   4. one thread per node: its children are a contiguous run of the level
//      below (binary search), written as child ids or leaf colours.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "internal.h"
#include "prims.h"

namespace yx {

constexpr int kColMax = 48;  // merged intervals kept per column

struct SbSphere {
    double cx, cy, cz, r;  // voxel units
    int xlo, xhi, ylo, yhi;  // columns the CPU builder visits
};

struct SbInterval {
    int lo, hi;  // inclusive
};

struct SbArgs {
    const SbSphere* spheres;
    int num_spheres;
    const int* perm;  // heightfield noise, 512 entries (null: no heightfield)
    int octaves;
    double base, amp;
    int n;
    SbInterval* iv;  // [n * n][kColMax]
    int* cnt;        // [n * n]
    int* overflow;
    const EditSphere* edits;  // applied in order after the recipe's solid
    const uint32_t* cell_off;   // edits per 64 x 64-column cell (EditSet's CSR)
    const uint32_t* cell_ids;
    int cells;
    const uint32_t* vcol;    // volume sources' voxel colours
    const uint8_t* vsolid;   // and solid flags
};

// Source voxel index of finest voxel (i, j, k) in a volume edit, or -1.
__device__ __forceinline__ int64_t sb_vol_index(const EditSphere& e, int i, int j, int k) {
    if (i < e.vx || j < e.vy || k < e.vz) return -1;
    const int a = (i - e.vx) / e.s, b = (j - e.vy) / e.s, c = (k - e.vz) / e.s;
    if (a >= e.sx || b >= e.sy || c >= e.sz) return -1;
    return (int64_t)e.off + a + (int64_t)e.sx * (b + (int64_t)e.sy * c);
}

__device__ __forceinline__ size_t sb_cell(int cells, int i, int j) {
    return (size_t)(j >> EditSet::kCellShift) * cells + (size_t)(i >> EditSet::kCellShift);
}

__device__ __forceinline__ bool sb_in_sphere(const SbSphere& s, int i, int j, int k) {
    const double dx = ((double)i + 0.5) - s.cx;
    const double dy = ((double)j + 0.5) - s.cy;
    const double dz = ((double)k + 0.5) - s.cz;
    return dx * dx + dy * dy + dz * dz <= s.r * s.r;
}

__device__ __forceinline__ SbInterval sb_sphere_column(const SbSphere& s, int i, int j, int n) {
    const double dx = ((double)i + 0.5) - s.cx;
    const double dy = ((double)j + 0.5) - s.cy;
    const double rem = s.r * s.r - dx * dx - dy * dy;
    SbInterval iv{1, 0};
    if (rem < -1.0) return iv;
    const double h = sqrt(fmax(rem, 0.0));
    int lo = (int)ceil(s.cz - h - 0.5) - 1;
    int hi = (int)floor(s.cz + h - 0.5) + 1;
    lo = max(lo, 0);
    hi = min(hi, n - 1);
    while (lo <= hi && !sb_in_sphere(s, i, j, lo)) ++lo;
    while (hi >= lo && !sb_in_sphere(s, i, j, hi)) --hi;
    iv.lo = lo;
    iv.hi = hi;
    return iv;
}

__device__ __forceinline__ double sb_fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
__device__ __forceinline__ double sb_lerp(double t, double a, double b) { return a + t * (b - a); }

__device__ __forceinline__ double sb_grad2(int h, double x, double y) {
    switch (h & 7) {
        case 0: return x + y;
        case 1: return -x + y;
        case 2: return x - y;
        case 3: return -x - y;
        case 4: return x;
        case 5: return -x;
        case 6: return y;
        default: return -y;
    }
}

__device__ double sb_noise2(const int* P, double x, double y) {
    const double fx = floor(x), fy = floor(y);
    const int xi = (int)((int64_t)fx & 255), yi = (int)((int64_t)fy & 255);
    const double xf = x - fx, yf = y - fy;
    const double u = sb_fade(xf), v = sb_fade(yf);
    const int aa = P[P[xi] + yi], ab = P[P[xi] + yi + 1];
    const int ba = P[P[xi + 1] + yi], bb = P[P[xi + 1] + yi + 1];
    const double x1 = sb_lerp(u, sb_grad2(aa, xf, yf), sb_grad2(ba, xf - 1.0, yf));
    const double x2 = sb_lerp(u, sb_grad2(ab, xf, yf - 1.0), sb_grad2(bb, xf - 1.0, yf - 1.0));
    return sb_lerp(v, x1, x2);
}

__device__ double sb_fbm2(const int* P, double x, double y, int octaves) {
    double sum = 0.0, amp = 1.0, freq = 1.0, norm = 0.0;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * sb_noise2(P, x * freq, y * freq);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm;
}

// Heightfield column: voxels with k + 0.5 < height are inside; returns the
// top inside voxel or -1.
__device__ int sb_top(const SbArgs& a, const int* P, int i, int j) {
    const int n = a.n;
    const double u = ((double)i + 0.5) / (double)n;
    const double v = ((double)j + 0.5) / (double)n;
    const double f = sb_fbm2(P, u * 4.0, v * 4.0, a.octaves);
    const double hgt = (double)n * (a.base + a.amp * f);
    int k = (int)ceil(hgt - 0.5) - 1;
    k = min(max(k, -1), n - 1);
    while (k + 1 <= n - 1 && (double)(k + 1) + 0.5 < hgt) ++k;
    while (k >= 0 && !((double)k + 0.5 < hgt)) --k;
    return k;
}

// 1. merged inside intervals of every column
__global__ __launch_bounds__(256) void sb_columns_kernel(SbArgs a) {
    __shared__ int P[512];
    if (a.perm)
        for (int q = threadIdx.x; q < 512; q += 256) P[q] = a.perm[q];
    __syncthreads();
    const int n = a.n;
    const size_t col = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= (size_t)n * n) return;
    const int i = (int)(col % (size_t)n), j = (int)(col / (size_t)n);
    SbInterval* out = a.iv + col * kColMax;
    int c = 0;
    bool over = false;
    // insertion by start keeps the list sorted; equal starts keep arrival
    // order (the merge gives the same union either way)
    auto insert = [&](SbInterval v) {
        if (c == kColMax) {
            over = true;
            return;
        }
        int p = c;
        while (p > 0 && out[p - 1].lo > v.lo) {
            out[p] = out[p - 1];
            --p;
        }
        out[p] = v;
        ++c;
    };
    for (int s = 0; s < a.num_spheres; ++s) {
        const SbSphere& sp = a.spheres[s];
        if (i < sp.xlo || i > sp.xhi || j < sp.ylo || j > sp.yhi) continue;
        const SbInterval v = sb_sphere_column(sp, i, j, n);
        if (v.lo <= v.hi) insert(v);
    }
    if (a.perm) {
        const int top = sb_top(a, P, i, j);
        if (top >= 0) insert(SbInterval{0, top});
    }
    // merge overlapping and touching intervals
    int m = 0;
    for (int q = 0; q < c; ++q) {
        const SbInterval v = out[q];
        if (m > 0 && v.lo <= out[m - 1].hi + 1)
            out[m - 1].hi = max(out[m - 1].hi, v.hi);
        else
            out[m++] = v;
    }
    // edits in order: union (GROW) or difference (CLEAR) of each source's
    // solid runs; the list stays sorted, disjoint and non-touching
    auto apply = [&](SbInterval v, bool grow) {
        if (v.lo > v.hi || over) return;
        if (grow) {
            if (m == kColMax) {
                over = true;
                return;
            }
            int p = m;
            while (p > 0 && out[p - 1].lo > v.lo) {
                out[p] = out[p - 1];
                --p;
            }
            out[p] = v;
            ++m;
            int w = 0;
            for (int q = 0; q < m; ++q) {
                const SbInterval u = out[q];
                if (w > 0 && u.lo <= out[w - 1].hi + 1)
                    out[w - 1].hi = max(out[w - 1].hi, u.hi);
                else
                    out[w++] = u;
            }
            m = w;
        } else {
            // overlapping intervals are the run [p, q]
            int p = 0;
            while (p < m && out[p].hi < v.lo) ++p;
            int q = p;
            while (q < m && out[q].lo <= v.hi) ++q;
            if (q == p) return;  // nothing inside
            const SbInterval first = out[p], last = out[q - 1];
            SbInterval keep[2];
            int nk = 0;
            if (first.lo < v.lo) keep[nk++] = SbInterval{first.lo, v.lo - 1};
            if (last.hi > v.hi) keep[nk++] = SbInterval{v.hi + 1, last.hi};
            const int removed = q - p, delta = nk - removed;
            if (m + delta > kColMax) {
                over = true;
                return;
            }
            if (delta > 0) {
                for (int r = m - 1; r >= q; --r) out[r + delta] = out[r];
            } else if (delta < 0) {
                for (int r = q; r < m; ++r) out[r + delta] = out[r];
            }
            for (int r = 0; r < nk; ++r) out[p + r] = keep[r];
            m += delta;
        }
    };
    const size_t cell = a.edits ? sb_cell(a.cells, i, j) : 0;
    const uint32_t e0 = a.edits ? a.cell_off[cell] : 0, e1 = a.edits ? a.cell_off[cell + 1] : 0;
    for (uint32_t e = e0; e < e1 && !over; ++e) {
        const EditSphere& x = a.edits[a.cell_ids[e]];
        if (i < x.xlo || i > x.xhi || j < x.ylo || j > x.yhi) continue;
        const bool grow = x.mode == YX_BUILD_GROW;
        if (x.kind == 0) {
            SbSphere sp;
            sp.cx = x.cx;
            sp.cy = x.cy;
            sp.cz = x.cz;
            sp.r = x.r;
            sp.xlo = x.xlo;
            sp.xhi = x.xhi;
            sp.ylo = x.ylo;
            sp.yhi = x.yhi;
            apply(sb_sphere_column(sp, i, j, n), grow);
            continue;
        }
        // a volume source: its solid runs along the column
        const int64_t base = sb_vol_index(x, i, j, x.vz);
        if (base < 0) continue;
        const int64_t stride = (int64_t)x.sx * x.sy;
        int c = 0;
        while (c < x.sz && !over) {
            while (c < x.sz && !a.vsolid[base + stride * c]) ++c;
            if (c == x.sz) break;
            const int c0 = c;
            while (c < x.sz && a.vsolid[base + stride * c]) ++c;
            const int64_t lo = (int64_t)x.vz + (int64_t)c0 * x.s, hi = (int64_t)x.vz + (int64_t)c * x.s - 1;
            apply(SbInterval{(int)max((int64_t)0, lo), (int)min((int64_t)(n - 1), hi)}, grow);
        }
    }
    a.cnt[col] = m;
    if (over) atomicOr(a.overflow, 1);
}

// a := a intersect b (both sorted, disjoint); -1 when the result would hold
// more than kColMax intervals
__device__ __forceinline__ int sb_intersect(SbInterval* a, int na, const SbInterval* b, int nb, SbInterval* tmp) {
    int r = 0, i = 0, j = 0;
    while (i < na && j < nb) {
        const int lo = max(a[i].lo, b[j].lo), hi = min(a[i].hi, b[j].hi);
        if (lo <= hi) {
            if (r == kColMax) return -1;
            tmp[r++] = SbInterval{lo, hi};
        }
        if (a[i].hi < b[j].hi)
            ++i;
        else
            ++j;
    }
    for (int q = 0; q < r; ++q) a[q] = tmp[q];
    return r;
}

__device__ __forceinline__ uint64_t sb_spread(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// 2. surface voxels of every column: counted (codes == null) or written at
// the column's offset
__global__ __launch_bounds__(128) void sb_surface_kernel(SbArgs a, const uint64_t* offsets, uint32_t* counts,
                                                         uint64_t* codes) {
    const int n = a.n;
    const size_t col = (size_t)blockIdx.x * 128 + threadIdx.x;
    if (col >= (size_t)n * n) return;
    const int i = (int)(col % (size_t)n), j = (int)(col / (size_t)n);
    const int cnt = a.cnt[col];
    const SbInterval* I = a.iv + col * kColMax;
    SbInterval interior[kColMax], tmp[kColMax];
    int ni = 0;
    for (int q = 0; q < cnt; ++q)
        if (I[q].lo + 1 <= I[q].hi - 1) interior[ni++] = SbInterval{I[q].lo + 1, I[q].hi - 1};
    const int nb[4][2] = {{i - 1, j}, {i + 1, j}, {i, j - 1}, {i, j + 1}};
    for (int q = 0; q < 4 && ni > 0; ++q) {
        const int x = nb[q][0], y = nb[q][1];
        if (x < 0 || y < 0 || x >= n || y >= n) {
            ni = 0;
            break;
        }
        const size_t c2 = (size_t)y * n + x;
        ni = sb_intersect(interior, ni, a.iv + c2 * kColMax, a.cnt[c2], tmp);
        if (ni < 0) {  // reported to the host, which fails the build
            atomicOr(a.overflow, 1);
            ni = 0;
        }
    }
    const uint64_t mxy = sb_spread((uint64_t)i) | (sb_spread((uint64_t)j) << 1);
    uint64_t w = codes ? offsets[col] : 0;
    uint32_t total = 0;
    int p = 0;
    for (int q = 0; q < cnt; ++q) {
        int k = I[q].lo;
        while (k <= I[q].hi) {
            while (p < ni && interior[p].hi < k) ++p;
            int stop = I[q].hi;
            if (p < ni && interior[p].lo <= k) {
                k = interior[p].hi + 1;
                continue;
            }
            if (p < ni) stop = min(stop, interior[p].lo - 1);
            if (codes) {
                for (; k <= stop; ++k) codes[w++] = mxy | (sb_spread((uint64_t)k) << 2);
            } else {
                total += (uint32_t)(stop - k + 1);
                k = stop + 1;
            }
        }
    }
    if (!codes) counts[col] = total;
}

// Leaf colour of a synthetic scene (the builder's hash of the Morton code:
// splitmix64's finaliser of code ^ seed * golden ratio, low 32 bits).
__device__ __forceinline__ uint32_t sb_leaf_color(uint64_t code, uint64_t seed) {
    uint64_t z = code ^ (seed * 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)(z ^ (z >> 31));
}

__device__ __forceinline__ uint32_t sb_compact3(uint64_t v) {
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
    v = (v ^ (v >> 32)) & 0x1fffffull;
    return (uint32_t)v;
}

// the last edit whose ball holds the voxel decides its colour
__device__ __forceinline__ uint32_t sb_edited_color(uint64_t code, uint64_t seed, const SbArgs& a) {
    if (!a.edits) return sb_leaf_color(code, seed);
    const uint32_t i = sb_compact3(code), j = sb_compact3(code >> 1), k = sb_compact3(code >> 2);
    const size_t cell = sb_cell(a.cells, (int)i, (int)j);
    for (uint32_t q = a.cell_off[cell + 1]; q-- > a.cell_off[cell];) {
        const EditSphere& e = a.edits[a.cell_ids[q]];
        if ((int)i < e.xlo || (int)i > e.xhi || (int)j < e.ylo || (int)j > e.yhi) continue;
        if (e.kind == 0) {
            const double dx = ((double)i + 0.5) - e.cx;
            const double dy = ((double)j + 0.5) - e.cy;
            const double dz = ((double)k + 0.5) - e.cz;
            if (dx * dx + dy * dy + dz * dz <= e.r * e.r)
                return e.mode == YX_BUILD_GROW ? e.color : sb_leaf_color(code, seed);
        } else {
            const int64_t v = sb_vol_index(e, (int)i, (int)j, (int)k);
            if (v >= 0 && a.vsolid[v]) return e.mode == YX_BUILD_GROW ? a.vcol[v] : sb_leaf_color(code, seed);
        }
    }
    return sb_leaf_color(code, seed);
}

// 4. node k of a level: children are the run of `below` whose parent code
// (>> 3) equals this node's code; child ids (internal levels) or leaf colours
// (the last node level, below = leaf codes).
__global__ __launch_bounds__(256) void sb_nodes_kernel(const uint64_t* cur, uint64_t n_cur, const uint64_t* below,
                                                       uint64_t n_below, uint64_t node_base, uint64_t child_base,
                                                       int leaf_level, uint64_t color_seed, SbArgs a,
                                                       uint32_t* nodes) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n_cur) return;
    const uint64_t code = cur[k];
    // lower_bound of code << 3 in below
    uint64_t lo = 0, hi = n_below;
    const uint64_t key = code << 3;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (below[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    uint32_t child[8];
    for (int s = 0; s < 8; ++s) child[s] = YX_EMPTY_NODE;
    uint32_t flags = 0;
    for (uint64_t q = lo; q < n_below && (below[q] >> 3) == code; ++q) {
        const int s = (int)(below[q] & 7u);
        if (leaf_level) {
            child[s] = sb_edited_color(below[q], color_seed, a);
            flags |= 1u << s;
        } else {
            child[s] = (uint32_t)(child_base + q);
        }
    }
    for (int s = 0; s < 8; ++s) {
        const bool present = leaf_level ? ((flags >> s) & 1u) != 0 : child[s] != YX_EMPTY_NODE;
        if (!present) flags |= 1u << (8 + s);
    }
    uint32_t* nd = nodes + (node_base + k) * 10;
    nd[0] = flags;
    nd[1] = 0;
    for (int s = 0; s < 8; ++s) nd[2 + s] = child[s];
}

}  // namespace yx

using namespace yx;

#define YX_HIP(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return fail(YX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

namespace {

// device buffers of one build, freed on every exit
struct SbBufs {
    std::vector<void*> ptrs;
    ~SbBufs() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** p, size_t n) {
        *p = nullptr;
        if (hipMalloc((void**)p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) {
            *p = nullptr;
            return fail(YX_ERR_OOM, "hipMalloc(%zu) in the GPU scene build", n * sizeof(T));
        }
        ptrs.push_back(*p);
        return YX_OK;
    }
};

}  // namespace

// Builds the scene of recipe d with edits into s (nodes, root, counts).
static int gpu_build(const yx_scene_desc* d, const EditSet& edits, yx_scene* s, void* stream) {
    if (d->depth < 1 || d->depth > 12) return fail(YX_ERR_INVALID_ARG, "depth %d: 1..12 on the GPU", d->depth);
    if (d->num_spheres < 0 || d->num_spheres > (1 << 20))
        return fail(YX_ERR_INVALID_ARG, "num_spheres %d", d->num_spheres);
    if (d->heightfield && (d->hf_octaves < 1 || d->hf_octaves > 16))
        return fail(YX_ERR_INVALID_ARG, "hf_octaves %d: 1..16", d->hf_octaves);
    const hipStream_t st = (hipStream_t)stream;
    const int D = d->depth;
    const int n = 1 << D;
    const size_t ncol = (size_t)n * n;

    // spheres exactly as the CPU builder draws them
    SplitMix64 rng(d->seed);
    std::vector<SbSphere> sph((size_t)d->num_spheres);
    for (auto& s : sph) {
        const double cx = rng.uniform(), cy = rng.uniform(), cz = rng.uniform();
        const double r = (double)d->radius_min + ((double)d->radius_max - (double)d->radius_min) * rng.uniform();
        s.cx = cx * n;
        s.cy = cy * n;
        s.cz = cz * n;
        s.r = r;
        s.ylo = std::max(0, (int)std::floor(s.cy - s.r - 1.0));
        s.yhi = std::min(n - 1, (int)std::ceil(s.cy + s.r + 1.0));
        s.xlo = std::max(0, (int)std::floor(s.cx - s.r - 1.0));
        s.xhi = std::min(n - 1, (int)std::ceil(s.cx + s.r + 1.0));
    }
    Perlin noise(d->seed ^ 0xD1B54A32D192ED03ull);

    SbBufs B;
    int rc;
    EditSphere* d_edits = nullptr;
    uint32_t *d_cell_off = nullptr, *d_cell_ids = nullptr, *d_vcol = nullptr;
    uint8_t* d_vsolid = nullptr;
    if (!edits.vcol.empty()) {
        const size_t nv = edits.vcol.size();
        if ((rc = B.alloc(&d_vcol, nv)) != YX_OK || (rc = B.alloc(&d_vsolid, nv)) != YX_OK) return rc;
        YX_HIP(hipMemcpyAsync(d_vcol, edits.vcol.data(), nv * 4, hipMemcpyHostToDevice, st));
        YX_HIP(hipMemcpyAsync(d_vsolid, edits.vsolid.data(), nv, hipMemcpyHostToDevice, st));
    }
    if (!edits.empty()) {
        const size_t ne = edits.edits.size(), no = edits.cell_off.size(), ni = std::max<size_t>(1, edits.cell_ids.size());
        if ((rc = B.alloc(&d_edits, ne)) != YX_OK || (rc = B.alloc(&d_cell_off, no)) != YX_OK ||
            (rc = B.alloc(&d_cell_ids, ni)) != YX_OK)
            return rc;
        YX_HIP(hipMemcpyAsync(d_edits, edits.edits.data(), ne * sizeof(EditSphere), hipMemcpyHostToDevice, st));
        YX_HIP(hipMemcpyAsync(d_cell_off, edits.cell_off.data(), no * 4, hipMemcpyHostToDevice, st));
        if (!edits.cell_ids.empty())
            YX_HIP(hipMemcpyAsync(d_cell_ids, edits.cell_ids.data(), edits.cell_ids.size() * 4,
                                  hipMemcpyHostToDevice, st));
    }
    SbSphere* d_sph = nullptr;
    int *d_perm = nullptr, *d_cnt = nullptr, *d_over = nullptr;
    SbInterval* d_iv = nullptr;
    if ((rc = B.alloc(&d_sph, sph.size())) != YX_OK || (rc = B.alloc(&d_perm, 512)) != YX_OK ||
        (rc = B.alloc(&d_iv, ncol * kColMax)) != YX_OK || (rc = B.alloc(&d_cnt, ncol)) != YX_OK ||
        (rc = B.alloc(&d_over, 1)) != YX_OK)
        return rc;
    if (!sph.empty())
        YX_HIP(hipMemcpyAsync(d_sph, sph.data(), sph.size() * sizeof(SbSphere), hipMemcpyHostToDevice, st));
    YX_HIP(hipMemcpyAsync(d_perm, noise.perm, 512 * sizeof(int), hipMemcpyHostToDevice, st));
    YX_HIP(hipMemsetAsync(d_over, 0, sizeof(int), st));
    SbArgs a;
    a.spheres = d_sph;
    a.num_spheres = (int)sph.size();
    a.perm = d->heightfield ? d_perm : nullptr;
    a.octaves = d->hf_octaves;
    a.base = (double)d->hf_base;
    a.amp = (double)d->hf_amplitude;
    a.n = n;
    a.iv = d_iv;
    a.cnt = d_cnt;
    a.overflow = d_over;
    a.edits = d_edits;
    a.cell_off = d_cell_off;
    a.cell_ids = d_cell_ids;
    a.cells = edits.cells;
    a.vcol = d_vcol;
    a.vsolid = d_vsolid;
    hipLaunchKernelGGL(sb_columns_kernel, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, st, a);
    YX_HIP(hipGetLastError());

    // surface voxel counts -> offsets -> codes
    uint32_t* d_counts = nullptr;
    uint64_t* d_offs = nullptr;
    if ((rc = B.alloc(&d_counts, ncol)) != YX_OK || (rc = B.alloc(&d_offs, ncol + 1)) != YX_OK) return rc;
    hipLaunchKernelGGL(sb_surface_kernel, dim3((unsigned)((ncol + 127) / 128)), dim3(128), 0, st, a,
                       (const uint64_t*)nullptr, d_counts, (uint64_t*)nullptr);
    YX_HIP(hipGetLastError());
    {
        const size_t bytes = scan_workspace_bytes(ncol);
        void* tmp = nullptr;
        if ((rc = B.alloc((uint8_t**)&tmp, bytes)) != YX_OK) return rc;
        // d_offs[ncol] receives the total
        if ((rc = scan_excl_u32_u64(d_counts, d_offs, ncol, d_offs + ncol, tmp, bytes, st)) != YX_OK) return rc;
    }
    int over = 0;
    uint64_t total = 0;
    YX_HIP(hipMemcpyAsync(&over, d_over, sizeof(int), hipMemcpyDeviceToHost, st));
    YX_HIP(hipMemcpyAsync(&total, d_offs + ncol, 8, hipMemcpyDeviceToHost, st));
    YX_HIP(hipStreamSynchronize(st));
    if (over) return fail(YX_ERR_RANGE, "a column holds more than %d inside intervals", kColMax);

    s->num_leaves = total;
    if (total == 0) {
        s->nodes.clear();
        s->root = YX_EMPTY_NODE;
        s->depth = 0;
        return YX_OK;
    }
    // the sort and the per-level unique passes take int counts, and every
    // level holds at most `total` codes: 2^31 - 1 bounds them all
    if (total >= 0x7FFFFFFFull)
        return fail(YX_ERR_RANGE, "%llu surface voxels (at most 2^31 - 2)", (unsigned long long)total);
    uint64_t *d_codes = nullptr, *d_sorted = nullptr, *d_swap = nullptr;
    if ((rc = B.alloc(&d_codes, total)) != YX_OK || (rc = B.alloc(&d_sorted, total)) != YX_OK ||
        (rc = B.alloc(&d_swap, total)) != YX_OK)
        return rc;
    hipLaunchKernelGGL(sb_surface_kernel, dim3((unsigned)((ncol + 127) / 128)), dim3(128), 0, st, a, d_offs,
                       d_counts, d_codes);
    if (hipGetLastError() != hipSuccess) return fail(YX_ERR_HIP, "surface kernel launch");
    const uint64_t* leaf_codes = nullptr;
    {
        const size_t bytes = sort_workspace_bytes(total);
        void* tmp = nullptr;
        if ((rc = B.alloc((uint8_t**)&tmp, bytes)) != YX_OK) return rc;
        uint64_t* const kbuf[2] = {d_sorted, d_swap};
        if ((rc = radix_sort<uint64_t>(d_codes, nullptr, kbuf, nullptr, total, nullptr, 0, 3 * D, &leaf_codes, nullptr,
                                       tmp, bytes, st)) != YX_OK)
            return rc;
    }

    // 3. levels: lv[L] = distinct (lv[L + 1] >> 3), lv[D] = leaf codes
    std::vector<uint64_t*> lv((size_t)D + 1, nullptr);
    std::vector<uint64_t> cnt((size_t)D + 1, 0);
    lv[(size_t)D] = const_cast<uint64_t*>(leaf_codes);
    cnt[(size_t)D] = total;
    uint64_t* d_nsel = nullptr;
    if ((rc = B.alloc(&d_nsel, 1)) != YX_OK) return rc;
    void* uniq_ws = nullptr;
    const size_t uniq_bytes = distinct_workspace_bytes(total);
    if ((rc = B.alloc((uint8_t**)&uniq_ws, uniq_bytes)) != YX_OK) return rc;
    for (int L = D - 1; L >= 0; --L) {
        if ((rc = B.alloc(&lv[(size_t)L], cnt[(size_t)L + 1])) != YX_OK) return rc;
        if ((rc = distinct_shifted_u64(lv[(size_t)L + 1], lv[(size_t)L], cnt[(size_t)L + 1], 3, d_nsel, uniq_ws,
                                       uniq_bytes, st)) != YX_OK)
            return rc;
        uint64_t m = 0;
        if (hipMemcpyAsync(&m, d_nsel, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return fail(YX_ERR_HIP, "level size copy");
        cnt[(size_t)L] = m;
    }
    std::vector<uint64_t> base((size_t)D + 1, 0);
    for (int L = 0; L < D; ++L) base[(size_t)L + 1] = base[(size_t)L] + cnt[(size_t)L];
    const uint64_t total_nodes = base[(size_t)D];
    if (total_nodes >= 0x7FFFFFFFull) return fail(YX_ERR_RANGE, "too many nodes");
    uint32_t* d_nodes = nullptr;
    if ((rc = B.alloc(&d_nodes, total_nodes * 10)) != YX_OK) return rc;
    for (int L = 0; L < D; ++L) {
        const uint64_t m = cnt[(size_t)L];
        hipLaunchKernelGGL(sb_nodes_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, lv[(size_t)L], m,
                           lv[(size_t)L + 1], cnt[(size_t)L + 1], base[(size_t)L], base[(size_t)L + 1],
                           L == D - 1 ? 1 : 0, (uint64_t)d->seed, a, d_nodes);
        if (hipGetLastError() != hipSuccess) return fail(YX_ERR_HIP, "node kernel launch");
    }
    try {
        // headroom for nodes that in-place edits append
        s->nodes.reserve((size_t)(total_nodes + total_nodes / 16 + 1024) * 10);
        s->nodes.resize((size_t)total_nodes * 10);
    } catch (const std::bad_alloc&) {
        return fail(YX_ERR_OOM, "out of host memory for %llu nodes", (unsigned long long)total_nodes);
    }
    if (hipMemcpyAsync(s->nodes.data(), d_nodes, total_nodes * 40, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(YX_ERR_HIP, "node download");
    s->root = 0;
    s->depth = D;
    return YX_OK;
}

extern "C" int yx_scene_build_synthetic_gpu(const yx_scene_desc* d, yx_scene** out, void* stream) {
    clear_error();
    if (!d || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    yx_scene* s = new (std::nothrow) yx_scene();
    if (!s) return fail(YX_ERR_OOM, "out of host memory");
    const int rc = gpu_build(d, EditSet{}, s, stream);
    if (rc != YX_OK) {
        delete s;
        return rc;
    }
    s->has_recipe = true;
    s->recipe = *d;
    *out = s;
    return YX_OK;
}

namespace yx {
int gpu_build_synthetic(const yx_scene_desc& d, const EditSet& edits, yx_scene& out, void* stream) {
    return gpu_build(&d, edits, &out, stream);
}
}  // namespace yx
