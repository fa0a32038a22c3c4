// Scene side of libyoxel: synthetic SVO builder, .vox I/O, and the re-layout of
// a VoxNode tree into the unified 8-byte GPU record array.
//
// Node format and .vox layout follow the reference: VoxNode
// (reaction/report/main.tex:36-56) and SVOData::Load (cell/svodata.h:31-50).
// The reference's builder (DynamicSVO::BuildRange, missing from the snapshot)
// is replaced by a column-interval voxelizer of sphere and heightfield
// sources, the sources the reference's scene scripts combine
// (scene_gen.py:87-130, gen_spheres.py:16-32).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <thread>

#include "internal.h"

namespace yx {

static thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

void clear_error() { g_error.clear(); }

int host_threads() {
    int n = (int)std::thread::hardware_concurrency();
    if (n <= 0) n = 1;
    const char* e = getenv("YX_THREADS");
    if (!e) e = getenv("OMP_NUM_THREADS");
    if (e && atoi(e) > 0) n = std::min(n, atoi(e));
    return std::max(1, std::min(n, 16));
}

template <class F>
static void parallel_for(int n, F&& f) {
    int nt = std::min(host_threads(), std::max(1, n));
    std::atomic<int> next{0};
    auto worker = [&]() {
        for (;;) {
            int i = next.fetch_add(1);
            if (i >= n) break;
            f(i);
        }
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < nt; ++t) ts.emplace_back(worker);
    worker();
    for (auto& t : ts) t.join();
}

// ---- Perlin improved noise ------------------------------------------------

Perlin::Perlin(uint64_t seed) {
    SplitMix64 rng(seed);
    int p[256];
    for (int i = 0; i < 256; ++i) p[i] = i;
    for (int i = 255; i > 0; --i) {
        int j = (int)(rng.next() % (uint64_t)(i + 1));
        std::swap(p[i], p[j]);
    }
    for (int i = 0; i < 512; ++i) perm[i] = p[i & 255];
}

static inline double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
static inline double lerp(double t, double a, double b) { return a + t * (b - a); }

static inline double grad2(int h, double x, double y) {
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

static inline double grad3(int h, double x, double y, double z) {
    // the 12 cube-edge directions of the improved noise, 4 repeated
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

double Perlin::noise2(double x, double y) const {
    double fx = std::floor(x), fy = std::floor(y);
    int xi = (int)((int64_t)fx & 255), yi = (int)((int64_t)fy & 255);
    double xf = x - fx, yf = y - fy;
    double u = fade(xf), v = fade(yf);
    int aa = perm[perm[xi] + yi], ab = perm[perm[xi] + yi + 1];
    int ba = perm[perm[xi + 1] + yi], bb = perm[perm[xi + 1] + yi + 1];
    double x1 = lerp(u, grad2(aa, xf, yf), grad2(ba, xf - 1.0, yf));
    double x2 = lerp(u, grad2(ab, xf, yf - 1.0), grad2(bb, xf - 1.0, yf - 1.0));
    return lerp(v, x1, x2);
}

double Perlin::noise3(double x, double y, double z) const {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int xi = (int)((int64_t)fx & 255), yi = (int)((int64_t)fy & 255), zi = (int)((int64_t)fz & 255);
    double xf = x - fx, yf = y - fy, zf = z - fz;
    double u = fade(xf), v = fade(yf), w = fade(zf);
    int a = perm[xi] + yi, aa = perm[a] + zi, ab = perm[a + 1] + zi;
    int b = perm[xi + 1] + yi, ba = perm[b] + zi, bb = perm[b + 1] + zi;
    double l1 = lerp(u, grad3(perm[aa], xf, yf, zf), grad3(perm[ba], xf - 1.0, yf, zf));
    double l2 = lerp(u, grad3(perm[ab], xf, yf - 1.0, zf), grad3(perm[bb], xf - 1.0, yf - 1.0, zf));
    double l3 = lerp(u, grad3(perm[aa + 1], xf, yf, zf - 1.0),
                     grad3(perm[ba + 1], xf - 1.0, yf, zf - 1.0));
    double l4 = lerp(u, grad3(perm[ab + 1], xf, yf - 1.0, zf - 1.0),
                     grad3(perm[bb + 1], xf - 1.0, yf - 1.0, zf - 1.0));
    return lerp(w, lerp(v, l1, l2), lerp(v, l3, l4));
}

double Perlin::fbm2(double x, double y, int octaves) const {
    double sum = 0.0, amp = 1.0, freq = 1.0, norm = 0.0;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * noise2(x * freq, y * freq);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm;
}

double Perlin::fbm3(double x, double y, double z, int octaves) const {
    double sum = 0.0, amp = 1.0, freq = 1.0, norm = 0.0;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * noise3(x * freq, y * freq, z * freq);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm;
}

// ---- tree from sorted leaf codes ------------------------------------------

static inline uint32_t compact3(uint64_t v) {
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
    v = (v ^ (v >> 32)) & 0x1fffffull;
    return (uint32_t)v;
}

// Source voxel index of finest voxel (i, j, k) in a volume edit, or -1.
static inline int64_t vol_index(const EditSphere& e, int i, int j, int k) {
    if (i < e.vx || j < e.vy || k < e.vz) return -1;
    const int a = (i - e.vx) / e.s, b = (j - e.vy) / e.s, c = (k - e.vz) / e.s;
    if (a >= e.sx || b >= e.sy || c >= e.sz) return -1;
    return (int64_t)e.off + a + (int64_t)e.sx * (b + (int64_t)e.sy * c);
}

// Whether the edit's source holds finest voxel (i, j, k); for a volume
// source, its colour in *col.
static inline bool in_edit(const EditSet& set, const EditSphere& e, uint32_t i, uint32_t j, uint32_t k,
                           uint32_t* col) {
    if (e.kind == 0) {
        double dx = ((double)i + 0.5) - e.cx;
        double dy = ((double)j + 0.5) - e.cy;
        double dz = ((double)k + 0.5) - e.cz;
        *col = e.color;
        return dx * dx + dy * dy + dz * dz <= e.r * e.r;
    }
    const int64_t v = vol_index(e, (int)i, (int)j, (int)k);
    if (v < 0 || !set.vsolid[(size_t)v]) return false;
    *col = set.vcol[(size_t)v];
    return true;
}

// Leaf colour: the last edit whose source holds the voxel decides (a GROW's
// colour, or after a CLEAR the recipe's); without one, the recipe's hash.
// Edits with ids >= limit are not applied yet.
static inline uint32_t edited_color(uint64_t code, uint64_t seed, const EditSet* edits, size_t limit) {
    if (edits && !edits->empty()) {
        const uint32_t i = compact3(code), j = compact3(code >> 1), k = compact3(code >> 2);
        // newest first; the pointer never moves below the cell's first id (an
        // empty cell of an empty id list has none at all)
        const uint32_t* b = edits->begin((int)i, (int)j);
        for (const uint32_t* p = edits->end((int)i, (int)j); p != b;) {
            --p;
            if (*p >= limit) continue;
            const EditSphere& e = edits->edits[*p];
            if ((int)i < e.xlo || (int)i > e.xhi || (int)j < e.ylo || (int)j > e.yhi) continue;
            uint32_t col = 0;
            if (in_edit(*edits, e, i, j, k, &col)) return e.mode == YX_BUILD_GROW ? col : leaf_color(code, seed);
        }
    }
    return leaf_color(code, seed);
}

void build_tree_from_codes(const std::vector<uint64_t>& codes, int depth, uint64_t color_seed,
                           yx_scene& out, const EditSet* edits) {
    out.nodes.clear();
    out.num_leaves = codes.size();
    out.depth = codes.empty() ? 0 : depth;
    if (codes.empty()) {
        out.root = YX_EMPTY_NODE;
        return;
    }
    // level L (0 = root) node codes: codes >> 3*(depth-L), unique
    std::vector<std::vector<uint64_t>> lv(depth);
    {
        const int L = depth - 1;
        auto& v = lv[L];
        v.reserve(codes.size() / 2 + 1);
        for (uint64_t c : codes) {
            uint64_t p = c >> 3;
            if (v.empty() || v.back() != p) v.push_back(p);
        }
    }
    for (int L = depth - 2; L >= 0; --L) {
        auto& v = lv[L];
        for (uint64_t c : lv[L + 1]) {
            uint64_t p = c >> 3;
            if (v.empty() || v.back() != p) v.push_back(p);
        }
    }
    std::vector<uint64_t> base(depth + 1, 0);
    for (int L = 0; L < depth; ++L) base[L + 1] = base[L] + lv[L].size();
    const uint64_t total = base[depth];
    // headroom for nodes that in-place edits append
    out.nodes.reserve((total + total / 16 + 1024) * 10);
    out.nodes.assign(total * 10, 0);
    out.root = 0;
    // internal levels
    for (int L = 0; L < depth - 1; ++L) {
        const auto& cur = lv[L];
        const auto& nxt = lv[L + 1];
        const int nchunks = std::max<int>(1, (int)(cur.size() / 65536));
        parallel_for(nchunks, [&](int ci) {
            size_t lo = cur.size() * (size_t)ci / nchunks, hi = cur.size() * (size_t)(ci + 1) / nchunks;
            if (lo >= hi) return;
            size_t j = std::lower_bound(nxt.begin(), nxt.end(), cur[lo] << 3) - nxt.begin();
            for (size_t k = lo; k < hi; ++k) {
                uint32_t* n = &out.nodes[(base[L] + k) * 10];
                uint32_t flags = 0;
                for (int s = 0; s < 8; ++s) n[2 + s] = YX_EMPTY_NODE;
                while (j < nxt.size() && (nxt[j] >> 3) == cur[k]) {
                    int s = (int)(nxt[j] & 7);
                    n[2 + s] = (uint32_t)(base[L + 1] + j);
                    ++j;
                }
                for (int s = 0; s < 8; ++s)
                    if (n[2 + s] == YX_EMPTY_NODE) flags |= 1u << (8 + s);
                n[0] = flags;
                n[1] = 0;
            }
        });
    }
    // last node level: leaf children
    {
        const int L = depth - 1;
        const auto& cur = lv[L];
        const int nchunks = std::max<int>(1, (int)(cur.size() / 65536));
        parallel_for(nchunks, [&](int ci) {
            size_t lo = cur.size() * (size_t)ci / nchunks, hi = cur.size() * (size_t)(ci + 1) / nchunks;
            if (lo >= hi) return;
            size_t j = std::lower_bound(codes.begin(), codes.end(), cur[lo] << 3) - codes.begin();
            for (size_t k = lo; k < hi; ++k) {
                uint32_t* n = &out.nodes[(base[L] + k) * 10];
                uint32_t flags = 0;
                for (int s = 0; s < 8; ++s) n[2 + s] = YX_EMPTY_NODE;
                while (j < codes.size() && (codes[j] >> 3) == cur[k]) {
                    int s = (int)(codes[j] & 7);
                    n[2 + s] = edited_color(codes[j], color_seed, edits, edits ? edits->edits.size() : 0);
                    flags |= 1u << s;
                    ++j;
                }
                for (int s = 0; s < 8; ++s)
                    if (!(flags & (1u << s))) flags |= 1u << (8 + s);
                n[0] = flags;
                n[1] = 0;
            }
        });
    }
}

// ---- synthetic voxelizer --------------------------------------------------

namespace {

struct Sphere {
    double cx, cy, cz, r;  // voxel units
};

struct Interval {
    int32_t lo, hi;  // inclusive
};

// Inside predicate of a sphere at voxel (i, j, k): centre distance^2 <= r^2,
// evaluated in double in this exact order.
inline bool in_sphere(const Sphere& s, int i, int j, int k) {
    double dx = ((double)i + 0.5) - s.cx;
    double dy = ((double)j + 0.5) - s.cy;
    double dz = ((double)k + 0.5) - s.cz;
    return dx * dx + dy * dy + dz * dz <= s.r * s.r;
}

// Column z-range of a sphere (empty when lo > hi).
inline Interval sphere_column(const Sphere& s, int i, int j, int n) {
    double dx = ((double)i + 0.5) - s.cx;
    double dy = ((double)j + 0.5) - s.cy;
    double rem = s.r * s.r - dx * dx - dy * dy;
    Interval iv{1, 0};
    if (rem < -1.0) return iv;
    double h = std::sqrt(std::max(rem, 0.0));
    int lo = (int)std::ceil(s.cz - h - 0.5) - 1;
    int hi = (int)std::floor(s.cz + h - 0.5) + 1;
    lo = std::max(lo, 0);
    hi = std::min(hi, n - 1);
    while (lo <= hi && !in_sphere(s, i, j, lo)) ++lo;
    while (hi >= lo && !in_sphere(s, i, j, hi)) --hi;
    iv.lo = lo;
    iv.hi = hi;
    return iv;
}

struct Heightfield {
    const Perlin* noise;
    int octaves;
    double base, amp;
    int n;
    // Height in voxel units at column (i, j).
    double height(int i, int j) const {
        double u = ((double)i + 0.5) / (double)n;
        double v = ((double)j + 0.5) / (double)n;
        double f = noise->fbm2(u * 4.0, v * 4.0, octaves);
        return (double)n * (base + amp * f);
    }
    // Voxels with k + 0.5 < height are inside.
    int top(int i, int j) const {
        double hgt = height(i, j);
        int k = (int)std::ceil(hgt - 0.5) - 1;
        k = std::min(std::max(k, -1), n - 1);
        while (k + 1 <= n - 1 && (double)(k + 1) + 0.5 < hgt) ++k;
        while (k >= 0 && !((double)k + 0.5 < hgt)) --k;
        return k;
    }
};

// a := a intersect b
void intersect(std::vector<Interval>& a, const Interval* b, size_t nb) {
    std::vector<Interval> r;
    size_t i = 0, j = 0;
    while (i < a.size() && j < nb) {
        int lo = std::max(a[i].lo, b[j].lo), hi = std::min(a[i].hi, b[j].hi);
        if (lo <= hi) r.push_back({lo, hi});
        if (a[i].hi < b[j].hi)
            ++i;
        else
            ++j;
    }
    a.swap(r);
}

// Surface voxels of column (i, j) with inside intervals I[0, cnt): inside, and
// k - 1 or k + 1 outside the column, or voxel k of a side neighbour outside.
// column(i, j, cnt) returns a column's intervals (none outside the domain).
template <class ColumnFn, class Emit>
void column_surface(int i, int j, const Interval* I, size_t cnt, ColumnFn&& column, std::vector<Interval>& interior,
                    Emit&& emit) {
    // erode along z inside the domain: k-1 and k+1 inside
    interior.clear();
    for (size_t q = 0; q < cnt; ++q)
        if (I[q].lo + 1 <= I[q].hi - 1) interior.push_back({I[q].lo + 1, I[q].hi - 1});
    const int nb[4][2] = {{i - 1, j}, {i + 1, j}, {i, j - 1}, {i, j + 1}};
    for (int q = 0; q < 4 && !interior.empty(); ++q) {
        size_t c2;
        const Interval* J = column(nb[q][0], nb[q][1], c2);
        intersect(interior, J, c2);
    }
    // surface = I \ interior, both sorted
    size_t p = 0;
    for (size_t q = 0; q < cnt; ++q) {
        int k = I[q].lo;
        while (k <= I[q].hi) {
            while (p < interior.size() && interior[p].hi < k) ++p;
            int stop = I[q].hi;
            if (p < interior.size() && interior[p].lo <= k) {
                k = interior[p].hi + 1;
                continue;
            }
            if (p < interior.size()) stop = std::min(stop, interior[p].lo - 1);
            for (; k <= stop; ++k) emit(k);
        }
    }
}

}  // namespace

// sorted, disjoint, non-touching interval lists
static void interval_union(std::vector<Interval>& a, Interval v) {
    size_t p = 0;
    while (p < a.size() && a[p].lo <= v.lo) ++p;
    a.insert(a.begin() + (ptrdiff_t)p, v);
    std::vector<Interval> r;
    r.reserve(a.size());
    for (const Interval& x : a) {
        if (!r.empty() && x.lo <= r.back().hi + 1)
            r.back().hi = std::max(r.back().hi, x.hi);
        else
            r.push_back(x);
    }
    a.swap(r);
}

static void interval_subtract(std::vector<Interval>& a, Interval v) {
    std::vector<Interval> r;
    r.reserve(a.size() + 1);
    for (const Interval& x : a) {
        if (x.hi < v.lo || x.lo > v.hi) {
            r.push_back(x);
            continue;
        }
        if (x.lo < v.lo) r.push_back({x.lo, v.lo - 1});
        if (x.hi > v.hi) r.push_back({v.hi + 1, x.hi});
    }
    a.swap(r);
}

void EditSet::index(int n) {
    cells = std::max(1, (n + (1 << kCellShift) - 1) >> kCellShift);
    cell_off.assign((size_t)cells * cells + 1, 0);
    auto range = [&](const EditSphere& e, int& cx0, int& cx1, int& cy0, int& cy1) {
        cx0 = e.xlo >> kCellShift;
        cx1 = e.xhi >> kCellShift;
        cy0 = e.ylo >> kCellShift;
        cy1 = e.yhi >> kCellShift;
        return e.xlo <= e.xhi && e.ylo <= e.yhi;
    };
    for (const EditSphere& e : edits) {
        int cx0, cx1, cy0, cy1;
        if (!range(e, cx0, cx1, cy0, cy1)) continue;
        for (int cy = cy0; cy <= cy1; ++cy)
            for (int cx = cx0; cx <= cx1; ++cx) cell_off[(size_t)cy * cells + cx + 1]++;
    }
    for (size_t c = 0; c + 1 < cell_off.size(); ++c) cell_off[c + 1] += cell_off[c];
    cell_ids.assign(cell_off.back(), 0);
    std::vector<uint32_t> fill(cell_off.begin(), cell_off.end() - 1);
    for (size_t q = 0; q < edits.size(); ++q) {
        int cx0, cx1, cy0, cy1;
        if (!range(edits[q], cx0, cx1, cy0, cy1)) continue;
        for (int cy = cy0; cy <= cy1; ++cy)
            for (int cx = cx0; cx <= cx1; ++cx) cell_ids[fill[(size_t)cy * cells + cx]++] = (uint32_t)q;
    }
}

// Applies edit x to the inside intervals of column (i, j): GROW unions the
// source's solid runs, CLEAR subtracts them.
static void apply_edit_column(const EditSet& set, const EditSphere& x, int i, int j, int n,
                              std::vector<Interval>& colv) {
    if (i < x.xlo || i > x.xhi || j < x.ylo || j > x.yhi) return;
    auto apply = [&](Interval iv) {
        if (iv.lo > iv.hi) return;
        if (x.mode == YX_BUILD_GROW)
            interval_union(colv, iv);
        else
            interval_subtract(colv, iv);
    };
    if (x.kind == 0) {
        apply(sphere_column(Sphere{x.cx, x.cy, x.cz, x.r}, i, j, n));
        return;
    }
    const int64_t base = vol_index(x, i, j, x.vz);
    if (base < 0) return;
    const int64_t stride = (int64_t)x.sx * x.sy;
    int c = 0;
    while (c < x.sz) {
        while (c < x.sz && !set.vsolid[(size_t)(base + stride * c)]) ++c;
        if (c == x.sz) break;
        const int c0 = c;
        while (c < x.sz && set.vsolid[(size_t)(base + stride * c)]) ++c;
        const int64_t lo = (int64_t)x.vz + (int64_t)c0 * x.s, hi = (int64_t)x.vz + (int64_t)c * x.s - 1;
        apply(Interval{(int32_t)std::max<int64_t>(0, lo), (int32_t)std::min<int64_t>(n - 1, hi)});
    }
}

int add_sphere_edits(const yx_sphere_edit* e, size_t n, int depth, EditSet& set) {
    std::vector<EditSphere> out;
    const int grid = 1 << depth;
    for (size_t q = 0; q < n; ++q) {
        const yx_sphere_edit& x = e[q];
        if (x.mode != YX_BUILD_GROW && x.mode != YX_BUILD_CLEAR)
            return fail(YX_ERR_INVALID_ARG, "edit %zu: unknown mode %d", q, x.mode);
        if (x.level < 1 || x.level > depth)
            return fail(YX_ERR_INVALID_ARG, "edit %zu: level %d outside 1..%d", q, x.level, depth);
        if (!(x.radius > 0.0f) || !std::isfinite(x.radius) || !std::isfinite(x.pos[0]) || !std::isfinite(x.pos[1]) ||
            !std::isfinite(x.pos[2]))
            return fail(YX_ERR_INVALID_ARG, "edit %zu: radius must be positive and finite", q);
        const double scale = std::ldexp(1.0, depth - x.level);
        // keeps the column bounds of the ball within int range
        const double lim = 16777216.0;
        if (std::fabs((double)x.pos[0] * scale) > lim || std::fabs((double)x.pos[1] * scale) > lim ||
            std::fabs((double)x.pos[2] * scale) > lim || (double)x.radius * scale > lim)
            return fail(YX_ERR_INVALID_ARG, "edit %zu: centre or radius beyond 2^24 voxels", q);
        EditSphere s;
        s.mode = x.mode;
        s.cx = (double)x.pos[0] * scale;
        s.cy = (double)x.pos[1] * scale;
        s.cz = (double)x.pos[2] * scale;
        s.r = (double)x.radius * scale;
        s.color = x.color;
        s.xlo = std::max(0, (int)std::max(-1.0, std::floor(s.cx - s.r - 1.0)));
        s.xhi = std::min(grid - 1, (int)std::min((double)grid, std::ceil(s.cx + s.r + 1.0)));
        s.ylo = std::max(0, (int)std::max(-1.0, std::floor(s.cy - s.r - 1.0)));
        s.yhi = std::min(grid - 1, (int)std::min((double)grid, std::ceil(s.cy + s.r + 1.0)));
        s.zlo = std::max(0, (int)std::max(-1.0, std::floor(s.cz - s.r - 1.0)));
        s.zhi = std::min(grid - 1, (int)std::min((double)grid, std::ceil(s.cz + s.r + 1.0)));
        s.kind = 0;
        s.vx = s.vy = s.vz = 0;
        s.s = 1;
        s.sx = s.sy = s.sz = 0;
        s.off = 0;
        out.push_back(s);
    }
    set.edits.insert(set.edits.end(), out.begin(), out.end());
    return YX_OK;
}

int add_volume_edit(const yx_volume_source& v, int depth, EditSet& set) {
    const int grid = 1 << depth;
    if (v.mode != YX_BUILD_GROW && v.mode != YX_BUILD_CLEAR) return fail(YX_ERR_INVALID_ARG, "unknown mode %d", v.mode);
    if (v.level < 1 || v.level > depth) return fail(YX_ERR_INVALID_ARG, "level %d outside 1..%d", v.level, depth);
    if (v.kind != YX_SOURCE_ISO && v.kind != YX_SOURCE_RAW)
        return fail(YX_ERR_INVALID_ARG, "unknown source kind %d", v.kind);
    for (int a = 0; a < 3; ++a) {
        if (v.size[a] < 1 || v.size[a] > 4096) return fail(YX_ERR_INVALID_ARG, "source size %d: 1..4096", v.size[a]);
        if (v.pos[a] < -(1 << 24) || v.pos[a] > (1 << 24))
            return fail(YX_ERR_INVALID_ARG, "source position beyond 2^24 voxels");
    }
    if (!v.data) return fail(YX_ERR_INVALID_ARG, "null source data");
    const int64_t sc = (int64_t)1 << (depth - v.level);
    const size_t nv = (size_t)v.size[0] * (size_t)v.size[1] * (size_t)v.size[2];
    EditSphere e{};
    e.mode = v.mode;
    e.kind = 1;
    e.color = v.color;
    e.s = (int)sc;
    e.sx = v.size[0];
    e.sy = v.size[1];
    e.sz = v.size[2];
    const int64_t lo[3] = {v.pos[0] * sc, v.pos[1] * sc, v.pos[2] * sc};
    const int64_t hi[3] = {lo[0] + v.size[0] * sc - 1, lo[1] + v.size[1] * sc - 1, lo[2] + v.size[2] * sc - 1};
    if (std::max({hi[0], hi[1], hi[2]}) >= ((int64_t)1 << 30) || std::min({lo[0], lo[1], lo[2]}) <= -((int64_t)1 << 30))
        return fail(YX_ERR_INVALID_ARG, "source extent beyond 2^30 voxels");
    e.vx = (int)lo[0];
    e.vy = (int)lo[1];
    e.vz = (int)lo[2];
    auto clampc = [&](int64_t x) { return (int)std::min<int64_t>(grid, std::max<int64_t>(-1, x)); };
    e.xlo = std::max(0, clampc(lo[0]));
    e.xhi = std::min(grid - 1, clampc(hi[0]));
    e.ylo = std::max(0, clampc(lo[1]));
    e.yhi = std::min(grid - 1, clampc(hi[1]));
    e.zlo = std::max(0, clampc(lo[2]));
    e.zhi = std::min(grid - 1, clampc(hi[2]));
    e.off = set.vcol.size();
    const size_t old = set.vcol.size();
    try {
        set.vcol.resize(old + nv);
        set.vsolid.resize(old + nv);
    } catch (const std::bad_alloc&) {
        set.vcol.resize(old);
        set.vsolid.resize(old);
        return fail(YX_ERR_OOM, "out of host memory for a %zu-voxel source", nv);
    }
    if (v.kind == YX_SOURCE_ISO) {
        const uint8_t* d = static_cast<const uint8_t*>(v.data);
        for (size_t q = 0; q < nv; ++q) {
            const bool above = (int)d[q] >= v.iso;
            set.vsolid[old + q] = (uint8_t)(v.inside ? above : !above);
            set.vcol[old + q] = v.color;
        }
    } else {
        const uint32_t* d = static_cast<const uint32_t*>(v.data);
        for (size_t q = 0; q < nv; ++q) {
            set.vsolid[old + q] = (uint8_t)((d[q] >> 24) != 0);
            set.vcol[old + q] = d[q];
        }
    }
    set.edits.push_back(e);
    return YX_OK;
}

// The recipe's spheres in voxel units: centres uniform in the domain, radii
// uniform in [radius_min, radius_max], drawn in that order from SplitMix64(seed).
static std::vector<Sphere> recipe_spheres(const yx_scene_desc& d) {
    const int n = 1 << d.depth;
    SplitMix64 rng(d.seed);
    std::vector<Sphere> spheres(d.num_spheres);
    for (auto& s : spheres) {
        double cx = rng.uniform(), cy = rng.uniform(), cz = rng.uniform();
        double r = (double)d.radius_min + ((double)d.radius_max - (double)d.radius_min) * rng.uniform();
        s.cx = cx * n;
        s.cy = cy * n;
        s.cz = cz * n;
        s.r = r;
    }
    return spheres;
}

static int build_synthetic(const yx_scene_desc& d, yx_scene& out, const EditSet* edits = nullptr) {
    const int D = d.depth;
    const int n = 1 << D;
    const std::vector<Sphere> spheres = recipe_spheres(d);
    Perlin noise(d.seed ^ 0xD1B54A32D192ED03ull);
    Heightfield hf{&noise, d.hf_octaves, (double)d.hf_base, (double)d.hf_amplitude, n};

    if (d.heightfield)
        out.hf_top.assign((size_t)n * n, 0);
    else
        out.hf_top.clear();
    // Pass 1: merged inside-interval lists per column, row bands in parallel.
    const int band = std::max(1, std::min(n, 16));
    const int nbands = (n + band - 1) / band;
    std::vector<std::vector<uint32_t>> band_off(nbands);
    std::vector<std::vector<Interval>> band_iv(nbands);
    parallel_for(nbands, [&](int b) {
        int j0 = b * band, j1 = std::min(n, j0 + band);
        struct E {
            uint32_t col;
            Interval iv;
        };
        std::vector<E> es;
        for (const auto& s : spheres) {
            int ylo = std::max(j0, (int)std::floor(s.cy - s.r - 1.0));
            int yhi = std::min(j1 - 1, (int)std::ceil(s.cy + s.r + 1.0));
            if (ylo > yhi) continue;
            int xlo = std::max(0, (int)std::floor(s.cx - s.r - 1.0));
            int xhi = std::min(n - 1, (int)std::ceil(s.cx + s.r + 1.0));
            for (int j = ylo; j <= yhi; ++j)
                for (int i = xlo; i <= xhi; ++i) {
                    Interval iv = sphere_column(s, i, j, n);
                    if (iv.lo <= iv.hi) es.push_back({(uint32_t)((j - j0) * n + i), iv});
                }
        }
        if (d.heightfield) {
            for (int j = j0; j < j1; ++j)
                for (int i = 0; i < n; ++i) {
                    int top = hf.top(i, j);
                    out.hf_top[(size_t)j * n + i] = top;  // kept for in-place edits
                    if (top >= 0) es.push_back({(uint32_t)((j - j0) * n + i), Interval{0, top}});
                }
        }
        std::sort(es.begin(), es.end(), [](const E& a, const E& b) {
            return a.col < b.col || (a.col == b.col && a.iv.lo < b.iv.lo);
        });
        auto& off = band_off[b];
        auto& ivs = band_iv[b];
        off.assign((size_t)(j1 - j0) * n + 1, 0);
        size_t e = 0;
        std::vector<Interval> colv;
        for (uint32_t col = 0; col < (uint32_t)((j1 - j0) * n); ++col) {
            off[col] = (uint32_t)ivs.size();
            colv.clear();
            while (e < es.size() && es[e].col == col) {
                Interval iv = es[e].iv;
                if (!colv.empty() && iv.lo <= colv.back().hi + 1)
                    colv.back().hi = std::max(colv.back().hi, iv.hi);
                else
                    colv.push_back(iv);
                ++e;
            }
            // sphere edits in order: union (GROW) or difference (CLEAR)
            const int i = (int)(col % (uint32_t)n), j = j0 + (int)(col / (uint32_t)n);
            if (edits && !edits->empty())
                for (const uint32_t* e = edits->begin(i, j); e != edits->end(i, j); ++e)
                    apply_edit_column(*edits, edits->edits[*e], i, j, n, colv);
            ivs.insert(ivs.end(), colv.begin(), colv.end());
        }
        off[(size_t)(j1 - j0) * n] = (uint32_t)ivs.size();
    });
    auto column = [&](int i, int j, size_t& cnt) -> const Interval* {
        if (i < 0 || j < 0 || i >= n || j >= n) {
            cnt = 0;
            return nullptr;
        }
        int b = j / band;
        size_t col = (size_t)(j - b * band) * n + i;
        cnt = band_off[b][col + 1] - band_off[b][col];
        return band_iv[b].data() + band_off[b][col];
    };

    // Pass 2: surface voxels = inside minus (inside and all 6 neighbours inside).
    const int top_bits = std::min(D, 3);
    const int shift = 3 * (D - top_bits);
    const int nbuckets = 1 << (3 * top_bits);
    std::vector<std::vector<uint64_t>> band_codes(nbands);
    parallel_for(nbands, [&](int b) {
        int j0 = b * band, j1 = std::min(n, j0 + band);
        auto& codes = band_codes[b];
        std::vector<Interval> interior;
        for (int j = j0; j < j1; ++j)
            for (int i = 0; i < n; ++i) {
                size_t cnt;
                const Interval* I = column(i, j, cnt);
                if (cnt == 0) continue;
                column_surface(i, j, I, cnt, column, interior, [&](int k) { codes.push_back(morton3(i, j, k)); });
            }
    });

    // Bucket by the top octree levels, then sort buckets in parallel.
    std::vector<std::vector<size_t>> counts(nbands, std::vector<size_t>(nbuckets, 0));
    parallel_for(nbands, [&](int b) {
        for (uint64_t c : band_codes[b]) counts[b][c >> shift]++;
    });
    std::vector<size_t> bstart(nbuckets + 1, 0);
    for (int k = 0; k < nbuckets; ++k) {
        size_t s = 0;
        for (int b = 0; b < nbands; ++b) s += counts[b][k];
        bstart[k + 1] = bstart[k] + s;
    }
    std::vector<std::vector<size_t>> pos(nbands, std::vector<size_t>(nbuckets));
    for (int k = 0; k < nbuckets; ++k) {
        size_t p = bstart[k];
        for (int b = 0; b < nbands; ++b) {
            pos[b][k] = p;
            p += counts[b][k];
        }
    }
    std::vector<uint64_t> codes(bstart[nbuckets]);
    parallel_for(nbands, [&](int b) {
        auto& ps = pos[b];
        for (uint64_t c : band_codes[b]) codes[ps[c >> shift]++] = c;
        std::vector<uint64_t>().swap(band_codes[b]);
    });
    parallel_for(nbuckets, [&](int k) { std::sort(codes.begin() + bstart[k], codes.begin() + bstart[k + 1]); });

    build_tree_from_codes(codes, D, d.seed, out, edits && !edits->empty() ? edits : nullptr);
    return YX_OK;
}

// ---- sphere edits in place -----------------------------------------------------

namespace {

// Rewrites the nodes under the voxel box [lo, hi] of a scene whose leaves all
// sit at level `depth`: leaves inside the box become exactly `codes` (sorted
// Morton codes, with `colors`), nodes outside the box keep their ids and bytes.
// Emptied nodes are blanked and their ids reused, lowest first, before the
// array grows. Every written node's page gets `version`.
struct Splice {
    yx_scene& s;
    int depth;
    int lo[3], hi[3];
    const std::vector<uint64_t>& codes;
    const std::vector<uint32_t>& colors;
    uint64_t version;
    int64_t leaf_delta = 0;

    void mark(uint32_t id) {
        const size_t p = id / kScenePageNodes;
        if (s.page_version.size() <= p) s.page_version.resize(p + 1, 1u);
        s.page_version[p] = version;
    }
    uint32_t alloc() {
        if (!s.free_nodes.empty()) {
            std::pop_heap(s.free_nodes.begin(), s.free_nodes.end(), std::greater<uint32_t>());
            const uint32_t id = s.free_nodes.back();
            s.free_nodes.pop_back();
            return id;
        }
        const size_t id = s.nodes.size() / 10;
        if (id >= 0x7FFFFFFFull) throw std::length_error("node ids must stay below 2^31");
        if (s.nodes.size() == s.nodes.capacity()) s.nodes.reserve(s.nodes.size() + s.nodes.size() / 4 + 10240);
        s.nodes.resize((id + 1) * 10, 0u);
        return (uint32_t)id;
    }
    void release(uint32_t id) {
        uint32_t* r = &s.nodes[(size_t)id * 10];
        r[0] = 0xFF00u;
        r[1] = 0;
        for (int c = 0; c < 8; ++c) r[2 + c] = YX_EMPTY_NODE;
        mark(id);
        s.free_nodes.push_back(id);
        std::push_heap(s.free_nodes.begin(), s.free_nodes.end(), std::greater<uint32_t>());
    }
    bool meets_box(uint64_t cube, int level) const {
        const int size = 1 << (depth - level);
        const int o[3] = {(int)compact3(cube) * size, (int)compact3(cube >> 1) * size, (int)compact3(cube >> 2) * size};
        for (int a = 0; a < 3; ++a)
            if (o[a] > hi[a] || o[a] + size - 1 < lo[a]) return false;
        return true;
    }
    // node `id` (or YX_EMPTY_NODE) covering cube `cube` of `level`; the new
    // leaves in the cube are codes[b, e). Returns the node's id afterwards.
    uint32_t visit(uint32_t id, int level, uint64_t cube, size_t b, size_t e) {
        uint32_t old_row[10], row[10];
        if (id != YX_EMPTY_NODE) {
            std::memcpy(old_row, &s.nodes[(size_t)id * 10], 40);
        } else {
            old_row[0] = 0xFF00u;
            old_row[1] = 0;
            for (int c = 0; c < 8; ++c) old_row[2 + c] = YX_EMPTY_NODE;
        }
        std::memcpy(row, old_row, 40);
        const int shift = 3 * (depth - level - 1);
        size_t q = b;
        for (int c = 0; c < 8; ++c) {
            const uint64_t child = (cube << 3) | (uint64_t)c;
            const size_t cb = q;
            while (q < e && (codes[q] >> shift) == child) ++q;
            if (!meets_box(child, level + 1)) continue;
            const bool was_leaf = (old_row[0] >> c) & 1u;
            if (level == depth - 1) {
                const bool now_leaf = q > cb;
                row[2 + c] = now_leaf ? colors[cb] : YX_EMPTY_NODE;
                row[0] = (row[0] & ~((1u << c) | (1u << (8 + c)))) | (now_leaf ? 1u << c : 1u << (8 + c));
                leaf_delta += (int64_t)now_leaf - (int64_t)was_leaf;
            } else {
                if (was_leaf) throw std::logic_error("a leaf above the finest level");
                const uint32_t nc = visit(old_row[2 + c], level + 1, child, cb, q);
                row[2 + c] = nc;
                row[0] = (row[0] & ~(1u << (8 + c))) | (nc == YX_EMPTY_NODE ? 1u << (8 + c) : 0u);
            }
        }
        if ((row[0] & 0xFFu) == 0 && (row[0] & 0xFF00u) == 0xFF00u) {
            if (id != YX_EMPTY_NODE) release(id);
            return YX_EMPTY_NODE;
        }
        if (id == YX_EMPTY_NODE) id = alloc();
        else if (std::memcmp(row, old_row, 40) == 0) return id;
        std::memcpy(&s.nodes[(size_t)id * 10], row, 40);
        mark(id);
        return id;
    }
};

}  // namespace

// Applies edits[first..] one at a time to the scene in place (the scene holds
// the recipe's tree with edits[0, first) applied).
static int edit_in_place(yx_scene& s, const EditSet& set, size_t first) {
    const std::vector<EditSphere>& edits = set.edits;
    const yx_scene_desc& d = s.recipe;
    const int D = d.depth, n = 1 << D;
    const std::vector<Sphere> spheres = recipe_spheres(d);
    Perlin noise(d.seed ^ 0xD1B54A32D192ED03ull);
    Heightfield hf{&noise, d.hf_octaves, (double)d.hf_base, (double)d.hf_amplitude, n};
    constexpr int32_t kTopUnknown = INT32_MIN;
    if (d.heightfield && s.hf_top.size() != (size_t)n * n) s.hf_top.assign((size_t)n * n, kTopUnknown);
    for (size_t q = first; q < edits.size(); ++q) {
        const EditSphere& e = edits[q];
        // voxels the edit can change: its box within the domain (a sphere's
        // with a voxel of slack), then one more for the surface test of its
        // neighbours
        int lo[3] = {e.xlo, e.ylo, e.zlo}, hi[3] = {e.xhi, e.yhi, e.zhi};
        bool empty = false;
        for (int a = 0; a < 3; ++a) {
            if (lo[a] > hi[a]) empty = true;
            lo[a] = std::max(0, lo[a] - 1);
            hi[a] = std::min(n - 1, hi[a] + 1);
        }
        if (empty) continue;
        // inside intervals of the columns of the box and its rim
        const int x0 = std::max(0, lo[0] - 1), x1 = std::min(n - 1, hi[0] + 1);
        const int y0 = std::max(0, lo[1] - 1), y1 = std::min(n - 1, hi[1] + 1);
        const int nx = x1 - x0 + 1, ny = y1 - y0 + 1;
        // only voxels k in [lo[2] - 1, hi[2] + 1] of these columns are ever
        // tested, so spheres that miss that slab can be left out
        const int z0 = lo[2] - 1, z1 = hi[2] + 1;
        std::vector<Sphere> near;
        for (const Sphere& sp : spheres)
            if (sp.cx + sp.r + 1.0 >= x0 && sp.cx - sp.r - 1.0 <= x1 + 1 && sp.cy + sp.r + 1.0 >= y0 &&
                sp.cy - sp.r - 1.0 <= y1 + 1 && sp.cz + sp.r + 1.0 >= z0 && sp.cz - sp.r - 1.0 <= z1 + 1)
                near.push_back(sp);
        std::vector<std::vector<Interval>> cols((size_t)nx * ny);
        auto rows = [&](int jj) {
            const int j = y0 + jj;
            for (int i = x0; i <= x1; ++i) {
                std::vector<Interval>& colv = cols[(size_t)jj * nx + (i - x0)];
                std::vector<Interval> ivs;
                for (const Sphere& sp : near) {
                    const Interval iv = sphere_column(sp, i, j, n);
                    if (iv.lo <= iv.hi) ivs.push_back(iv);
                }
                if (d.heightfield) {
                    int32_t& top = s.hf_top[(size_t)j * n + i];
                    if (top == kTopUnknown) top = hf.top(i, j);
                    if (top >= 0) ivs.push_back(Interval{0, top});
                }
                std::sort(ivs.begin(), ivs.end(), [](const Interval& a, const Interval& b) { return a.lo < b.lo; });
                for (const Interval& iv : ivs) {
                    if (!colv.empty() && iv.lo <= colv.back().hi + 1)
                        colv.back().hi = std::max(colv.back().hi, iv.hi);
                    else
                        colv.push_back(iv);
                }
                for (const uint32_t* ek = set.begin(i, j); ek != set.end(i, j) && *ek <= q; ++ek)
                    apply_edit_column(set, edits[*ek], i, j, n, colv);
            }
        };
        // threads only pay off past a few thousand columns
        if ((size_t)nx * ny >= 4096)
            parallel_for(ny, rows);
        else
            for (int jj = 0; jj < ny; ++jj) rows(jj);
        auto column = [&](int i, int j, size_t& cnt) -> const Interval* {
            if (i < x0 || j < y0 || i > x1 || j > y1) {
                cnt = 0;
                return nullptr;
            }
            const std::vector<Interval>& v = cols[(size_t)(j - y0) * nx + (i - x0)];
            cnt = v.size();
            return v.data();
        };
        // the new leaves of the box
        std::vector<uint64_t> codes;
        std::vector<Interval> interior;
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
                size_t cnt;
                const Interval* I = column(i, j, cnt);
                if (cnt == 0) continue;
                column_surface(i, j, I, cnt, column, interior, [&](int k) {
                    if (k >= lo[2] && k <= hi[2]) codes.push_back(morton3(i, j, k));
                });
            }
        std::sort(codes.begin(), codes.end());
        std::vector<uint32_t> colors(codes.size());
        for (size_t k = 0; k < codes.size(); ++k) colors[k] = edited_color(codes[k], d.seed, &set, q + 1);
        s.version += 1;
        Splice sp{s, D, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}, codes, colors, s.version};
        s.root = sp.visit(s.root, 0, 0, 0, codes.size());
        s.num_leaves = (uint64_t)((int64_t)s.num_leaves + sp.leaf_delta);
    }
    const size_t pages = (s.nodes.size() / 10 + kScenePageNodes - 1) / kScenePageNodes;
    if (s.page_version.size() < pages) s.page_version.resize(pages, s.version);
    s.depth = s.root == YX_EMPTY_NODE ? 0 : D;
    s.depth_stale = false;
    return YX_OK;
}

// ---- GPU re-layout --------------------------------------------------------

// RendererBase::InitRayDir (cell/renderer_base.h:50-61): fwd = normalize(dir),
// right = normalize(fwd x up), up' = right x fwd, da = tan(fov/2) / w,
// du = 2 right da, dv = -2 up' da, dir0 = fwd - du w/2 - dv h/2. Integer
// factors are converted to float first, as the cg point operators do.
// forward, right and up' of InitRayDir (cell/renderer_base.h:50-55)
static void camera_vectors(const float dir[3], const float up[3], float fwd[3], float right[3], float upv[3]) {
    auto norm = [](float* v) {
        float n2 = 0.0f;
        n2 += v[0] * v[0];
        n2 += v[1] * v[1];
        n2 += v[2] * v[2];
        const float n = std::sqrt(n2);
        v[0] = v[0] / n;
        v[1] = v[1] / n;
        v[2] = v[2] / n;
    };
    auto cross = [](const float* a, const float* b, float* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    for (int i = 0; i < 3; ++i) fwd[i] = dir[i];
    norm(fwd);
    cross(fwd, up, right);
    norm(right);
    cross(right, fwd, upv);
}

void camera_basis(const float dir[3], const float up[3], float fov, int w, int h, float basis[9]) {
    float fwd[3], right[3], upv[3];
    camera_vectors(dir, up, fwd, right, upv);
    const float deg2rad = (float)(3.14159265358979323846 / (double)180.0f);
    const float da = tanf((fov / 2.0f) * deg2rad) / (float)w;
    const float wf = (float)w, hf = (float)h;
    for (int i = 0; i < 3; ++i) {
        const float du = (right[i] * 2.0f) * da;
        const float dv = (upv[i] * -2.0f) * da;
        basis[3 + i] = du;
        basis[6 + i] = dv;
        basis[i] = (fwd[i] - (du * wf) / 2.0f) - (dv * hf) / 2.0f;
    }
}

void camera_frame(const float dir[3], const float up[3], float frame[9]) {
    float fwd[3], right[3], upv[3];
    camera_vectors(dir, up, fwd, right, upv);
    for (int i = 0; i < 3; ++i) {
        frame[i] = right[i];
        frame[3 + i] = -upv[i];
        frame[6 + i] = fwd[i];
    }
}

int build_gpu_tree(const yx_scene& s, GpuTree& out, uint32_t base, bool with_lod) {
    out.recs.clear();
    out.lod.clear();
    out.depth = 0;
    out.empty = true;
    const uint32_t nn = (uint32_t)(s.nodes.size() / 10);
    if (s.root & 0x80000000u) return YX_OK;  // EmptyNode / FullNode root: nothing to hit
    if (s.root >= nn) return fail(YX_ERR_FORMAT, "root id %u out of range (%u nodes)", s.root, nn);
    std::vector<uint8_t> seen(nn, 0);
    // queue of (canonical node id, record index, level)
    std::vector<uint32_t> q_id, q_rec;
    std::vector<uint8_t> q_lvl;
    out.recs.reserve((size_t)nn * 2 * 3);
    out.recs.push_back(0);
    out.recs.push_back(0);
    q_id.push_back(s.root);
    q_rec.push_back(0);
    q_lvl.push_back(0);
    seen[s.root] = 1;
    int depth = 0, max_node_level = 0;
    out.level_records.assign(1, 1u);  // level 0: the root record
    for (size_t h = 0; h < q_id.size(); ++h) {
        const uint32_t id = q_id[h];
        const uint32_t* n = &s.nodes[(size_t)id * 10];
        const uint32_t leaf_mask = n[0] & 0xffu;
        uint32_t child_mask = 0;
        const int lvl = q_lvl[h];
        // the first node of a level: every record of that level is appended
        while ((int)out.level_records.size() <= lvl)
            out.level_records.push_back((uint32_t)(out.recs.size() / 2));
        max_node_level = std::max(max_node_level, lvl);
        for (int c = 0; c < 8; ++c) {
            if (leaf_mask & (1u << c)) continue;
            uint32_t ch = n[2 + c];
            if (ch & 0x80000000u) continue;
            if (ch >= nn) return fail(YX_ERR_FORMAT, "node %u child %d id %u out of range", id, c, ch);
            if (seen[ch]) return fail(YX_ERR_FORMAT, "node %u reached twice (not a tree)", ch);
            seen[ch] = 1;
            child_mask |= 1u << c;
        }
        if ((leaf_mask | child_mask) == 0) {
            out.recs[2 * (size_t)q_rec[h]] = 0;
            out.recs[2 * (size_t)q_rec[h] + 1] = 0;
            continue;
        }
        if (lvl + 1 > kMaxDepth)
            return fail(YX_ERR_RANGE, "tree deeper than %d levels", kMaxDepth);
        const size_t first = out.recs.size() / 2;
        if ((uint64_t)base + first + 8 > 0xFFFFFFF0ull) return fail(YX_ERR_RANGE, "more than 2^32 records");
        out.recs[2 * (size_t)q_rec[h]] = (uint32_t)(base + first);
        out.recs[2 * (size_t)q_rec[h] + 1] = child_mask | (leaf_mask << 8);
        for (int c = 0; c < 8; ++c) {
            const uint32_t bit = 1u << c;
            if (leaf_mask & bit) {
                out.recs.push_back(n[2 + c]);
                out.recs.push_back(id);
                depth = std::max(depth, lvl + 1);
            } else if (child_mask & bit) {
                const uint32_t rec = (uint32_t)(out.recs.size() / 2);
                out.recs.push_back(0);
                out.recs.push_back(0);
                q_id.push_back(n[2 + c]);
                q_rec.push_back(rec);
                q_lvl.push_back((uint8_t)(lvl + 1));
            }
        }
    }
    out.depth = depth;
    out.node_levels = max_node_level + 1;
    out.empty = (depth == 0);
    out.level_records.push_back((uint32_t)(out.recs.size() / 2));  // the deepest level's records
    if (with_lod) {
        // reaction/report/main.tex:61: a node's data is the mean of the voxels
        // of its subtree. Per channel (sum + count / 2) / count in integers
        // (0 without leaves). The queue in reverse visits children before
        // their parents.
        struct Sums {
            uint64_t c[4];
            uint64_t n;
        };
        std::vector<Sums> sums(nn);
        out.lod.assign(out.recs.size() / 2, 0u);
        for (size_t h = q_id.size(); h-- > 0;) {
            const uint32_t id = q_id[h];
            const uint32_t* n = &s.nodes[(size_t)id * 10];
            Sums acc{{0, 0, 0, 0}, 0};
            for (int c = 0; c < 8; ++c) {
                const uint32_t ch = n[2 + c];
                if (n[0] & (1u << c)) {
                    for (int k = 0; k < 4; ++k) acc.c[k] += (ch >> (8 * k)) & 0xffu;
                    acc.n += 1;
                } else if (!(ch & 0x80000000u)) {
                    const Sums& cs = sums[ch];
                    for (int k = 0; k < 4; ++k) acc.c[k] += cs.c[k];
                    acc.n += cs.n;
                }
            }
            sums[id] = acc;
            uint32_t mean = 0;
            if (acc.n > 0)
                for (int k = 0; k < 4; ++k) mean |= (uint32_t)((acc.c[k] + acc.n / 2) / acc.n) << (8 * k);
            out.lod[q_rec[h]] = mean;
        }
        // leaf records: their own colour
        for (size_t h = 0; h < q_id.size(); ++h) {
            const uint32_t rec = q_rec[h];
            const uint32_t first = out.recs[2 * (size_t)rec] - base;
            const uint32_t masks = out.recs[2 * (size_t)rec + 1];
            const uint32_t lm = (masks >> 8) & 0xffu, all = (masks & 0xffu) | lm;
            for (int c = 0, k = 0; c < 8; ++c) {
                if (!(all & (1u << c))) continue;
                if (lm & (1u << c)) out.lod[first + k] = out.recs[2 * (size_t)(first + k)];
                ++k;
            }
        }
    }
    return YX_OK;
}

}  // namespace yx

using namespace yx;

uint64_t next_scene_uid() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

// ---- C ABI: scenes --------------------------------------------------------

extern "C" int yx_scene_build_synthetic(const yx_scene_desc* desc, yx_scene** out) {
    clear_error();
    if (!desc || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (desc->depth < 1 || desc->depth > kMaxDepth)
        return fail(YX_ERR_RANGE, "depth %d outside 1..%d", desc->depth, kMaxDepth);
    if (desc->num_spheres < 0 || desc->radius_min < 0 || desc->radius_max < desc->radius_min)
        return fail(YX_ERR_INVALID_ARG, "bad sphere parameters");
    if (desc->heightfield && (desc->hf_octaves < 1 || desc->hf_octaves > 16))
        return fail(YX_ERR_INVALID_ARG, "hf_octaves must be 1..16");
    yx_scene* s = new (std::nothrow) yx_scene();
    if (!s) return fail(YX_ERR_OOM, "out of host memory");
    try {
        int rc = build_synthetic(*desc, *s);
        if (rc != YX_OK) {
            delete s;
            return rc;
        }
    } catch (const std::bad_alloc&) {
        delete s;
        return fail(YX_ERR_OOM, "out of host memory building the scene");
    }
    s->has_recipe = true;
    s->recipe = *desc;
    *out = s;
    return YX_OK;
}

// ---- edits of scenes without a recipe --------------------------------------------

namespace {

// A child slot: EMPTY, FULL (FullNode: solid, no surface), a leaf payload, or a
// node id.
struct Slot {
    enum Kind : uint8_t { kEmptySlot, kFullSlot, kLeafSlot, kNodeSlot } kind;
    uint32_t v;  // colour for kLeafSlot, id for kNodeSlot
};

inline Slot slot_of(const uint32_t* row, int c) {
    const uint32_t ch = row[2 + c];
    if ((row[0] >> c) & 1u) return Slot{Slot::kLeafSlot, ch};
    if (ch == YX_FULL_NODE) return Slot{Slot::kFullSlot, ch};
    if (ch & 0x80000000u) return Slot{Slot::kEmptySlot, YX_EMPTY_NODE};
    return Slot{Slot::kNodeSlot, ch};
}

// Voxel states of a local grid: empty, solid inside, solid on the surface.
enum : uint8_t { kVoxEmpty = 0, kVoxFull = 1, kVoxLeaf = 2 };

// Dense state and colour of the voxels of the box [g0, g1] (finest level D).
struct LocalGrid {
    int g0[3], g1[3], ext[3];
    std::vector<uint8_t> state;
    std::vector<uint32_t> color;
    size_t at(int i, int j, int k) const {
        return (size_t)(i - g0[0]) + (size_t)ext[0] * ((size_t)(j - g0[1]) + (size_t)ext[1] * (size_t)(k - g0[2]));
    }
    bool inside(int i, int j, int k) const {
        return i >= g0[0] && i <= g1[0] && j >= g0[1] && j <= g1[1] && k >= g0[2] && k <= g1[2];
    }
    void fill(const int o[3], int size, uint8_t st, uint32_t col) {
        int a0[3], a1[3];
        for (int a = 0; a < 3; ++a) {
            a0[a] = std::max(o[a], g0[a]);
            a1[a] = std::min(o[a] + size - 1, g1[a]);
            if (a0[a] > a1[a]) return;
        }
        for (int k = a0[2]; k <= a1[2]; ++k)
            for (int j = a0[1]; j <= a1[1]; ++j)
                for (int i = a0[0]; i <= a1[0]; ++i) {
                    const size_t x = at(i, j, k);
                    state[x] = st;
                    color[x] = col;
                }
    }
};

// Re-encodes the voxels of box [lo, hi] of a tree whose leaves are at most at
// level `depth`: rewrites only the nodes whose cubes meet the box.
struct TreeSplice {
    Splice base;  // page marks, allocation and blanking
    const LocalGrid& g;
    const std::vector<uint8_t>& nstate;   // new state per grid voxel
    const std::vector<uint32_t>& ncolor;  // and colour

    int depth() const { return base.depth; }
    bool meets(const int o[3], int size) const {
        for (int a = 0; a < 3; ++a)
            if (o[a] > base.hi[a] || o[a] + size - 1 < base.lo[a]) return false;
        return true;
    }
    static bool is_leaf(const Slot& x) { return x.kind == Slot::kLeafSlot; }
    void release_tree(uint32_t id) {
        uint32_t row[10];
        std::memcpy(row, &base.s.nodes[(size_t)id * 10], 40);
        for (int c = 0; c < 8; ++c) {
            const Slot k = slot_of(row, c);
            if (k.kind == Slot::kNodeSlot) release_tree(k.v);
            if (k.kind == Slot::kLeafSlot) base.leaf_delta -= 1;
        }
        base.release(id);
    }

    // slot `x` covering the cube at origin o of `level` (its children are
    // voxels when level == depth - 1); returns the slot that replaces it
    Slot visit(Slot x, int level, const int o[3]) {
        const int size = 1 << (depth() - level);
        if (!meets(o, size)) return x;
        if (level == depth()) {
            if (x.kind == Slot::kNodeSlot) release_tree(x.v);  // an empty chain below the leaf level
            const size_t at = g.at(o[0], o[1], o[2]);
            const uint8_t st = nstate[at];
            const Slot y = st == kVoxLeaf ? Slot{Slot::kLeafSlot, ncolor[at]}
                                          : (st == kVoxFull ? Slot{Slot::kFullSlot, YX_FULL_NODE}
                                                            : Slot{Slot::kEmptySlot, YX_EMPTY_NODE});
            base.leaf_delta += (int64_t)is_leaf(y) - (int64_t)is_leaf(x);
            return y;
        }
        uint32_t old_row[10], row[10];
        Slot kids[8];
        if (x.kind == Slot::kNodeSlot) {
            std::memcpy(old_row, &base.s.nodes[(size_t)x.v * 10], 40);
            for (int c = 0; c < 8; ++c) kids[c] = slot_of(old_row, c);
        } else {
            // an empty, full or leaf cube seen as a node of 8 such children
            for (int c = 0; c < 8; ++c) kids[c] = x;
            if (x.kind == Slot::kLeafSlot) base.leaf_delta += 7;
        }
        const int half = size >> 1;
        bool all_empty = true, all_full = true;
        for (int c = 0; c < 8; ++c) {
            const int co[3] = {o[0] + (c & 1) * half, o[1] + ((c >> 1) & 1) * half, o[2] + ((c >> 2) & 1) * half};
            kids[c] = visit(kids[c], level + 1, co);
            all_empty = all_empty && kids[c].kind == Slot::kEmptySlot;
            all_full = all_full && kids[c].kind == Slot::kFullSlot;
        }
        if (all_empty || all_full) {
            if (x.kind == Slot::kNodeSlot) base.release(x.v);
            return all_empty ? Slot{Slot::kEmptySlot, YX_EMPTY_NODE} : Slot{Slot::kFullSlot, YX_FULL_NODE};
        }
        row[0] = 0;
        row[1] = x.kind == Slot::kNodeSlot ? old_row[1] : 0u;
        for (int c = 0; c < 8; ++c) {
            row[2 + c] = kids[c].v;
            if (kids[c].kind == Slot::kLeafSlot) row[0] |= 1u << c;
            if (kids[c].kind == Slot::kEmptySlot || kids[c].kind == Slot::kFullSlot) row[0] |= 1u << (8 + c);
        }
        uint32_t id;
        if (x.kind == Slot::kNodeSlot) {
            id = x.v;
            row[0] |= old_row[0] & ~0xFFFFu;  // keep the flag bits above the masks
            if (std::memcmp(row, old_row, 40) == 0) return x;
        } else {
            id = base.alloc();
        }
        std::memcpy(&base.s.nodes[(size_t)id * 10], row, 40);
        base.mark(id);
        return Slot{Slot::kNodeSlot, id};
    }
};

// Fills g with the solid of the subtree in slot x (cube at origin o of
// `level`): leaf cubes and FullNode cubes are solid.
void decode_into(const yx_scene& s, LocalGrid& g, Slot x, int level, int depth, const int o[3]) {
    const int size = 1 << (depth - level);
    for (int a = 0; a < 3; ++a)
        if (o[a] > g.g1[a] || o[a] + size - 1 < g.g0[a]) return;
    if (x.kind == Slot::kEmptySlot) return;
    if (x.kind == Slot::kFullSlot) return g.fill(o, size, kVoxFull, 0u);
    if (x.kind == Slot::kLeafSlot) return g.fill(o, size, kVoxLeaf, x.v);
    if (level >= depth) return;  // a chain of empty nodes below the deepest leaf
    const uint32_t* row = &s.nodes[(size_t)x.v * 10];
    const int half = size >> 1;
    for (int c = 0; c < 8; ++c) {
        const int co[3] = {o[0] + (c & 1) * half, o[1] + ((c >> 1) & 1) * half, o[2] + ((c >> 2) & 1) * half};
        decode_into(s, g, slot_of(row, c), level + 1, depth, co);
    }
}

}  // namespace

// Applies edits to a scene without a recipe in place. Its solid is its leaves
// (a leaf slot covers its child cube) and its FullNode cubes; an edit unions
// (GROW) or removes (CLEAR) the source's solid voxels, and the voxels of the
// edit's box plus one are encoded again at the leaf level `depth`: surface
// voxels (solid, with a 6-neighbour empty or outside the domain) as leaves,
// other solid voxels as FullNode (whole FullNode cubes collapse), the rest
// empty. A leaf keeps its colour; a grown voxel takes the source's; a solid
// voxel that a CLEAR exposes takes the CLEAR's colour.
static int edit_tree_in_place(yx_scene& s, const EditSet& set, int depth) {
    const int n = 1 << depth;
    for (size_t q = 0; q < set.edits.size(); ++q) {
        const EditSphere& e = set.edits[q];
        int lo[3] = {e.xlo, e.ylo, e.zlo}, hi[3] = {e.xhi, e.yhi, e.zhi};
        bool empty = false;
        for (int a = 0; a < 3; ++a) {
            if (lo[a] > hi[a]) empty = true;
            lo[a] = std::max(0, lo[a] - 1);
            hi[a] = std::min(n - 1, hi[a] + 1);
        }
        if (empty) continue;
        LocalGrid g;
        size_t cells = 1;
        for (int a = 0; a < 3; ++a) {
            g.g0[a] = std::max(0, lo[a] - 1);
            g.g1[a] = std::min(n - 1, hi[a] + 1);
            g.ext[a] = g.g1[a] - g.g0[a] + 1;
            cells *= (size_t)g.ext[a];
        }
        if (cells > ((size_t)1 << 28))
            return fail(YX_ERR_RANGE, "edit %zu covers %zu voxels; a scene without a recipe takes at most 2^28 per edit",
                        q, cells);
        g.state.assign(cells, kVoxEmpty);
        g.color.assign(cells, 0u);
        const int o0[3] = {0, 0, 0};
        const Slot root = s.root == YX_FULL_NODE ? Slot{Slot::kFullSlot, YX_FULL_NODE}
                                                 : ((s.root & 0x80000000u) ? Slot{Slot::kEmptySlot, YX_EMPTY_NODE}
                                                                           : Slot{Slot::kNodeSlot, s.root});
        decode_into(s, g, root, 0, depth, o0);
        // the edit, then the new encoding of the box
        std::vector<uint8_t> solid(cells), nstate(cells, kVoxEmpty);
        std::vector<uint32_t> ncolor(g.color);
        for (int k = g.g0[2]; k <= g.g1[2]; ++k)
            for (int j = g.g0[1]; j <= g.g1[1]; ++j)
                for (int i = g.g0[0]; i <= g.g1[0]; ++i) {
                    const size_t x = g.at(i, j, k);
                    bool sol = g.state[x] != kVoxEmpty;
                    uint32_t col = 0;
                    if (i >= e.xlo && i <= e.xhi && j >= e.ylo && j <= e.yhi && k >= e.zlo && k <= e.zhi &&
                        in_edit(set, e, (uint32_t)i, (uint32_t)j, (uint32_t)k, &col)) {
                        sol = e.mode == YX_BUILD_GROW;
                        if (sol) ncolor[x] = col;
                        else ncolor[x] = 0;
                        if (sol) g.state[x] = kVoxLeaf;  // grown: coloured
                    }
                    solid[x] = sol ? 1 : 0;
                }
        auto solid_at = [&](int i, int j, int k) -> bool {
            if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n) return false;
            return solid[g.at(i, j, k)] != 0;
        };
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    const size_t x = g.at(i, j, k);
                    if (!solid[x]) continue;
                    const bool surf = !solid_at(i - 1, j, k) || !solid_at(i + 1, j, k) || !solid_at(i, j - 1, k) ||
                                      !solid_at(i, j + 1, k) || !solid_at(i, j, k - 1) || !solid_at(i, j, k + 1);
                    if (!surf) {
                        nstate[x] = kVoxFull;
                    } else {
                        nstate[x] = kVoxLeaf;
                        if (g.state[x] != kVoxLeaf) ncolor[x] = e.color;  // exposed by a CLEAR
                    }
                }
        s.version += 1;
        static const std::vector<uint64_t> no_codes;
        static const std::vector<uint32_t> no_colors;
        TreeSplice ts{Splice{s, depth, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}, no_codes, no_colors, s.version},
                      g, nstate, ncolor};
        const Slot r = ts.visit(root, 0, o0);
        if (r.kind == Slot::kNodeSlot) {
            s.root = r.v;
        } else if (r.kind == Slot::kLeafSlot) {
            // a root that is one leaf cube: keep a node above it
            throw std::logic_error("a leaf root");
        } else {
            s.root = r.v;
        }
        s.num_leaves = (uint64_t)((int64_t)s.num_leaves + ts.base.leaf_delta);
    }
    const size_t pages = (s.nodes.size() / 10 + kScenePageNodes - 1) / kScenePageNodes;
    if (s.page_version.size() < pages) s.page_version.resize(pages, s.version);
    return YX_OK;
}

static int tree_depth(yx_scene* s, int* depth);

// ---- solid interior of a surface-only tree ---------------------------------------

namespace {

struct BitVol {
    int n = 0;
    std::vector<uint64_t> w;
    void init(int side) {
        n = side;
        w.assign(((size_t)side * side * side + 63) / 64, 0ull);
    }
    size_t at(int x, int y, int z) const { return (size_t)x + (size_t)n * ((size_t)y + (size_t)n * (size_t)z); }
    bool get(size_t i) const { return (w[i >> 6] >> (i & 63)) & 1ull; }
    void set(size_t i) { w[i >> 6] |= 1ull << (i & 63); }
    void fill(const int o[3], int size) {
        for (int z = o[2]; z < o[2] + size; ++z)
            for (int y = o[1]; y < o[1] + size; ++y)
                for (int x = o[0]; x < o[0] + size; ++x) set(at(x, y, z));
    }
};

void mark_solid(const yx_scene& s, BitVol& b, Slot x, int level, int depth, const int o[3]) {
    const int size = 1 << (depth - level);
    if (x.kind == Slot::kEmptySlot) return;
    if (x.kind == Slot::kFullSlot || x.kind == Slot::kLeafSlot) return b.fill(o, size);
    if (level >= depth) return;  // an empty chain below the deepest leaf
    const uint32_t* row = &s.nodes[(size_t)x.v * 10];
    const int half = size >> 1;
    for (int c = 0; c < 8; ++c) {
        const int co[3] = {o[0] + (c & 1) * half, o[1] + ((c >> 1) & 1) * half, o[2] + ((c >> 2) & 1) * half};
        mark_solid(s, b, slot_of(row, c), level + 1, depth, co);
    }
}

// 6-connected flood of the empty voxels reachable from the domain's faces,
// by x-spans: a span is filled whole, then the rows y +- 1 and z +- 1 over it
// are scanned for unfilled empty runs. Rows of 64 voxels or more are whole
// 64-bit words, and spans are found and filled a word at a time.
void flood_outside(const BitVol& solid, BitVol& out) {
    const int n = solid.n;
    struct Span {
        int x0, x1, y, z;
    };
    std::vector<Span> stack;
    const bool words = n % 64 == 0;
    const int wpr = n / 64;  // words per row
    auto row_word = [&](int y, int z) { return ((size_t)y + (size_t)n * (size_t)z) * (size_t)wpr; };
    auto open_bits = [&](size_t w) { return ~solid.w[w] & ~out.w[w]; };
    auto open = [&](int x, int y, int z) {
        const size_t i = solid.at(x, y, z);
        return !solid.get(i) && !out.get(i);
    };
    // fills the open run through x of row (y, z), pushes it as a span
    auto fill_from = [&](int x, int y, int z) {
        int x0 = x, x1 = x;
        if (!words) {
            while (x0 > 0 && open(x0 - 1, y, z)) --x0;
            while (x1 < n - 1 && open(x1 + 1, y, z)) ++x1;
            for (int k = x0; k <= x1; ++k) out.set(out.at(k, y, z));
        } else {
            const size_t base = row_word(y, z);
            // left end: highest closed bit below x, word by word
            int w = x >> 6, b = x & 63;
            for (;;) {
                const uint64_t below = b >= 64 ? ~0ull : (b ? ((1ull << b) - 1ull) : 0ull);
                const uint64_t closed = ~open_bits(base + w) & below;
                if (closed) {
                    x0 = (w << 6) + (63 - __builtin_clzll(closed)) + 1;
                    break;
                }
                if (w == 0) {
                    x0 = 0;
                    break;
                }
                --w;
                b = 64;
                if (open_bits(base + w) == ~0ull) continue;  // whole word open: keep going
                const uint64_t c2 = ~open_bits(base + w);
                x0 = (w << 6) + (63 - __builtin_clzll(c2)) + 1;
                break;
            }
            // right end: lowest closed bit above x
            w = x >> 6;
            b = x & 63;
            for (;;) {
                const uint64_t above = b == 63 ? 0ull : (~0ull << (b + 1));
                const uint64_t closed = ~open_bits(base + w) & above;
                if (closed) {
                    x1 = (w << 6) + __builtin_ctzll(closed) - 1;
                    break;
                }
                if (w == wpr - 1) {
                    x1 = n - 1;
                    break;
                }
                ++w;
                b = -1;
                if (open_bits(base + w) == ~0ull) {
                    if (w == wpr - 1) {
                        x1 = n - 1;
                        break;
                    }
                    b = 63;  // whole word open: test the next one
                    continue;
                }
                const uint64_t c2 = ~open_bits(base + w);
                x1 = (w << 6) + __builtin_ctzll(c2) - 1;
                break;
            }
            for (int ww = x0 >> 6; ww <= (x1 >> 6); ++ww) {
                const int lo = std::max(x0, ww << 6) - (ww << 6), hi = std::min(x1, (ww << 6) + 63) - (ww << 6);
                const uint64_t m = (hi == 63 ? ~0ull : ((1ull << (hi + 1)) - 1ull)) & (~0ull << lo);
                out.w[base + ww] |= m;
            }
        }
        stack.push_back(Span{x0, x1, y, z});
    };
    auto scan_row = [&](int x0, int x1, int y, int z) {
        if (y < 0 || z < 0 || y >= n || z >= n) return;
        if (!words) {
            for (int x = x0; x <= x1; ++x)
                if (open(x, y, z)) fill_from(x, y, z);
            return;
        }
        const size_t base = row_word(y, z);
        for (int ww = x0 >> 6; ww <= (x1 >> 6); ++ww) {
            const int lo = std::max(x0, ww << 6) - (ww << 6), hi = std::min(x1, (ww << 6) + 63) - (ww << 6);
            const uint64_t range = (hi == 63 ? ~0ull : ((1ull << (hi + 1)) - 1ull)) & (~0ull << lo);
            uint64_t m;
            while ((m = open_bits(base + ww) & range) != 0) fill_from((ww << 6) + __builtin_ctzll(m), y, z);
        }
    };
    auto drain = [&]() {
        while (!stack.empty()) {
            const Span sp = stack.back();
            stack.pop_back();
            scan_row(sp.x0, sp.x1, sp.y - 1, sp.z);
            scan_row(sp.x0, sp.x1, sp.y + 1, sp.z);
            scan_row(sp.x0, sp.x1, sp.y, sp.z - 1);
            scan_row(sp.x0, sp.x1, sp.y, sp.z + 1);
        }
    };
    // seeds: every open voxel on the six faces (x faces through the y, z rows)
    for (int a = 0; a < n; ++a)
        for (int b2 = 0; b2 < n; ++b2) {
            if (open(0, a, b2)) {
                fill_from(0, a, b2);
                drain();
            }
            if (open(n - 1, a, b2)) {
                fill_from(n - 1, a, b2);
                drain();
            }
        }
    for (int a = 0; a < n; ++a) {
        scan_row(0, n - 1, 0, a);
        drain();
        scan_row(0, n - 1, n - 1, a);
        drain();
        scan_row(0, n - 1, a, 0);
        drain();
        scan_row(0, n - 1, a, n - 1);
        drain();
    }
}

// Turns the empty slots whose cube the flood did not reach into FullNode.
// An empty cube is one connected empty region, so one voxel decides it.
uint64_t mark_enclosed(yx_scene& s, const BitVol& out, Splice& sp, uint32_t id, int level, int depth, const int o[3]) {
    uint32_t* row = &s.nodes[(size_t)id * 10];
    const int half = (1 << (depth - level)) >> 1;
    uint64_t changed = 0;
    bool dirty = false;
    for (int c = 0; c < 8; ++c) {
        const int co[3] = {o[0] + (c & 1) * half, o[1] + ((c >> 1) & 1) * half, o[2] + ((c >> 2) & 1) * half};
        const Slot k = slot_of(row, c);
        if (k.kind == Slot::kEmptySlot) {
            if (!out.get(out.at(co[0], co[1], co[2]))) {
                row[2 + c] = YX_FULL_NODE;
                row[0] |= 1u << (8 + c);
                dirty = true;
                ++changed;
            }
        } else if (k.kind == Slot::kNodeSlot && level + 1 < depth) {
            changed += mark_enclosed(s, out, sp, k.v, level + 1, depth, co);
            row = &s.nodes[(size_t)id * 10];
        }
    }
    if (dirty) sp.mark(id);
    return changed;
}

}  // namespace

extern "C" int yx_scene_fill_interior(yx_scene* s, uint64_t* filled) {
    clear_error();
    if (!s) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (s->has_recipe)
        return fail(YX_ERR_INVALID_ARG, "a synthetic scene keeps its solid in its recipe; fill only trees without one");
    int depth = 0;
    int rc = tree_depth(s, &depth);
    if (rc != YX_OK) return rc;
    if (depth > 11) return fail(YX_ERR_RANGE, "depth %d: the dense flood handles up to 2048^3 voxels", depth);
    uint64_t count = 0;
    try {
        BitVol solid, out;
        solid.init(1 << depth);
        out.init(1 << depth);
        const int o0[3] = {0, 0, 0};
        const Slot root = s->root == YX_FULL_NODE ? Slot{Slot::kFullSlot, YX_FULL_NODE}
                                                  : ((s->root & 0x80000000u) ? Slot{Slot::kEmptySlot, YX_EMPTY_NODE}
                                                                             : Slot{Slot::kNodeSlot, s->root});
        mark_solid(*s, solid, root, 0, depth, o0);
        flood_outside(solid, out);
        if (root.kind == Slot::kNodeSlot) {
            s->version += 1;
            static const std::vector<uint64_t> no_codes;
            static const std::vector<uint32_t> no_colors;
            Splice sp{*s, depth, {0, 0, 0}, {0, 0, 0}, no_codes, no_colors, s->version};
            count = mark_enclosed(*s, out, sp, root.v, 0, depth, o0);
        }
    } catch (const std::bad_alloc&) {
        return fail(YX_ERR_OOM, "out of host memory for the %d-level flood", depth);
    }
    if (filled) *filled = count;
    return YX_OK;
}

// Applies the edits appended to s->edits from index `first` (volume data from
// `vfirst`) by method `how`; on failure they are taken off again.
static int apply_new_edits(yx_scene* s, size_t first, size_t vfirst, int32_t how, void* stream) {
    EditSet& es = s->edits;
    auto rollback = [&](int code) {
        es.edits.resize(first);
        es.vcol.resize(vfirst);
        es.vsolid.resize(vfirst);
        es.index(1 << s->recipe.depth);
        return code;
    };
    try {
        es.index(1 << s->recipe.depth);
    } catch (const std::bad_alloc&) {
        return rollback(fail(YX_ERR_OOM, "out of host memory"));
    }
    int rc;
    if (how == YX_RANGE_IN_PLACE) {
        try {
            rc = edit_in_place(*s, es, first);
        } catch (const std::bad_alloc&) {
            return fail(YX_ERR_OOM, "out of host memory editing the scene (the scene is partly edited)");
        } catch (const std::exception& x) {
            return fail(YX_ERR_RANGE, "%s (the scene is partly edited)", x.what());
        }
        return rc;
    }
    yx_scene t;
    try {
        rc = how == YX_RANGE_REBUILD_GPU ? gpu_build_synthetic(s->recipe, es, t, stream)
                                         : build_synthetic(s->recipe, t, &es);
    } catch (const std::bad_alloc&) {
        return rollback(fail(YX_ERR_OOM, "out of host memory rebuilding the scene"));
    }
    if (rc != YX_OK) return rollback(rc);
    // new version on the pages whose bytes changed (and on new pages)
    const size_t old_n = s->nodes.size() / 10, new_n = t.nodes.size() / 10;
    const size_t pages = (new_n + kScenePageNodes - 1) / kScenePageNodes;
    s->version += 1;
    try {
        s->page_version.resize(std::max(pages, s->page_version.size()), 1u);
    } catch (const std::bad_alloc&) {
        return rollback(fail(YX_ERR_OOM, "out of host memory"));
    }
    for (size_t p = 0; p < pages; ++p) {
        const size_t n0 = p * kScenePageNodes, n1 = std::min(new_n, n0 + kScenePageNodes);
        const bool same = n1 <= old_n && std::memcmp(&s->nodes[n0 * 10], &t.nodes[n0 * 10], (n1 - n0) * 40) == 0;
        if (!same) s->page_version[p] = s->version;
    }
    s->page_version.resize(pages);
    s->nodes.swap(t.nodes);
    s->root = t.root;
    s->num_leaves = t.num_leaves;
    s->depth = t.depth;
    s->depth_stale = false;
    s->free_nodes.clear();
    if (!t.hf_top.empty()) s->hf_top.swap(t.hf_top);
    return YX_OK;
}

static int check_range_call(yx_scene* s, int32_t how) {
    if (how != YX_RANGE_IN_PLACE && how != YX_RANGE_REBUILD && how != YX_RANGE_REBUILD_GPU)
        return fail(YX_ERR_INVALID_ARG, "unknown build-range method %d", how);
    if (!s->has_recipe && how != YX_RANGE_IN_PLACE)
        return fail(YX_ERR_INVALID_ARG, "a scene without a synthetic recipe (loaded, or edited node by node) is "
                                        "edited in place only");
    return YX_OK;
}

// Depth of the leaf level of a scene without a recipe (validating it if it
// was edited node by node since).
static int tree_depth(yx_scene* s, int* depth) {
    if (s->depth_stale) {
        yx_scene_info info;
        const int rc = yx_scene_get_info(s, &info);
        if (rc != YX_OK) return rc;
    }
    if (s->depth < 1) return fail(YX_ERR_INVALID_ARG, "the scene has no leaves, so no leaf level to edit at");
    *depth = s->depth;
    return YX_OK;
}

// Sphere or volume edits on a scene without a recipe (in place).
template <class Add>
static int edit_tree(yx_scene* s, Add&& add) {
    int depth = 0;
    int rc = tree_depth(s, &depth);
    if (rc != YX_OK) return rc;
    EditSet es;
    try {
        rc = add(depth, es);
        if (rc != YX_OK) return rc;
        rc = edit_tree_in_place(*s, es, depth);
    } catch (const std::bad_alloc&) {
        return fail(YX_ERR_OOM, "out of host memory editing the scene (the scene is partly edited)");
    } catch (const std::exception& x) {
        return fail(YX_ERR_RANGE, "%s (the scene is partly edited)", x.what());
    }
    return rc;
}

extern "C" int yx_scene_build_range(yx_scene* s, const yx_sphere_edit* edits, uint32_t count, int32_t how,
                                    void* stream) {
    clear_error();
    if (!s || (!edits && count)) return fail(YX_ERR_INVALID_ARG, "null argument");
    int rc = check_range_call(s, how);
    if (rc != YX_OK || count == 0) return rc;
    if (!s->has_recipe)
        return edit_tree(s, [&](int depth, EditSet& es) { return add_sphere_edits(edits, count, depth, es); });
    const size_t first = s->edits.edits.size(), vfirst = s->edits.vcol.size();
    try {
        rc = add_sphere_edits(edits, count, s->recipe.depth, s->edits);
    } catch (const std::bad_alloc&) {
        s->edits.edits.resize(first);
        return fail(YX_ERR_OOM, "out of host memory");
    }
    if (rc != YX_OK) return rc;
    return apply_new_edits(s, first, vfirst, how, stream);
}

extern "C" int yx_scene_build_range_volume(yx_scene* s, const yx_volume_source* src, int32_t how, void* stream) {
    clear_error();
    if (!s || !src) return fail(YX_ERR_INVALID_ARG, "null argument");
    int rc = check_range_call(s, how);
    if (rc != YX_OK) return rc;
    if (!s->has_recipe) return edit_tree(s, [&](int depth, EditSet& es) { return add_volume_edit(*src, depth, es); });
    const size_t first = s->edits.edits.size(), vfirst = s->edits.vcol.size();
    try {
        rc = add_volume_edit(*src, s->recipe.depth, s->edits);
    } catch (const std::bad_alloc&) {
        s->edits.edits.resize(first);
        s->edits.vcol.resize(vfirst);
        s->edits.vsolid.resize(vfirst);
        return fail(YX_ERR_OOM, "out of host memory");
    }
    if (rc != YX_OK) return rc;
    return apply_new_edits(s, first, vfirst, how, stream);
}

extern "C" int yx_scene_from_nodes(const uint32_t* nodes, uint32_t num_nodes, uint32_t root,
                                   yx_scene** out) {
    clear_error();
    if (!out || (!nodes && num_nodes)) return fail(YX_ERR_INVALID_ARG, "null argument");
    yx_scene* s = new (std::nothrow) yx_scene();
    if (!s) return fail(YX_ERR_OOM, "out of host memory");
    try {
        s->nodes.assign(nodes, nodes + (size_t)num_nodes * 10);
    } catch (const std::bad_alloc&) {
        delete s;
        return fail(YX_ERR_OOM, "out of host memory");
    }
    s->root = root;
    GpuTree t;  // validates ids, computes depth and leaf count
    int rc = build_gpu_tree(*s, t);
    if (rc != YX_OK) {
        delete s;
        return rc;
    }
    s->depth = t.depth;
    uint64_t leaves = 0;
    for (uint32_t i = 0; i < num_nodes; ++i) leaves += __builtin_popcount(nodes[(size_t)i * 10] & 0xffu);
    s->num_leaves = leaves;
    *out = s;
    return YX_OK;
}

extern "C" int yx_scene_load_vox(const char* path, yx_scene** out) {
    clear_error();
    if (!path || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(YX_ERR_IO, "cannot open %s", path);
    uint32_t hdr[4];
    if (fread(hdr, 4, 4, f) != 4) {
        fclose(f);
        return fail(YX_ERR_IO, "%s: short header", path);
    }
    const uint32_t size = hdr[3];
    // the header's node count must fit the file before anything is allocated:
    // a truncated or foreign file cannot ask for gigabytes
    long here = ftell(f), end = -1;
    if (here >= 0 && fseek(f, 0, SEEK_END) == 0) end = ftell(f);
    if (end < 0 || fseek(f, here, SEEK_SET) != 0) {
        fclose(f);
        return fail(YX_ERR_IO, "%s: cannot seek", path);
    }
    if ((uint64_t)(end - here) < (uint64_t)size * 40u) {
        fclose(f);
        return fail(YX_ERR_IO, "%s: short file (%ld bytes of nodes, the header names %u nodes)", path,
                    end - here, size);
    }
    std::vector<uint32_t> nodes;
    try {
        nodes.resize((size_t)size * 10);
    } catch (const std::bad_alloc&) {
        fclose(f);
        return fail(YX_ERR_OOM, "%s: %u nodes do not fit in host memory", path, size);
    }
    size_t got = size ? fread(nodes.data(), 40, size, f) : 0;
    fclose(f);
    if (got != size) return fail(YX_ERR_IO, "%s: short file (%zu of %u nodes)", path, got, size);
    return yx_scene_from_nodes(nodes.data(), size, hdr[0], out);
}

extern "C" int yx_scene_save_vox(const yx_scene* s, const char* path) {
    clear_error();
    if (!s || !path) return fail(YX_ERR_INVALID_ARG, "null argument");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(YX_ERR_IO, "cannot create %s", path);
    uint32_t hdr[4] = {s->root, 0, 0, (uint32_t)(s->nodes.size() / 10)};
    bool ok = fwrite(hdr, 4, 4, f) == 4;
    if (ok && hdr[3]) ok = fwrite(s->nodes.data(), 40, hdr[3], f) == hdr[3];
    ok = (fclose(f) == 0) && ok;
    return ok ? YX_OK : fail(YX_ERR_IO, "write to %s failed", path);
}

extern "C" int yx_scene_write_nodes(yx_scene* s, uint32_t first, const uint32_t* nodes, uint32_t count) {
    clear_error();
    if (!s || (!nodes && count)) return fail(YX_ERR_INVALID_ARG, "null argument");
    const size_t size = s->nodes.size() / 10;
    if (first > size) return fail(YX_ERR_RANGE, "first node %u beyond the end (%zu nodes)", first, size);
    const uint64_t end = (uint64_t)first + count;
    if (end >= 0x80000000ull) return fail(YX_ERR_RANGE, "node ids must stay below 2^31");
    if (count == 0) return YX_OK;
    try {
        if (end > size) s->nodes.resize((size_t)end * 10, 0u);
        const size_t pages = (size_t)((end + kScenePageNodes - 1) / kScenePageNodes);
        if (s->page_version.size() < pages) s->page_version.resize(pages, 1u);
    } catch (const std::bad_alloc&) {
        return fail(YX_ERR_OOM, "out of host memory growing the scene");
    }
    uint64_t leaves = s->num_leaves;
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t* dst = &s->nodes[((size_t)first + i) * 10];
        if ((size_t)first + i < size) leaves -= (uint64_t)__builtin_popcount(dst[0] & 0xffu);
        std::memcpy(dst, nodes + (size_t)i * 10, 40);
        leaves += (uint64_t)__builtin_popcount(dst[0] & 0xffu);
    }
    s->num_leaves = leaves;
    s->version += 1;
    for (uint64_t p = first / kScenePageNodes; p <= (end - 1) / kScenePageNodes; ++p)
        s->page_version[(size_t)p] = s->version;
    s->depth_stale = true;
    s->has_recipe = false;  // the nodes no longer follow the recipe
    s->free_nodes.clear();
    return YX_OK;
}

extern "C" int yx_scene_set_root(yx_scene* s, uint32_t root) {
    clear_error();
    if (!s) return fail(YX_ERR_INVALID_ARG, "null argument");
    s->root = root;
    s->version += 1;
    s->depth_stale = true;
    s->has_recipe = false;
    return YX_OK;
}

extern "C" int yx_scene_get_info(const yx_scene* s, yx_scene_info* info) {
    clear_error();
    if (!s || !info) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (s->depth_stale) {
        // edited since the depth was known: validate and measure again
        GpuTree t;
        int rc;
        try {
            rc = build_gpu_tree(*s, t);
        } catch (const std::bad_alloc&) {
            return fail(YX_ERR_OOM, "out of host memory");
        }
        if (rc != YX_OK) return rc;
        s->depth = t.depth;
        s->depth_stale = false;
    }
    info->root = s->root;
    info->num_nodes = (uint32_t)(s->nodes.size() / 10);
    info->num_leaves = s->num_leaves;
    info->depth = s->depth;
    return YX_OK;
}

extern "C" int yx_scene_nodes(const yx_scene* s, const uint32_t** nodes, uint32_t* num_nodes,
                              uint32_t* root) {
    clear_error();
    if (!s || !nodes || !num_nodes || !root) return fail(YX_ERR_INVALID_ARG, "null argument");
    *nodes = s->nodes.data();
    *num_nodes = (uint32_t)(s->nodes.size() / 10);
    *root = s->root;
    return YX_OK;
}

// Leaf payload decoding of a dumped frame (the host half of
// SVORenderer::DumpTraceData, demo/SVORenderer.cpp:158-192: UnpackVoxData /
// UnpackColor / UnpackNormal, whose bodies the reference does not ship; the
// colour16 + normal16 layout is the one the kernels shade with).
static uint32_t host_rgba_from_565(uint32_t c) {
    const uint32_t r5 = (c >> 11) & 31u, g6 = (c >> 5) & 63u, b5 = c & 31u;
    const uint32_t r = (r5 << 3) | (r5 >> 2), g = (g6 << 2) | (g6 >> 4), b = (b5 << 3) | (b5 >> 2);
    return r | (g << 8) | (b << 16) | 0xff000000u;
}

static void host_normal_from_16(uint32_t q, float* n) {
    int u = (int)((q >> 7) & 63u), v = (int)(q & 127u);
    if (u + v > 126) {
        u = 127 - u;
        v = 127 - v;
    }
    const float x = (float)u, y = (float)v, z = (float)(126 - u - v);
    const float len = std::sqrt(x * x + y * y + z * z);
    n[0] = (q & 0x8000u) ? -(x / len) : x / len;
    n[1] = (q & 0x4000u) ? -(y / len) : y / len;
    n[2] = (q & 0x2000u) ? -(z / len) : z / len;
}

extern "C" int yx_scene_decode_hits(const yx_scene* s, const uint32_t* hit_id, uint64_t n, int32_t payload,
                                    uint32_t* color, float* normal) {
    clear_error();
    if (!s || (!hit_id && n)) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (payload != YX_PAYLOAD_RGBA8 && payload != YX_PAYLOAD_C16N16)
        return fail(YX_ERR_INVALID_ARG, "unknown payload format %d", payload);
    const uint64_t num_nodes = s->nodes.size() / 10;
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t h = hit_id[i];
        uint32_t c = 0;
        float nn[3] = {0.0f, 0.0f, 0.0f};
        if (h != YX_MISS) {
            const uint64_t node = h >> 3;
            const uint32_t slot = h & 7u;
            if (node >= num_nodes)
                return fail(YX_ERR_FORMAT, "pixel %llu: node %llu beyond the %llu nodes", (unsigned long long)i,
                            (unsigned long long)node, (unsigned long long)num_nodes);
            const uint32_t* rec = s->nodes.data() + node * 10;
            if (!((rec[0] >> slot) & 1u))
                return fail(YX_ERR_FORMAT, "pixel %llu: child %u of node %llu is not a leaf", (unsigned long long)i,
                            slot, (unsigned long long)node);
            const uint32_t leaf = rec[2 + slot];
            if (payload == YX_PAYLOAD_C16N16) {
                c = host_rgba_from_565(leaf & 0xffffu);
                host_normal_from_16(leaf >> 16, nn);
            } else {
                c = leaf;
            }
        }
        if (color) color[i] = c;
        if (normal) {
            normal[3 * i] = nn[0];
            normal[3 * i + 1] = nn[1];
            normal[3 * i + 2] = nn[2];
        }
    }
    return YX_OK;
}

extern "C" void yx_scene_free(yx_scene* s) { delete s; }

extern "C" const char* yx_last_error(void) { return g_error.c_str(); }

extern "C" const char* yx_version(void) { return "yoxel-mi355x 0.1"; }
