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

void build_tree_from_codes(const std::vector<uint64_t>& codes, int depth, uint64_t color_seed,
                           yx_scene& out) {
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
                    n[2 + s] = leaf_color(codes[j], color_seed);
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

}  // namespace

static int build_synthetic(const yx_scene_desc& d, yx_scene& out) {
    const int D = d.depth;
    const int n = 1 << D;
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
    Perlin noise(d.seed ^ 0xD1B54A32D192ED03ull);
    Heightfield hf{&noise, d.hf_octaves, (double)d.hf_base, (double)d.hf_amplitude, n};

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
        for (uint32_t col = 0; col < (uint32_t)((j1 - j0) * n); ++col) {
            off[col] = (uint32_t)ivs.size();
            while (e < es.size() && es[e].col == col) {
                Interval iv = es[e].iv;
                if (!ivs.empty() && off[col] < ivs.size() && iv.lo <= ivs.back().hi + 1)
                    ivs.back().hi = std::max(ivs.back().hi, iv.hi);
                else
                    ivs.push_back(iv);
                ++e;
            }
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
                        for (; k <= stop; ++k) codes.push_back(morton3(i, j, k));
                    }
                }
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

    build_tree_from_codes(codes, D, d.seed, out);
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
    *out = s;
    return YX_OK;
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
    return YX_OK;
}

extern "C" int yx_scene_set_root(yx_scene* s, uint32_t root) {
    clear_error();
    if (!s) return fail(YX_ERR_INVALID_ARG, "null argument");
    s->root = root;
    s->version += 1;
    s->depth_stale = true;
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

extern "C" void yx_scene_free(yx_scene* s) { delete s; }

extern "C" const char* yx_last_error(void) { return g_error.c_str(); }

extern "C" const char* yx_version(void) { return "yoxel-mi355x 0.1"; }
