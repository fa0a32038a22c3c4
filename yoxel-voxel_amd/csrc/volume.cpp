// Synthetic brick volume for differentiable compositing.
//
// A grid of N^3 voxels held as B^3 bricks (the brick grid G = N/B per axis,
// brick table = slot index or -1, the marked-brick test of castRay,
// reaction/trace/trace.cu:122-137, and the brick pool of BrickManager,
// ntree/BrickManager.h:5-81). Allocated bricks: those whose fBm value at the
// brick centre is at or above the (1 - occupancy) quantile over all bricks.
// Per voxel of an allocated brick: float4 {raw density, r, g, b};
// raw = 4 * (fbm3(centre) - threshold) - 1, rgb = position colour plus seeded
// noise. Voxel order inside a brick: x fastest, then y, then z.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <new>
#include <thread>

#include "internal.h"

struct yx_volume {
    int32_t grid = 0, brick = 0, bgrid = 0;
    uint32_t num_slots = 0;
    std::vector<int32_t> table;  // bgrid^3
    std::vector<float> params;   // num_slots * brick^3 * 4
};

namespace yx {

template <class F>
static void par_for(int n, F&& f) {
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

static int build_volume(const yx_volume_desc& d, yx_volume& v) {
    const int N = d.grid, B = d.brick, G = N / B;
    v.grid = N;
    v.brick = B;
    v.bgrid = G;
    Perlin noise(d.seed ^ 0x2545F4914F6CDD1Dull);
    const double freq = (double)d.frequency;
    const size_t nb = (size_t)G * G * G;
    std::vector<double> bval(nb);
    par_for(G, [&](int z) {
        for (int y = 0; y < G; ++y)
            for (int x = 0; x < G; ++x) {
                const double cx = ((double)x + 0.5) / G, cy = ((double)y + 0.5) / G, cz = ((double)z + 0.5) / G;
                bval[((size_t)z * G + y) * G + x] = noise.fbm3(cx * freq, cy * freq, cz * freq, d.octaves);
            }
    });
    std::vector<double> sorted(bval);
    size_t k = (size_t)std::floor((1.0 - (double)d.occupancy) * (double)nb);
    k = std::min(k, nb - 1);
    std::nth_element(sorted.begin(), sorted.begin() + (ptrdiff_t)k, sorted.end());
    const double thr = sorted[k];
    v.table.assign(nb, -1);
    uint32_t slots = 0;
    for (size_t i = 0; i < nb; ++i)
        if (bval[i] >= thr) v.table[i] = (int32_t)slots++;
    v.num_slots = slots;
    const int B3 = B * B * B;
    v.params.assign((size_t)slots * B3 * 4, 0.0f);
    par_for(G, [&](int bz) {
        for (int by = 0; by < G; ++by)
            for (int bx = 0; bx < G; ++bx) {
                const int32_t s = v.table[((size_t)bz * G + by) * G + bx];
                if (s < 0) continue;
                float* out = &v.params[(size_t)s * B3 * 4];
                for (int lz = 0; lz < B; ++lz)
                    for (int ly = 0; ly < B; ++ly)
                        for (int lx = 0; lx < B; ++lx) {
                            const int vx = bx * B + lx, vy = by * B + ly, vz = bz * B + lz;
                            const double px = ((double)vx + 0.5) / N, py = ((double)vy + 0.5) / N,
                                         pz = ((double)vz + 0.5) / N;
                            const double f = noise.fbm3(px * freq, py * freq, pz * freq, d.octaves);
                            const uint64_t key = ((uint64_t)vz * N + vy) * N + vx;
                            const uint64_t h = mix64(key ^ (d.seed * 0x9E3779B97F4A7C15ull));
                            const double u0 = (double)(h & 0xffff) / 65535.0;
                            const double u1 = (double)((h >> 16) & 0xffff) / 65535.0;
                            const double u2 = (double)((h >> 32) & 0xffff) / 65535.0;
                            float* o = out + 4 * (size_t)(lx + B * (ly + B * lz));
                            o[0] = (float)(4.0 * (f - thr) - 1.0);
                            o[1] = (float)std::min(1.0, std::max(0.0, 0.2 + 0.6 * px + 0.2 * (u0 - 0.5)));
                            o[2] = (float)std::min(1.0, std::max(0.0, 0.2 + 0.6 * py + 0.2 * (u1 - 0.5)));
                            o[3] = (float)std::min(1.0, std::max(0.0, 0.2 + 0.6 * pz + 0.2 * (u2 - 0.5)));
                        }
            }
    });
    return YX_OK;
}

}  // namespace yx

using namespace yx;

extern "C" int yx_volume_build_synthetic(const yx_volume_desc* d, yx_volume** out) {
    clear_error();
    if (!d || !out) return fail(YX_ERR_INVALID_ARG, "null argument");
    if (d->brick < 1 || d->brick > 16 || d->grid < d->brick || d->grid % d->brick != 0 || d->grid > 4096)
        return fail(YX_ERR_INVALID_ARG, "grid %d / brick %d: brick 1..16 dividing grid <= 4096", d->grid, d->brick);
    if (!(d->occupancy > 0.0f && d->occupancy <= 1.0f)) return fail(YX_ERR_INVALID_ARG, "occupancy in (0, 1]");
    if (d->octaves < 1 || d->octaves > 12) return fail(YX_ERR_INVALID_ARG, "octaves 1..12");
    yx_volume* v = new (std::nothrow) yx_volume();
    if (!v) return fail(YX_ERR_OOM, "out of host memory");
    try {
        build_volume(*d, *v);
    } catch (const std::bad_alloc&) {
        delete v;
        return fail(YX_ERR_OOM, "out of host memory building the volume");
    }
    *out = v;
    return YX_OK;
}

extern "C" int yx_volume_get_info(const yx_volume* v, yx_volume_info* info) {
    clear_error();
    if (!v || !info) return fail(YX_ERR_INVALID_ARG, "null argument");
    info->grid = v->grid;
    info->brick = v->brick;
    info->num_slots = v->num_slots;
    return YX_OK;
}

extern "C" int yx_volume_arrays(const yx_volume* v, const int32_t** table, const float** params) {
    clear_error();
    if (!v || !table || !params) return fail(YX_ERR_INVALID_ARG, "null argument");
    *table = v->table.data();
    *params = v->params.data();
    return YX_OK;
}

extern "C" void yx_volume_free(yx_volume* v) { delete v; }
