// The C ABI (include/yoxel.h) seen from compiled C++, as a caller of the
// reference's renderer would see it. Built with g++ against the header and
// linked to libyoxel.so; the tests run it.
//
//   abi_check layout
//     prints sizeof and every field's offsetof of each yx_* struct as JSON
//     lines, for comparison with the Python ctypes mirrors and the header.
//   abi_check render <scene.vox> <frame.raw>
//     the reference's driver sequence (testRenderer, cell/main.cpp:21-40)
//     through an ISVORenderer implementation that forwards to the C ABI: the
//     interface of cell/svorenderer.h:5-31 restated here with stand-in point
//     and colour types (the reference's own headers live in its missing cpp/
//     directory). Checks that RenderFrame returns NULL without a scene, loads
//     the scene with SVOData::Load's file layout (cell/svodata.h:31-54), renders
//     1024x768 from (0.5, 0.5, 0.3) along (-1, -1, -1.5) and writes the RGBA8
//     frame to <frame.raw>; prints the renderer state as JSON.
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "yoxel.h"

// ---- layout ------------------------------------------------------------------

#define STRUCT_BEGIN(T) std::printf("{\"struct\": \"%s\", \"size\": %zu, \"fields\": {", #T, sizeof(T)); sep = "";
#define FIELD(T, f) std::printf("%s\"%s\": %zu", sep, #f, offsetof(T, f)); sep = ", ";
#define STRUCT_END() std::printf("}}\n");

static int layout() {
    const char* sep = "";
    STRUCT_BEGIN(yx_vox_node) FIELD(yx_vox_node, flags) FIELD(yx_vox_node, data) FIELD(yx_vox_node, child)
    STRUCT_END()
    STRUCT_BEGIN(yx_scene_desc) FIELD(yx_scene_desc, depth) FIELD(yx_scene_desc, num_spheres)
    FIELD(yx_scene_desc, radius_min) FIELD(yx_scene_desc, radius_max) FIELD(yx_scene_desc, heightfield)
    FIELD(yx_scene_desc, hf_octaves) FIELD(yx_scene_desc, hf_amplitude) FIELD(yx_scene_desc, hf_base)
    FIELD(yx_scene_desc, seed) STRUCT_END()
    STRUCT_BEGIN(yx_scene_info) FIELD(yx_scene_info, root) FIELD(yx_scene_info, num_nodes)
    FIELD(yx_scene_info, num_leaves) FIELD(yx_scene_info, depth) STRUCT_END()
    STRUCT_BEGIN(yx_sphere_edit) FIELD(yx_sphere_edit, mode) FIELD(yx_sphere_edit, level) FIELD(yx_sphere_edit, pos)
    FIELD(yx_sphere_edit, radius) FIELD(yx_sphere_edit, color) STRUCT_END()
    STRUCT_BEGIN(yx_volume_source) FIELD(yx_volume_source, mode) FIELD(yx_volume_source, level)
    FIELD(yx_volume_source, pos) FIELD(yx_volume_source, size) FIELD(yx_volume_source, kind)
    FIELD(yx_volume_source, iso) FIELD(yx_volume_source, inside) FIELD(yx_volume_source, color)
    FIELD(yx_volume_source, data) STRUCT_END()
    STRUCT_BEGIN(yx_outputs) FIELD(yx_outputs, hit_id) FIELD(yx_outputs, voxel) FIELD(yx_outputs, t)
    FIELD(yx_outputs, rgba) FIELD(yx_outputs, shadow) FIELD(yx_outputs, visits) STRUCT_END()
    STRUCT_BEGIN(yx_tuning) FIELD(yx_tuning, kernel) FIELD(yx_tuning, block) FIELD(yx_tuning, tile_w)
    FIELD(yx_tuning, tile_h) FIELD(yx_tuning, refill_min) FIELD(yx_tuning, queues) FIELD(yx_tuning, window)
    FIELD(yx_tuning, waves_per_cu) FIELD(yx_tuning, lds_levels) STRUCT_END()
    STRUCT_BEGIN(yx_tune_trial) FIELD(yx_tune_trial, tuning) FIELD(yx_tune_trial, ms) STRUCT_END()
    STRUCT_BEGIN(yx_light_params) FIELD(yx_light_params, enabled) FIELD(yx_light_params, pos)
    FIELD(yx_light_params, diffuse) FIELD(yx_light_params, specular) FIELD(yx_light_params, attenuation)
    STRUCT_END()
    STRUCT_BEGIN(yx_pipeline_outputs) FIELD(yx_pipeline_outputs, rgba) FIELD(yx_pipeline_outputs, z)
    FIELD(yx_pipeline_outputs, t) FIELD(yx_pipeline_outputs, albedo) FIELD(yx_pipeline_outputs, level)
    FIELD(yx_pipeline_outputs, visits) STRUCT_END()
    STRUCT_BEGIN(yx_batch_job) FIELD(yx_batch_job, scene) FIELD(yx_batch_job, width) FIELD(yx_batch_job, height)
    FIELD(yx_batch_job, pos) FIELD(yx_batch_job, dir) FIELD(yx_batch_job, up) FIELD(yx_batch_job, fov)
    FIELD(yx_batch_job, light_mode) FIELD(yx_batch_job, light) STRUCT_END()
    STRUCT_BEGIN(yx_volume_desc) FIELD(yx_volume_desc, grid) FIELD(yx_volume_desc, brick)
    FIELD(yx_volume_desc, occupancy) FIELD(yx_volume_desc, octaves) FIELD(yx_volume_desc, frequency)
    FIELD(yx_volume_desc, seed) STRUCT_END()
    STRUCT_BEGIN(yx_volume_info) FIELD(yx_volume_info, grid) FIELD(yx_volume_info, brick)
    FIELD(yx_volume_info, num_slots) STRUCT_END()
    STRUCT_BEGIN(yx_volume_dev) FIELD(yx_volume_dev, table) FIELD(yx_volume_dev, params) FIELD(yx_volume_dev, grid)
    FIELD(yx_volume_dev, brick) FIELD(yx_volume_dev, num_slots) STRUCT_END()
    STRUCT_BEGIN(yx_ntree_desc) FIELD(yx_ntree_desc, size) FIELD(yx_ntree_desc, levels)
    FIELD(yx_ntree_desc, octaves) FIELD(yx_ntree_desc, frequency) FIELD(yx_ntree_desc, threshold)
    FIELD(yx_ntree_desc, scale) FIELD(yx_ntree_desc, seed) STRUCT_END()
    STRUCT_BEGIN(yx_ntree_info) FIELD(yx_ntree_info, size) FIELD(yx_ntree_info, levels)
    FIELD(yx_ntree_info, root_side) FIELD(yx_ntree_info, num_nodes) FIELD(yx_ntree_info, num_bricks)
    FIELD(yx_ntree_info, bytes) STRUCT_END()
    STRUCT_BEGIN(yx_ntree_outputs) FIELD(yx_ntree_outputs, rgba) FIELD(yx_ntree_outputs, t)
    FIELD(yx_ntree_outputs, voxel) FIELD(yx_ntree_outputs, nodes) FIELD(yx_ntree_outputs, voxels) STRUCT_END()
    return 0;
}

// ---- the reference's renderer interface over the C ABI ----------------------

struct point_3f {
    float x, y, z;
    point_3f(float a, float b, float c) : x(a), y(b), z(c) {}
};
struct point_2i {
    int x, y;
    point_2i(int a, int b) : x(a), y(b) {}
};
struct Color32 {
    uint8_t r, g, b, a;
};
static_assert(sizeof(Color32) == 4, "Color32 is 4 bytes");

// SVOData (cell/svodata.h:22-55): a .vox file is root, two skipped words,
// the node count, then the 40-byte VoxNodes.
class SVOData {
    std::vector<yx_vox_node> nodes_;
    uint32_t root_ = YX_EMPTY_NODE;

public:
    bool Load(const char* fn) {
        FILE* f = std::fopen(fn, "rb");
        if (!f) return false;
        uint32_t hdr[4];
        bool ok = std::fread(hdr, 4, 4, f) == 4;
        if (ok) {
            nodes_.resize(hdr[3]);
            ok = hdr[3] == 0 || std::fread(nodes_.data(), sizeof(yx_vox_node), hdr[3], f) == hdr[3];
            root_ = hdr[0];
        }
        std::fclose(f);
        return ok;
    }
    uint32_t GetRoot() const { return root_; }
    const yx_vox_node& operator[](size_t i) const { return nodes_[i]; }
    uint32_t Size() const { return (uint32_t)nodes_.size(); }
};

// ISVORenderer (cell/svorenderer.h:5-31)
class ISVORenderer {
public:
    virtual ~ISVORenderer() {}
    virtual void SetScene(SVOData* svo) = 0;
    virtual void SetViewPos(const point_3f& p) = 0;
    virtual void SetViewDir(const point_3f& d) = 0;
    virtual void SetViewUp(const point_3f& u) = 0;
    virtual void SetResolution(int width, int height) = 0;
    virtual point_2i GetResolution() const = 0;
    virtual void SetFOV(float fov) = 0;
    virtual const Color32* RenderFrame() = 0;
};

// The factory a maintainer adds beside CreateSimpleRenderer /
// CreateThreadedRenderer (cell/ppu_renderer.cpp:88-107).
class MI355XRenderer : public ISVORenderer {
    yx_renderer* r_ = nullptr;
    yx_scene* s_ = nullptr;

public:
    MI355XRenderer() { yx_renderer_create(&r_); }
    ~MI355XRenderer() override {
        yx_renderer_destroy(r_);
        yx_scene_free(s_);
    }
    void SetScene(SVOData* svo) override {
        yx_scene_free(s_);
        s_ = nullptr;
        if (yx_scene_from_nodes(reinterpret_cast<const uint32_t*>(&(*svo)[0]), svo->Size(), svo->GetRoot(), &s_) ==
            YX_OK)
            yx_renderer_set_scene(r_, s_);
    }
    void SetViewPos(const point_3f& p) override { yx_renderer_set_view_pos(r_, &p.x); }
    void SetViewDir(const point_3f& d) override { yx_renderer_set_view_dir(r_, &d.x); }
    void SetViewUp(const point_3f& u) override { yx_renderer_set_view_up(r_, &u.x); }
    void SetResolution(int w, int h) override { yx_renderer_set_resolution(r_, w, h); }
    point_2i GetResolution() const override {
        int32_t w = 0, h = 0;
        yx_renderer_get_resolution(r_, &w, &h);
        return point_2i(w, h);
    }
    void SetFOV(float fov) override { yx_renderer_set_fov(r_, fov); }
    const Color32* RenderFrame() override {
        const uint32_t* c = nullptr;
        return yx_render_frame(r_, &c) == YX_OK ? reinterpret_cast<const Color32*>(c) : nullptr;
    }
};

static std::shared_ptr<ISVORenderer> CreateMI355XRenderer() { return std::make_shared<MI355XRenderer>(); }

// testRenderer (cell/main.cpp:21-40), writing the frame's bytes instead of a
// JPEG (the reference uses Magick++).
static int render(const char* scene_fn, const char* out_fn) {
    std::shared_ptr<ISVORenderer> renderer = CreateMI355XRenderer();
    const bool null_without_scene = renderer->RenderFrame() == nullptr;
    const point_2i dflt = renderer->GetResolution();
    SVOData scene;
    if (!scene.Load(scene_fn)) {
        std::fprintf(stderr, "cannot load %s\n", scene_fn);
        return 2;
    }
    renderer->SetScene(&scene);
    renderer->SetResolution(1024, 768);
    renderer->SetViewPos(point_3f(0.5f, 0.5f, 0.3f));
    renderer->SetViewDir(point_3f(-1, -1, -1.5));
    const Color32* frameBuf = renderer->RenderFrame();
    if (!frameBuf) {
        std::fprintf(stderr, "RenderFrame failed: %s\n", yx_last_error());
        return 3;
    }
    const point_2i size = renderer->GetResolution();
    FILE* f = std::fopen(out_fn, "wb");
    if (!f) return 4;
    const size_t n = (size_t)size.x * (size_t)size.y;
    const bool ok = std::fwrite(frameBuf, sizeof(Color32), n, f) == n;
    std::fclose(f);
    std::printf("{\"null_without_scene\": %s, \"default_resolution\": [%d, %d], \"resolution\": [%d, %d], "
                "\"nodes\": %u, \"root\": %u}\n",
                null_without_scene ? "true" : "false", dflt.x, dflt.y, size.x, size.y, scene.Size(),
                scene.GetRoot());
    return ok ? 0 : 5;
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "layout") == 0) return layout();
    if (argc >= 4 && std::strcmp(argv[1], "render") == 0) return render(argv[2], argv[3]);
    std::fprintf(stderr, "usage: %s layout | render <scene.vox> <frame.raw>\n", argv[0]);
    return 1;
}
