/*
 * yoxel.h - C ABI of the MI355X sparse-voxel-octree ray caster.
 *
 * Every entry point returns an int status (YX_OK == 0, negative on error) unless
 * stated otherwise; yx_last_error() gives a message for the calling thread's last
 * failure. Handles are opaque. No C++ exceptions cross this boundary. Device
 * pointers are plain HIP device addresses; `stream` arguments are hipStream_t
 * values passed as void* (NULL = the default stream).
 *
 * Reference interfaces replaced (paths are inside znah/yoxel-voxel):
 *   ISVORenderer            cell/svorenderer.h:5-31
 *   RendererBase            cell/renderer_base.h:7-62
 *   SVOData                 cell/svodata.h:22-55
 *   SVORenderer (GPU)       demo/SVORenderer.h:8-64, demo/SVORenderer.cpp:95-149
 *   RunTrace (C ABI)        ntree/ntree_trace.cuh:22-28
 *   VoxNode format          reaction/report/main.tex:36-56
 */
#ifndef YOXEL_H
#define YOXEL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define YX_OK              0
#define YX_ERR_INVALID_ARG (-1) /* NULL handle, zero resolution, bad config   */
#define YX_ERR_OOM         (-2) /* host or device allocation failed           */
#define YX_ERR_IO          (-3) /* missing / short / unwritable .vox file     */
#define YX_ERR_FORMAT      (-4) /* child id out of range, cycle, bad header   */
#define YX_ERR_HIP         (-5) /* a HIP runtime call failed                   */
#define YX_ERR_NO_SCENE    (-6) /* render without a scene (reference: NULL)   */
#define YX_ERR_RANGE       (-7) /* size beyond a documented limit             */

/* ---- node format (reaction/report/main.tex:36-56) ------------------------ */
/* flags: bits 0..7 leafFlags, bits 8..15 nullFlags, bit 19 emptyFlag.
 * child[i] is a node id, a leaf payload (VoxData) when leafFlags bit i is set,
 * or a special id >= 0x80000000 (EmptyNode / FullNode). Child index bit 0 is x,
 * bit 1 is y, bit 2 is z. Leaf payloads of synthetic scenes are RGBA8 (r in the
 * low byte). */
#define YX_EMPTY_NODE 0x80000000u
#define YX_FULL_NODE  0x80000001u
#define YX_MISS       0xFFFFFFFFu /* hit_id of a ray that hits nothing */

typedef struct yx_vox_node {
    uint32_t flags;
    uint32_t data;
    uint32_t child[8];
} yx_vox_node; /* 40 bytes, the reference's VoxNode */

typedef struct yx_scene yx_scene;       /* host node array + root */
typedef struct yx_renderer yx_renderer; /* camera state + device scene copy */

/* ---- scenes (SVOData, cell/svodata.h:22-55) ------------------------------ */

/* Synthetic scene recipe: spheres (centres uniform in [0,1)^3, radii uniform in
 * [radius_min, radius_max] voxels) united with an optional fBm-Perlin
 * heightfield; only surface voxels (inside, with at least one 6-neighbour
 * outside; outside the domain counts as outside) become leaves. */
typedef struct yx_scene_desc {
    int32_t depth;          /* leaf grid is 2^depth per axis, 1..16            */
    int32_t num_spheres;
    float radius_min;       /* in voxels */
    float radius_max;       /* in voxels */
    int32_t heightfield;    /* 0 or 1 */
    int32_t hf_octaves;     /* fBm octaves */
    float hf_amplitude;     /* fraction of the extent */
    float hf_base;          /* base height, fraction of the extent */
    uint64_t seed;
} yx_scene_desc;

typedef struct yx_scene_info {
    uint32_t root;
    uint32_t num_nodes;
    uint64_t num_leaves;
    int32_t depth;          /* deepest leaf level (root children are level 1) */
} yx_scene_info;

/* SVOData::Load (cell/svodata.h:31-50): header root:u32, 2 x u32 (ignored),
 * size:u32, then size x 40-byte VoxNode, little-endian. */
int yx_scene_load_vox(const char* path, yx_scene** out);
int yx_scene_save_vox(const yx_scene* scene, const char* path);
/* Copies `num_nodes` VoxNode records (10 u32 each). */
int yx_scene_from_nodes(const uint32_t* nodes, uint32_t num_nodes, uint32_t root,
                        yx_scene** out);
int yx_scene_build_synthetic(const yx_scene_desc* desc, yx_scene** out);
/* The same scene built on the GPU (depth <= 12): the identical VoxNode array,
 * root and counts as yx_scene_build_synthetic, returned as a host scene.
 * SVO building is the step before the trace (DynamicSVO::BuildRange,
 * ore/src/main.cpp:37-129); no GPU counterpart in the reference. */
int yx_scene_build_synthetic_gpu(const yx_scene_desc* desc, yx_scene** out, void* stream);
int yx_scene_get_info(const yx_scene* scene, yx_scene_info* info);
/* Edits of the node array in place (the node pool of DynamicSVO / HomoStorage,
 * reaction/report/main.tex:67-71): writes `count` VoxNode records (10 u32 each)
 * from index `first` (at most the current size; the array grows as needed) and
 * stamps their 256-node pages with a new version. The tree is validated when it
 * is next laid out (yx_renderer_set_scene / yx_renderer_update_scene /
 * yx_scene_get_info), so a sequence of writes may pass through invalid states.
 * Invalidates pointers from yx_scene_nodes. */
int yx_scene_write_nodes(yx_scene* scene, uint32_t first, const uint32_t* nodes, uint32_t count);
int yx_scene_set_root(yx_scene* scene, uint32_t root);

/* DynamicSVO::BuildRange(level, pos, mode, SphereSource) (ore/src/main.cpp:
 * 66-70 and 121-127 bind it; demo/Demo.cpp:82-114 and qtview.py:144-153 call
 * it) on a synthetic scene. The scene keeps its recipe and edit list; the call
 * appends `count` edits. The solid is the recipe's, with each edit applied in
 * order: GROW unions the ball of voxel centres within `radius` of `pos`, CLEAR
 * subtracts it; leaves are its surface voxels as for yx_scene_build_synthetic.
 * A leaf's colour is that of the last edit whose ball contains it when that
 * edit is a GROW, else the recipe's colour. `how`:
 *   YX_RANGE_IN_PLACE     edit the node array in place on the host, one edit
 *                         at a time: nodes outside an edit's ball (plus one
 *                         voxel) keep their ids and bytes, emptied nodes are
 *                         blanked (flags 0xFF00, no children) and their ids
 *                         reused lowest first before the array grows;
 *   YX_RANGE_REBUILD      rebuild the canonical breadth-first tree on the host;
 *   YX_RANGE_REBUILD_GPU  the same tree, built on the GPU on `stream`.
 * Every method gives the same voxels and colours; the two rebuilds give the
 * same bytes. Pages whose bytes changed get a new version.
 * Scenes without a recipe (loaded, or edited with yx_scene_write_nodes) are
 * edited in place only (YX_ERR_INVALID_ARG for the rebuilds), as trees: their
 * solid is their leaves (a leaf slot covers its child cube) and FullNode
 * cubes, levels count from the tree's deepest leaf level, and the voxels of
 * each edit's box plus one are encoded again at that level: surface voxels
 * as leaves, other solid voxels as FullNode (whole FullNode cubes collapse),
 * the rest empty. Leaves keep their colour, grown voxels take the source's,
 * and solid voxels that a CLEAR exposes take the CLEAR's `color`. */
#define YX_BUILD_GROW  0
#define YX_BUILD_CLEAR 1
#define YX_RANGE_IN_PLACE    0
#define YX_RANGE_REBUILD     1
#define YX_RANGE_REBUILD_GPU 2
typedef struct yx_sphere_edit {
    int32_t mode;   /* YX_BUILD_GROW or YX_BUILD_CLEAR                        */
    int32_t level;  /* grid level of pos and radius, 1..depth (BuildRange's)  */
    float pos[3];   /* centre, in voxels of that level                         */
    float radius;   /* in voxels of that level, > 0                            */
    uint32_t color; /* RGBA8 of grown voxels (SphereSource's colour); of voxels
                       a CLEAR exposes in a scene without a recipe             */
} yx_sphere_edit;
int yx_scene_build_range(yx_scene* scene, const yx_sphere_edit* edits, uint32_t count, int32_t how,
                         void* stream);
/* BuildRange with a volume source: IsoSource (scene_gen.py:111-130, a uint8
 * volume thresholded at an iso level, ore/src/main.cpp:53-63,71-75,115-119) or
 * RawSource (scene_gen.py:74-82, RGBA8 voxels, ore/src/main.cpp:37-50,109-110). The
 * source's voxel (a, b, c), a fastest in memory, covers voxel pos + (a, b, c)
 * of `level` (2^(depth - level) finest voxels per axis); parts outside the
 * domain are dropped. Solid voxels: IsoSource value >= iso (inside != 0) or
 * value < iso (inside == 0), coloured `color`; RawSource alpha != 0, coloured
 * by the voxel. GROW unions them, CLEAR subtracts them, in order with the
 * sphere edits, by the same methods. The data are copied. */
/* Solid interior of a tree without a recipe whose solid is only its surface
 * (a saved synthetic scene, for one): the empty space that a 6-connected flood
 * from the domain's faces does not reach becomes FullNode, so that edits see
 * a solid (yx_scene_build_range). Leaves and rendering are unchanged. Each
 * empty child cube is one connected empty region and changes whole; *filled
 * (may be NULL) receives how many. Up to depth 11 (two bit volumes of the
 * leaf grid in host memory). No counterpart in the reference. */
int yx_scene_fill_interior(yx_scene* scene, uint64_t* filled);
#define YX_SOURCE_ISO 1
#define YX_SOURCE_RAW 2
typedef struct yx_volume_source {
    int32_t mode;     /* YX_BUILD_GROW or YX_BUILD_CLEAR                         */
    int32_t level;    /* 1..depth                                                 */
    int32_t pos[3];   /* where source voxel (0, 0, 0) lands, voxels of `level`    */
    int32_t size[3];  /* source voxels per axis, each 1..4096                     */
    int32_t kind;     /* YX_SOURCE_ISO: data = uint8 values; YX_SOURCE_RAW: data =
                         RGBA8 (r low byte) per voxel                              */
    int32_t iso;      /* IsoSource level (SetIsoLevel)                            */
    int32_t inside;   /* IsoSource: 1 solid at or above iso, 0 below (SetInside)  */
    uint32_t color;   /* IsoSource colour (SetColor)                              */
    const void* data; /* host memory, size[0] * size[1] * size[2] voxels          */
} yx_volume_source;
int yx_scene_build_range_volume(yx_scene* scene, const yx_volume_source* source, int32_t how, void* stream);
/* SVOData::GetRoot / operator[]: borrowed view of the node array (valid until
 * yx_scene_free). */
int yx_scene_nodes(const yx_scene* scene, const uint32_t** nodes, uint32_t* num_nodes,
                   uint32_t* root);
void yx_scene_free(yx_scene* scene);

/* ---- renderer (ISVORenderer cell/svorenderer.h:5-31, SVORenderer) -------- */

/* Light modes. The reference CPU shader puts the light at the eye
 * (RendererBase::SetViewPos, cell/renderer_base.h:30-35): YX_LIGHT_HEADLIGHT.
 * YX_LIGHT_SUN_SHADOW adds one shadow ray per lit hit toward a fixed
 * directional light. */
#define YX_LIGHT_HEADLIGHT  0
#define YX_LIGHT_SUN        1
#define YX_LIGHT_SUN_SHADOW 2

/* Per-pixel outputs; any pointer may be NULL to skip that output. Device
 * pointers for yx_render, row-major (y * width + x). */
typedef struct yx_outputs {
    uint32_t* hit_id; /* canonical parent VoxNode id * 8 + child slot, YX_MISS  */
    int32_t* voxel;   /* 3 x int32 per pixel: integer cell coords at hit level,
                         (-1,-1,-1) on a miss                                    */
    float* t;         /* hit distance along the unit ray, -1 on a miss          */
    uint32_t* rgba;   /* shaded RGBA8 (r low byte); (0,0,0,0) on a miss         */
    uint8_t* shadow;  /* 0 lit, 1 shadowed, 2 no shadow ray cast                */
    uint32_t* visits; /* node records read (primary + shadow), for byte counts  */
} yx_outputs;

int yx_renderer_create(yx_renderer** out);
void yx_renderer_destroy(yx_renderer* r);
/* SetScene (CudaSVO::SetSVO, demo/SVORenderer.cpp:39-53): uploads the
 * scene's VoxNode array to HBM and lays it out for traversal on the GPU (the
 * first-sync path of yx_renderer_update_scene), then waits for it. The scene
 * may be freed afterwards; the renderer keeps its own device copies. */
int yx_renderer_set_scene(yx_renderer* r, const yx_scene* scene);
/* The same layout computed on the host and uploaded (no device VoxNode
 * copy). Bit-identical records and LOD colours; slower for large scenes. No
 * counterpart in the reference: kept as a second, independent layout path. */
int yx_renderer_set_scene_host(yx_renderer* r, const yx_scene* scene);
int yx_renderer_set_view_pos(yx_renderer* r, const float pos[3]);
int yx_renderer_set_view_dir(yx_renderer* r, const float dir[3]);
int yx_renderer_set_view_up(yx_renderer* r, const float up[3]);
int yx_renderer_set_resolution(yx_renderer* r, int32_t width, int32_t height);
int yx_renderer_get_resolution(const yx_renderer* r, int32_t* width, int32_t* height);
int yx_renderer_set_fov(yx_renderer* r, float fov_deg); /* horizontal, default 70 */
int yx_renderer_set_light(yx_renderer* r, const float dir[3], int32_t mode);
/* How leaf payloads (VoxData) are read when shading yx_render /
 * yx_render_frame / yx_render_noalloc frames:
 *   YX_PAYLOAD_RGBA8 (default): RGBA8 colour (r in the low byte), Lambert over
 *     the entry face of the hit cell;
 *   YX_PAYLOAD_C16N16: the reference's colour16 + normal16 (VoxData,
 *     reaction/report/main.tex:38,65; UnpackVoxData / UnpackColor /
 *     UnpackNormal, demo/SVORenderer.cpp:181-185, bodies missing). Layout
 *     defined here: bits 0-15 RGB565 (red in 11-15, widened by bit
 *     replication); bits 16-31 a normal quantized after Baptista, "Higher
 *     Accuracy Quantized Normals" (the thesis's citation): bits 15/14/13 of the
 *     half the signs of x/y/z, bits 7-12 u and 0-6 v of the triangle
 *     u + v <= 126 (points with u >= 64 folded to 127 - u, 127 - v); normal =
 *     normalize(u, v, 126 - u - v) with the signs. Lambert over that normal;
 *     shadow rays as for RGBA8.
 * Batches take the format from yx_batch_set_payload; the shading pipeline
 * (yx_render_pipeline) reads RGBA8. */
#define YX_PAYLOAD_RGBA8 0
#define YX_PAYLOAD_C16N16 1
int yx_renderer_set_payload(yx_renderer* r, int32_t format);
/* SVORenderer::DumpTraceData (demo/SVORenderer.cpp:158-192), its decoding
 * step: for each of `n` pixels of a frame's hit_id output (a host copy), the
 * hit leaf's VoxData (node hit_id / 8, child hit_id % 8 of `scene`) read as
 * `payload`. color: RGBA8 (C16N16: the RGB565 colour widened as above, alpha
 * 255); normal: 3 floats per pixel (C16N16: the stored normal; RGBA8 stores
 * none and gives 0, 0, 0). Misses (YX_MISS) give zeros, as the reference skips
 * them. Either output may be NULL. YX_ERR_FORMAT for an id that names no leaf
 * of the scene. The reference also writes a distance file that it leaves at
 * zero (its fill line is commented out); the hit distance is yx_outputs.t. */
int yx_scene_decode_hits(const yx_scene* scene, const uint32_t* hit_id, uint64_t n, int32_t payload,
                         uint32_t* color, float* normal);
/* Ray directions of the current camera (RendererBase::InitRayDir,
 * cell/renderer_base.h:50-61): dir0, du, dv as 9 floats. */
int yx_renderer_ray_basis(const yx_renderer* r, float basis[9]);

/* Render into caller-owned device buffers on `stream` (asynchronous). The GPU
 * variant of RenderFrame with the device-buffer shape of SVORenderer::Render
 * (demo/SVORenderer.cpp:95-149); the reference's full GPU pipeline is
 * yx_render_pipeline below. */
int yx_render(yx_renderer* r, const yx_outputs* out, void* stream);
/* yx_render without any device allocation. No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops that forbid hipMalloc). The scene must already be set (set_scene
 * allocates at set-up). The persistent kernels' scratch (ray-queue heads and
 * retired-wave count: 65 words, 260 bytes) is a fixed per-device array inside
 * the library instead of renderer-owned memory; a tuning with more queues than
 * it holds is refused with YX_ERR_RANGE. Outputs equal yx_render's. Every call
 * on one GPU shares that array: calls that overlap on one GPU (other streams,
 * other threads, other renderers) race and must be serialised by the caller. */
int yx_render_noalloc(yx_renderer* r, const yx_outputs* out, void* stream);
/* RenderFrame (cell/ppu_renderer.cpp:76-85): synchronous; returns a pointer to
 * a renderer-owned host RGBA8 buffer valid until the next call or resolution
 * change. Fails with YX_ERR_NO_SCENE when no scene is set (the reference
 * returns NULL). */
int yx_render_frame(yx_renderer* r, const uint32_t** colors);

/* Incremental scene sync (CudaSVO::Update, demo/SVORenderer.cpp:39-53,
 * qtview.py:69-73): the renderer keeps a device copy of the scene's VoxNode
 * array and uploads only the 256-node pages edited since its last sync (all of
 * them for another scene object), then rebuilds its traversal records and LOD
 * colours on the GPU. Results are identical to yx_renderer_set_scene_host. On a
 * format error the previous records stay in place. */
int yx_renderer_update_scene(yx_renderer* r, const yx_scene* scene, void* stream);
/* DynamicSVO::CountChangedPages / CountTransfrerSize (ore/src/main.cpp:127-128):
 * pages and bytes the next yx_renderer_update_scene would upload. */
int yx_renderer_count_changed_pages(const yx_renderer* r, const yx_scene* scene, uint32_t* pages,
                                    uint64_t* bytes);
/* Pages and bytes uploaded by the last yx_renderer_update_scene. */
int yx_renderer_last_update(const yx_renderer* r, uint32_t* pages, uint64_t* bytes);
/* Copies the device traversal layout to host memory: 2 u32 per record (node:
 * first child record, child mask | leaf mask << 8; leaf: RGBA8, parent id) and
 * the LOD colour per record. Either array may be NULL; *n_recs is always set.
 * No reference counterpart (inspection and tests of the re-layout). */
int yx_renderer_layout(const yx_renderer* r, uint32_t* recs, uint32_t* lod, uint64_t cap, uint64_t* n_recs);

/* Device memory used by the uploaded scene, in bytes. */
int yx_renderer_scene_bytes(const yx_renderer* r, uint64_t* bytes);

/* Launch shape of the ray-cast kernels. No reference counterpart (the
 * reference fixes 16x16 / 16x32 blocks, trace_cuda.py:26); set by the
 * autotuner. Every setting renders identical outputs. */
#define YX_KERNEL_PER_PIXEL  0 /* one thread per pixel, 16x16 tiles           */
#define YX_KERNEL_PERSISTENT 1 /* persistent waves on atomic ray queues, per-lane
                                  traversal stack in LDS                       */
#define YX_KERNEL_REPLAY     2 /* persistent waves, no t stack: ancestors are
                                  rebuilt exactly by replay (deep trees)       */
typedef struct yx_tuning {
    int32_t kernel;       /* YX_KERNEL_*                                        */
    int32_t block;        /* threads per workgroup: 64, 128, 256 (stack kernel)
                             or 256, 512 (replay kernel)                        */
    int32_t tile_w;       /* ray packet (queue tile) width in pixels, power of 2 */
    int32_t tile_h;       /* ray packet height, power of 2; tile_w*tile_h <= 4096 */
    int32_t refill_min;   /* idle lanes before a wave takes new rays, 1..64     */
    int32_t queues;       /* ray queues, 1..64 (8 = one per XCD)                */
    int32_t window;       /* stack levels kept in registers: 0..6 (stack kernel
                             uses 0, 2, 4, 6; replay kernel 0, 1, 2)           */
    int32_t waves_per_cu; /* persistent launch size; 0 = occupancy limit        */
    int32_t lds_levels;   /* stack kernel, single scene: node records of tree
                             levels 0..lds_levels-1 staged in LDS per workgroup
                             (0 = none; fewer levels when LDS is short)        */
} yx_tuning;
int yx_renderer_set_tuning(yx_renderer* r, const yx_tuning* tuning);
int yx_renderer_get_tuning(const yx_renderer* r, yx_tuning* tuning);

/* ---- autotuner ------------------------------------------------------------
 * Coordinate descent over the tuning space of the stack kernel: ray packet
 * shape (tile_w x tile_h), queue chunk (refill_min: idle lanes a wave gathers
 * before it takes that many ray ids with one atomic), register window (levels
 * kept in registers; the rest of the traversal stack is LDS-resident),
 * workgroup size and queue count. Each candidate renders `reps` timed frames
 * (after one untimed one) into `out` on `stream`; the best is kept in the
 * object. `trials` (capacity `cap`, may be NULL) receives every candidate in
 * the order tried with its mean frame time in ms. No reference counterpart
 * (the reference hard-codes its launch shape, trace_cuda.py:26). */
typedef struct yx_tune_trial {
    yx_tuning tuning;
    float ms;
} yx_tune_trial;
int yx_renderer_autotune(yx_renderer* r, const yx_outputs* out, int32_t reps, yx_tune_trial* trials,
                         int32_t cap, int32_t* n_trials, void* stream);

/* ---- the reference GPU pipeline: LOD trace, depth blur, Phong shade -------
 * SVORenderer::Render (demo/SVORenderer.cpp:95-149) runs Run_Trace with a
 * level-of-detail stop, five depth-limited blurs of the z buffer (Run_BlurZ,
 * "SSNA") and Run_ShadeSimple with point lights and normals taken from the z
 * buffer. Its kernels are missing from the reference; the definitions this
 * library implements:
 *  - LOD stop: a node below the root whose cell edge 2^-level is smaller than
 *    its entry distance times detail = coef * rad(fov/2) / width
 *    (demo/SVORenderer.cpp:104), and which has a non-null child, ends the ray;
 *    its colour is the mean RGBA8 of the leaves under it (VoxNode::data,
 *    reaction/report/main.tex:61). coef 0 traces exactly.
 *  - z = t / |dir0 + du x + dv y|, the view depth (+inf on a miss).
 *  - blur pass i (i = 0..4, blurSize = 3 + 3i): pixels with z < 5 vox /
 *    (pixelAng blurSize) (vox = 2^-depth, pixelAng = rad(fov) / width) become
 *    the 5x5 Gaussian (SVORenderer::InitBlur) mean of the window pixels whose
 *    z differs by less than 4 vox, renormalised by their weights.
 *  - normals from z by one-sided differences of the smaller magnitude
 *    (demo/dumps/ztools.py:23-42); Phong with ambient 0.1 and specular
 *    exponent 10 (demo/SVORenderer.cpp:112-113) summed over the enabled lights,
 *    each attenuated by 1 / (c0 + c1 d + c2 d^2).
 * Primary rays only; no shadow rays. */
#define YX_MAX_LIGHTS 4
typedef struct yx_light_params { /* LightParams, demo/Demo.cpp:141-146 */
    int32_t enabled;
    float pos[3];         /* world position                                  */
    float diffuse[3];     /* per-channel diffuse intensity                   */
    float specular[3];    /* per-channel specular intensity                  */
    float attenuation[3]; /* constant, linear, quadratic coefficients        */
} yx_light_params;
typedef struct yx_pipeline_outputs {
    uint32_t* rgba;   /* shaded RGBA8 (r low byte); (0,0,0,0) on a miss     */
    float* z;         /* view depth after the blur passes; +inf on a miss   */
    float* t;         /* hit distance along the unit ray, -1 on a miss      */
    uint32_t* albedo; /* leaf colour, or the node's mean colour on an LOD
                         stop; 0 on a miss                                  */
    uint8_t* level;   /* level of the hit cell (1 = child of the root); 0 on
                         a miss                                             */
    uint32_t* visits; /* node records read                                  */
} yx_pipeline_outputs;
/* SVORenderer::SetDetailCoef / GetDetailCoef (demo/SVORenderer.h:25-26);
 * default 1. coef >= 0. */
int yx_renderer_set_detail_coef(yx_renderer* r, float coef);
int yx_renderer_get_detail_coef(const yx_renderer* r, float* coef);
/* SVORenderer::SetSSNA / GetSSNA (demo/SVORenderer.h:28-29); default on. */
int yx_renderer_set_ssna(yx_renderer* r, int32_t enable);
int yx_renderer_get_ssna(const yx_renderer* r, int32_t* enable);
/* SVORenderer::SetShowNormals (demo/SVORenderer.h:31): RGB = normal * 0.5 + 0.5. */
int yx_renderer_set_show_normals(yx_renderer* r, int32_t enable);
/* SVORenderer::SetLigth (demo/SVORenderer.h:34); i in 0..YX_MAX_LIGHTS-1, all
 * disabled by default. */
int yx_renderer_set_point_light(yx_renderer* r, int32_t i, const yx_light_params* light);
/* SVORenderer::Render(void*) (demo/SVORenderer.cpp:95-149) into caller-owned
 * device buffers on `stream` (asynchronous). Any output may be NULL; the
 * renderer keeps its own z and colour work buffers. */
int yx_render_pipeline(yx_renderer* r, const yx_pipeline_outputs* out, void* stream);

/* ---- batched rendering: many scenes x many cameras in one launch ---------
 * Replaces a loop of per-scene SetScene / SetView / Render calls
 * (demo/Demo.cpp:138-174 drives one SVORenderer per scene and view) with one
 * persistent launch over the rays of every job. A job is one camera on one
 * scene of the batch; results are identical to rendering each job alone with
 * yx_render. Outputs are per pixel as in yx_outputs, jobs concatenated in job
 * order (job j starts at pixel yx_batch_job_offset(j)). */
typedef struct yx_batch yx_batch;
typedef struct yx_batch_job {
    int32_t scene;         /* index into the batch's scene list             */
    int32_t width, height;
    float pos[3], dir[3], up[3];
    float fov;             /* horizontal, degrees                           */
    int32_t light_mode;    /* YX_LIGHT_*                                     */
    float light[3];        /* unit direction toward the light               */
} yx_batch_job;
int yx_batch_create(yx_batch** out);
void yx_batch_destroy(yx_batch* b);
/* Uploads all scenes into one device record pool (each re-laid out as for
 * yx_renderer_set_scene). Replaces any previous scenes and jobs. */
int yx_batch_set_scenes(yx_batch* b, const yx_scene* const* scenes, int32_t n);
/* Sets the jobs; total_pixels (may be NULL) receives the output size. Calling
 * it again with only cameras and lights changed keeps the captured graph. */
int yx_batch_set_jobs(yx_batch* b, const yx_batch_job* jobs, int32_t n, uint64_t* total_pixels);
int yx_batch_job_offset(const yx_batch* b, int32_t job, uint64_t* first_pixel);
int yx_batch_set_tuning(yx_batch* b, const yx_tuning* tuning); /* stack kernel only */
int yx_batch_get_tuning(const yx_batch* b, yx_tuning* tuning);
/* use_graph != 0 (default): each frame replays a hipGraph captured from the
 * frame's launch (re-captured when outputs, tuning or job layout change);
 * 0: plain kernel launch. */
int yx_batch_set_graph(yx_batch* b, int32_t use_graph);
/* One frame: every job's rays in one persistent launch on `stream`. */
int yx_render_batch(yx_batch* b, const yx_outputs* out, void* stream);
int yx_batch_autotune(yx_batch* b, const yx_outputs* out, int32_t reps, yx_tune_trial* trials, int32_t cap,
                      int32_t* n_trials, void* stream);
int yx_batch_scene_bytes(const yx_batch* b, uint64_t* bytes);
/* How the leaf payloads of every scene of the batch are read (YX_PAYLOAD_*,
 * as yx_renderer_set_payload; default RGBA8). No counterpart in the
 * reference, whose renderer handles one scene at a time. */
int yx_batch_set_payload(yx_batch* b, int32_t format);

/* ---- differentiable front-to-back compositing over brick voxels ----------
 * Forward model: reaction/trace/trace.cu:42-139 (castRay / walkBrick: brick
 * grid with occupancy table, samples on the lattice t = k*dt with dt = one
 * voxel, front-to-back accumulation); the thesis' recurrence
 * reaction/report/voxel.tex:59-87. Per sample: sigma = softplus(raw),
 * alpha = 1 - exp(-sigma), C += T*alpha*rgb, T *= exp(-sigma); nearest-voxel
 * lookup; no early exit (the loss stays smooth in the densities). */

typedef struct yx_volume_desc {
    int32_t grid;      /* voxels per axis, power of two (512)               */
    int32_t brick;     /* voxels per brick edge, power of two (8)            */
    float occupancy;   /* fraction of bricks allocated (0.1)                 */
    int32_t octaves;   /* fBm octaves of the density field                   */
    float frequency;   /* fBm base frequency over the unit cube              */
    uint64_t seed;
} yx_volume_desc;

typedef struct yx_volume_info {
    int32_t grid;
    int32_t brick;
    uint32_t num_slots; /* allocated bricks */
} yx_volume_info;

typedef struct yx_volume yx_volume; /* host arrays of a synthetic volume */

int yx_volume_build_synthetic(const yx_volume_desc* desc, yx_volume** out);
int yx_volume_get_info(const yx_volume* v, yx_volume_info* info);
/* Borrowed host views: brick table int32[(grid/brick)^3] (slot or -1, x
 * fastest) and params float[num_slots * brick^3 * 4] = {raw, r, g, b} per
 * voxel, x fastest inside a brick. */
int yx_volume_arrays(const yx_volume* v, const int32_t** brick_table, const float** params);
void yx_volume_free(yx_volume* v);

/* A camera as 12 floats: pos[3] then RendererBase::InitRayDir's dir0, du, dv
 * (cell/renderer_base.h:50-61); pixel (x, y) looks along
 * normalize(dir0 + du*x + dv*y). */
int yx_camera_basis(const float pos[3], const float dir[3], const float up[3], float fov_deg, int32_t width,
                    int32_t height, float view[12]);

/* Device views of a volume (caller-owned memory). */
typedef struct yx_volume_dev {
    const int32_t* table;
    const float* params; /* float4 per voxel slot */
    int32_t grid;
    int32_t brick;
    uint32_t num_slots;
} yx_volume_dev;

/* Holds the backward's workspace and a summary of the brick table (one byte
 * per 4^3 bricks, rebuilt by every forward and backward; allocated on the
 * first call and whenever the grid grows): calls on one compositor must not
 * overlap (use one stream, or one compositor per stream). */
typedef struct yx_compositor yx_compositor;
int yx_compositor_create(yx_compositor** out);
void yx_compositor_destroy(yx_compositor* c);

/* Forward over n_views cameras (device float[12 * n_views]) of width x height
 * rays each; outputs per ray, view-major then row-major: rgb float[3],
 * transmittance float, samples in allocated bricks uint32 (any may be NULL
 * except nsamp when a backward follows). Asynchronous on `stream`. */
int yx_composite_forward(yx_compositor* c, const yx_volume_dev* vol, const float* views, int32_t n_views,
                         int32_t width, int32_t height, float* rgb, float* trans, uint32_t* nsamp,
                         void* stream);
/* Adjoint of the forward: given the forward's rgb and nsamp and dL/drgb,
 * writes dL/dparams (float4 per voxel slot: d raw, d r, d g, d b). Recomputes
 * the exact sample sequence, emits one (voxel key, gradient) record per
 * sample, sorts the records by brick slot (bricks up to 8^3; larger bricks by
 * the whole key) and sums each voxel's records in emission order: no atomics
 * on gradients, deterministic. nsamp must be the forward's.
 * The record total (sum of nsamp) stays on the device. The workspace holds a
 * capacity of records: it is sized with one wait on `stream` the first time a
 * geometry (brick table and views pointers, sizes) is seen, unless
 * yx_compositor_reserve set it; calls on a known geometry never wait (a
 * training step can be captured in a graph). A call whose records exceed the
 * capacity (its table was edited in place to hold more samples) writes a NaN
 * gradient; the next call grows the workspace and yx_compositor_last_stats
 * reports it. */
int yx_composite_backward(yx_compositor* c, const yx_volume_dev* vol, const float* views, int32_t n_views,
                          int32_t width, int32_t height, const float* rgb, const uint32_t* nsamp,
                          const float* grad_rgb, float* grad_params, void* stream);
/* Capacity of the backward's workspace in records (samples), so that no
 * backward waits to size it. No reference counterpart (the reference has no
 * adjoint). */
int yx_compositor_reserve(yx_compositor* c, uint64_t records);
/* Records emitted by the last backward (sum of nsamp) and distinct keys.
 * Statistics only: waits for the device, then reads them. YX_ERR_RANGE when
 * the last backward overflowed its capacity. */
int yx_compositor_last_stats(const yx_compositor* c, uint64_t* records, uint64_t* unique_keys);

/* ---- brick pool manager (CuBrickPoolManager::allocMap,
 * reaction/voxtree/brickpool/BrickPool.cpp:21-62; UpdateBrickPool,
 * reaction/voxtree/density.cu:240-284) ---------------------------------------
 * Allocates and frees parameter slots of a brick volume in place: the pool
 * manages the caller's device table and params of `vol` (writing both) and
 * keeps a used-slot map. Policy: the listed bricks that have no slot, in order
 * of their first appearance in the list, get the lowest free slots in
 * increasing order; bricks that have one keep it. The reference fills the
 * least-used slices of a 3D texture pool; this pool is one linear array.
 * Calls on one pool must not overlap. */
typedef struct yx_brick_pool yx_brick_pool;
/* The table may already map bricks to slots (YX_ERR_FORMAT if a slot is used
 * twice or lies beyond num_slots). */
int yx_brick_pool_create(const yx_volume_dev* vol, void* stream, yx_brick_pool** out);
void yx_brick_pool_destroy(yx_brick_pool* p);
/* bricks: device int32[n], linear brick indices x + g (y + g z). slots_out
 * (device int32[n], may be NULL) receives every listed brick's slot. Fails
 * with YX_ERR_RANGE and changes nothing when the free slots do not suffice or
 * an index is out of range. */
int yx_brick_pool_alloc(yx_brick_pool* p, const int32_t* bricks, uint32_t n, int32_t* slots_out,
                        uint32_t* new_count, void* stream);
/* Frees the listed bricks' slots (table entries become -1); bricks without a
 * slot are skipped. */
int yx_brick_pool_release(yx_brick_pool* p, const int32_t* bricks, uint32_t n, uint32_t* released, void* stream);
/* UpdateBrickPool: copies packed brick contents (device float4[n][brick^3], in
 * list order) into the listed bricks' slots; bricks without a slot are
 * skipped, and a brick listed more than once takes its last entry. Fails with
 * YX_ERR_RANGE (nothing copied) on an index out of range. */
int yx_brick_pool_update(yx_brick_pool* p, const int32_t* bricks, uint32_t n, const float* packed, void* stream);
int yx_brick_pool_stats(const yx_brick_pool* p, uint32_t* used, uint32_t* capacity);

/* ---- brick classification of a bit volume (reaction/voxtree) ---------------
 * The voxtree pipeline (reaction/voxtree/voxtree.py:80-95): 2x2x2 density of
 * a bit volume, uniform / mixed classification of 4^3 bricks with one voxel of
 * overlap, and the packed list of mixed bricks. All pointers are device
 * pointers; B = bit volume side (a power of two, 128..1024), V = B / 2,
 * G = V / 4.
 *   bits         uint32 [B/128][B (y)][B (x)][4]: bit b of word l of slice s
 *                is voxel z = 128 s + 32 l + b (the voxelizer's layout)
 *   density      uint32 [V (y)][V (x)][V/8]: 4-bit counts 0..8, word w of a
 *                column holds z = 8w..8w+7 from the low nibble up
 *   brick_state  uint32 [G (y)][G (x)][G (z)]: 0x40000000 empty, 0x80000000
 *                solid (over the brick's 5^3 counts; counts past the volume
 *                are 0), ORed with the number of mixed bricks below it in its
 *                column
 *   column_count uint32 [G (y)][G (x)]: mixed bricks per column
 *   column_start uint32 exclusive prefix sum of column_count
 *   packed       uint32 per mixed brick: x | y << 8 | z << 16, ordered by
 *                column, then z
 * bits, density, brick_state, column_count and column_start must be 16-byte
 * aligned (YX_ERR_INVALID_ARG). */
/* CalcDensity (reaction/voxtree/density.cu:32-60). */
int yx_voxtree_calc_density(const uint32_t* bits, int32_t bit_size, uint32_t* density, void* stream);
/* MarkBricks (reaction/voxtree/density.cu:131-176); vol_size = V. */
int yx_voxtree_mark_bricks(const uint32_t* density, int32_t vol_size, uint32_t* brick_state,
                           uint32_t* column_count, void* stream);
/* PackBricks (reaction/voxtree/density.cu:210-221); writes below `capacity`. */
int yx_voxtree_pack_bricks(const uint32_t* brick_state, const uint32_t* column_start, int32_t vol_size,
                           uint32_t* packed, uint32_t capacity, void* stream);
/* All three and the column scan that voxtree.py:88-89 does on the host
 * (cumsum_ex), on the GPU. column_start holds G*G + 1 entries, the last the
 * total; *num_packed (host) receives it after a synchronisation of `stream`.
 * YX_ERR_RANGE, with nothing packed, when it exceeds `capacity`. */
int yx_voxtree_classify(const uint32_t* bits, int32_t bit_size, uint32_t* density, uint32_t* brick_state,
                        uint32_t* column_count, uint32_t* column_start, uint32_t* packed, uint32_t capacity,
                        uint32_t* num_packed, void* stream);

/* ---- brick n-tree volumes ---------------------------------------------------
 * The n-tree of ntree/ntree/nodes.h:13-100 (Grid nodes of 4^3 cells with the
 * ch2ofs cell order (z*4 + y)*4 + x, :99; Brick leaves, ntree/BrickManager.h
 * brick pool) with 8^3 u8 voxel bricks: size = 8 * 4^levels * R voxels per
 * axis, an R^3 root grid (R <= 16) above `levels` levels of 4^3 nodes.
 * Built on the GPU (MarkBricks / PackBricks, reaction/voxtree/density.cu:
 * 131-221, as classify -> compact -> pack); device-resident. */
typedef struct yx_ntree yx_ntree;
/* u8 density from 3D fBm: f = sum_{o < octaves} 0.5^o noise3(p * frequency *
 * 2^o), p = (voxel + 0.5) / size, improved Perlin noise over a permutation
 * seeded by `seed`; density 0 where f < threshold, else 1 + min(254,
 * (int)((f - threshold) * scale)); float32 throughout. */
typedef struct yx_ntree_desc {
    int32_t size, levels, octaves;
    float frequency, threshold, scale;
    uint64_t seed;
} yx_ntree_desc;
typedef struct yx_ntree_info {
    int32_t size, levels, root_side;
    uint64_t num_nodes, num_bricks, bytes; /* bytes: root + nodes + bricks in HBM */
} yx_ntree_info;
int yx_ntree_build_fbm(const yx_ntree_desc* desc, yx_ntree** out, void* stream);
/* From a host dense volume (size^3 u8, x fastest) - RangeBuilder::build
 * (ntree/ntree/ntree.h:11-63) over the whole range. */
int yx_ntree_from_dense(const uint8_t* dense, int32_t size, int32_t levels, yx_ntree** out);
int yx_ntree_get_info(const yx_ntree* t, yx_ntree_info* info);
/* Host copies: root (R^3 u32: level-1 node / brick index, 0xFFFFFFFF empty),
 * nodes (4 u32 each: child mask lo, hi, first child index, 0), bricks (512 B
 * each, voxel x + 8y + 64z); any pointer may be NULL. */
int yx_ntree_download(const yx_ntree* t, uint32_t* root, uint32_t* nodes, uint8_t* bricks);
void yx_ntree_free(yx_ntree* t);
/* Per-pixel outputs of the n-tree cast (device pointers, any may be NULL). */
typedef struct yx_ntree_outputs {
    uint32_t* rgba;   /* front-to-back composite, float2uint packing            */
    float* t;         /* entry t of the first non-empty voxel, -1 if none       */
    int32_t* voxel;   /* 3 x int32: that voxel, (-1,-1,-1) if none              */
    uint32_t* nodes;  /* node records read                                      */
    uint32_t* voxels; /* brick voxels read                                      */
} yx_ntree_outputs;
/* castRay (reaction/trace/trace.cu:83-139) over the n-tree: pinhole rays of
 * `view` (eye position + InitRayDir basis dir0, du, dv as from
 * yx_camera_basis), empty cells skipped at every level, brick voxels crossed
 * by a 3D DDA and composited front to back through `palette` (256 RGBA8,
 * device pointer; entry 0 unused) until alpha > 0.99. Asynchronous. */
int yx_ntree_render(const yx_ntree* t, const float view[12], int32_t width, int32_t height,
                    const uint32_t* palette, const yx_ntree_outputs* out, void* stream);
/* Launch of the cast: persistent != 0 runs persistent waves over an atomic ray
 * queue whose lanes take new rays once refill_min (1..64, default 16) of them
 * wait; 0 (default) runs one lane per pixel, one 8x8 tile per wave. Both give
 * identical outputs.
 * Casts on one tree must not overlap (the queue lives in the tree). No
 * reference counterpart (launch shape). */
int yx_ntree_set_launch(yx_ntree* t, int32_t persistent, int32_t refill_min);

/* ---- device primitives ----
 * The scans, stable radix sort and select-distinct that the scene builder,
 * the re-layout, the n-tree builder, the brick pool and the compositor's
 * adjoint run, exposed so that they can be checked against a host sort or
 * scan. No reference counterpart: the reference's GPU paths would call a
 * library for these. Every pointer is device memory; each call allocates its
 * workspace, runs on `stream` and waits for it to finish. */
/* Stable LSD sort of n keys by bits [begin_bit, end_bit) (8-bit digits) into
 * keys_out, carrying ids (NULL: the positions 0..n-1) into ids_out (NULL:
 * keys only). n_dev (NULL: n) is a device count read by the kernels; only the
 * first min(*n_dev, n) entries are sorted and defined. n < 2^31. */
int yx_prim_sort_u32(const uint32_t* keys, const uint32_t* ids, uint64_t n, const uint32_t* n_dev,
                     int32_t begin_bit, int32_t end_bit, uint32_t* keys_out, uint32_t* ids_out, void* stream);
int yx_prim_sort_u64(const uint64_t* keys, const uint32_t* ids, uint64_t n, const uint32_t* n_dev,
                     int32_t begin_bit, int32_t end_bit, uint64_t* keys_out, uint32_t* ids_out, void* stream);
/* Exclusive prefix sums out[i] = in[0] + ... + in[i-1] (out may alias in);
 * *total (may be NULL) receives the sum of all n. u32 sums wrap. n < 2^32. */
int yx_prim_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total, void* stream);
int yx_prim_scan_u8(const uint8_t* in, uint32_t* out, uint64_t n, uint32_t* total, void* stream);
int yx_prim_scan_u32_u64(const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* total, void* stream);
/* *total = sum of the n items (u32 wrap). */
int yx_prim_reduce_u32(const uint32_t* in, uint64_t n, uint32_t* total, void* stream);
/* The distinct values of in[i] >> shift of a sorted array, in order, into
 * out; *count receives their number. */
int yx_prim_distinct_u64(const uint64_t* in, uint64_t* out, uint64_t n, int32_t shift, uint64_t* count,
                         void* stream);

const char* yx_last_error(void);
const char* yx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* YOXEL_H */
