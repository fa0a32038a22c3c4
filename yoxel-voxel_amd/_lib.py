"""ctypes binding of libyoxel (include/yoxel.h).

Loads the in-tree library built by build.py and fails loudly when it is
missing: there is no fallback path.
"""
import ctypes
import os
import pathlib

_PKG = pathlib.Path(__file__).resolve().parent
# YOXEL_LIB names another build of the same library (the host-sanitized
# build that tests/sanitize/run.py makes); the default is the in-tree build.
LIB_PATH = pathlib.Path(os.environ["YOXEL_LIB"]) if os.environ.get("YOXEL_LIB") else _PKG / "lib" / "libyoxel.so"

YX_OK = 0
YX_ERR_INVALID_ARG = -1
YX_ERR_OOM = -2
YX_ERR_IO = -3
YX_ERR_FORMAT = -4
YX_ERR_HIP = -5
YX_ERR_NO_SCENE = -6
YX_ERR_RANGE = -7

YX_EMPTY_NODE = 0x80000000
YX_FULL_NODE = 0x80000001
YX_MISS = 0xFFFFFFFF

YX_LIGHT_HEADLIGHT = 0
YX_LIGHT_SUN = 1
YX_LIGHT_SUN_SHADOW = 2
YX_PAYLOAD_RGBA8 = 0
YX_PAYLOAD_C16N16 = 1

_ERRNAMES = {
    YX_ERR_INVALID_ARG: "YX_ERR_INVALID_ARG",
    YX_ERR_OOM: "YX_ERR_OOM",
    YX_ERR_IO: "YX_ERR_IO",
    YX_ERR_FORMAT: "YX_ERR_FORMAT",
    YX_ERR_HIP: "YX_ERR_HIP",
    YX_ERR_NO_SCENE: "YX_ERR_NO_SCENE",
    YX_ERR_RANGE: "YX_ERR_RANGE",
}


class YoxelError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("depth", ctypes.c_int32),
        ("num_spheres", ctypes.c_int32),
        ("radius_min", ctypes.c_float),
        ("radius_max", ctypes.c_float),
        ("heightfield", ctypes.c_int32),
        ("hf_octaves", ctypes.c_int32),
        ("hf_amplitude", ctypes.c_float),
        ("hf_base", ctypes.c_float),
        ("seed", ctypes.c_uint64),
    ]


class SceneInfo(ctypes.Structure):
    _fields_ = [
        ("root", ctypes.c_uint32),
        ("num_nodes", ctypes.c_uint32),
        ("num_leaves", ctypes.c_uint64),
        ("depth", ctypes.c_int32),
    ]


class Outputs(ctypes.Structure):
    _fields_ = [
        ("hit_id", ctypes.c_void_p),
        ("voxel", ctypes.c_void_p),
        ("t", ctypes.c_void_p),
        ("rgba", ctypes.c_void_p),
        ("shadow", ctypes.c_void_p),
        ("visits", ctypes.c_void_p),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("tile_w", ctypes.c_int32),
        ("tile_h", ctypes.c_int32),
        ("refill_min", ctypes.c_int32),
        ("queues", ctypes.c_int32),
        ("window", ctypes.c_int32),
        ("waves_per_cu", ctypes.c_int32),
        ("lds_levels", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneTrial(ctypes.Structure):
    _fields_ = [("tuning", Tuning), ("ms", ctypes.c_float)]


YX_MAX_LIGHTS = 4
YX_BUILD_GROW = 0
YX_BUILD_CLEAR = 1
YX_RANGE_IN_PLACE = 0
YX_RANGE_REBUILD = 1
YX_RANGE_REBUILD_GPU = 2


YX_SOURCE_ISO = 1
YX_SOURCE_RAW = 2


class VolumeSource(ctypes.Structure):
    """yx_volume_source (BuildRange with an IsoSource / RawSource)."""
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("level", ctypes.c_int32),
        ("pos", ctypes.c_int32 * 3),
        ("size", ctypes.c_int32 * 3),
        ("kind", ctypes.c_int32),
        ("iso", ctypes.c_int32),
        ("inside", ctypes.c_int32),
        ("color", ctypes.c_uint32),
        ("data", ctypes.c_void_p),
    ]


class SphereEdit(ctypes.Structure):
    """yx_sphere_edit (BuildRange with a SphereSource)."""
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("level", ctypes.c_int32),
        ("pos", ctypes.c_float * 3),
        ("radius", ctypes.c_float),
        ("color", ctypes.c_uint32),
    ]


class LightParams(ctypes.Structure):
    """yx_light_params (LightParams, demo/Demo.cpp:141-146)."""
    _fields_ = [
        ("enabled", ctypes.c_int32),
        ("pos", ctypes.c_float * 3),
        ("diffuse", ctypes.c_float * 3),
        ("specular", ctypes.c_float * 3),
        ("attenuation", ctypes.c_float * 3),
    ]


class PipelineOutputs(ctypes.Structure):
    _fields_ = [
        ("rgba", ctypes.c_void_p),
        ("z", ctypes.c_void_p),
        ("t", ctypes.c_void_p),
        ("albedo", ctypes.c_void_p),
        ("level", ctypes.c_void_p),
        ("visits", ctypes.c_void_p),
    ]


class BatchJob(ctypes.Structure):
    _fields_ = [
        ("scene", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("pos", ctypes.c_float * 3),
        ("dir", ctypes.c_float * 3),
        ("up", ctypes.c_float * 3),
        ("fov", ctypes.c_float),
        ("light_mode", ctypes.c_int32),
        ("light", ctypes.c_float * 3),
    ]


class NTreeDesc(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int32), ("levels", ctypes.c_int32), ("octaves", ctypes.c_int32),
                ("frequency", ctypes.c_float), ("threshold", ctypes.c_float), ("scale", ctypes.c_float),
                ("seed", ctypes.c_uint64)]


class NTreeInfo(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int32), ("levels", ctypes.c_int32), ("root_side", ctypes.c_int32),
                ("num_nodes", ctypes.c_uint64), ("num_bricks", ctypes.c_uint64), ("bytes", ctypes.c_uint64)]


class NTreeOutputs(ctypes.Structure):
    _fields_ = [("rgba", ctypes.c_void_p), ("t", ctypes.c_void_p), ("voxel", ctypes.c_void_p),
                ("nodes", ctypes.c_void_p), ("voxels", ctypes.c_void_p)]


class VolumeDesc(ctypes.Structure):
    _fields_ = [
        ("grid", ctypes.c_int32),
        ("brick", ctypes.c_int32),
        ("occupancy", ctypes.c_float),
        ("octaves", ctypes.c_int32),
        ("frequency", ctypes.c_float),
        ("seed", ctypes.c_uint64),
    ]


class VolumeInfo(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_int32), ("brick", ctypes.c_int32), ("num_slots", ctypes.c_uint32)]


class VolumeDev(ctypes.Structure):
    _fields_ = [
        ("table", ctypes.c_void_p),
        ("params", ctypes.c_void_p),
        ("grid", ctypes.c_int32),
        ("brick", ctypes.c_int32),
        ("num_slots", ctypes.c_uint32),
    ]


YX_KERNEL_PER_PIXEL = 0
YX_KERNEL_PERSISTENT = 1
YX_KERNEL_REPLAY = 2

_P = ctypes.c_void_p
_F3 = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes); every symbol include/yoxel.h declares.
SIGNATURES = {
    "yx_scene_load_vox": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "yx_scene_save_vox": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "yx_scene_from_nodes": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_P)]),
    "yx_scene_build_synthetic": (ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.POINTER(_P)]),
    "yx_scene_build_synthetic_gpu": (ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.POINTER(_P), _P]),
    "yx_scene_get_info": (ctypes.c_int, [_P, ctypes.POINTER(SceneInfo)]),
    "yx_scene_nodes": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32)]),
    "yx_scene_free": (None, [_P]),
    "yx_renderer_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "yx_renderer_destroy": (None, [_P]),
    "yx_renderer_set_scene": (ctypes.c_int, [_P, _P]),
    "yx_renderer_set_scene_host": (ctypes.c_int, [_P, _P]),
    "yx_renderer_set_view_pos": (ctypes.c_int, [_P, _F3]),
    "yx_renderer_set_view_dir": (ctypes.c_int, [_P, _F3]),
    "yx_renderer_set_view_up": (ctypes.c_int, [_P, _F3]),
    "yx_renderer_set_resolution": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "yx_renderer_get_resolution": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32),
                                                  ctypes.POINTER(ctypes.c_int32)]),
    "yx_renderer_set_fov": (ctypes.c_int, [_P, ctypes.c_float]),
    "yx_renderer_set_light": (ctypes.c_int, [_P, _F3, ctypes.c_int32]),
    "yx_renderer_ray_basis": (ctypes.c_int, [_P, _F3]),
    "yx_render": (ctypes.c_int, [_P, ctypes.POINTER(Outputs), _P]),
    "yx_render_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(Outputs), _P]),
    "yx_renderer_set_payload": (ctypes.c_int, [_P, ctypes.c_int32]),
    "yx_scene_decode_hits": (ctypes.c_int, [_P, _P, ctypes.c_uint64, ctypes.c_int32, _P, _P]),
    "yx_render_frame": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "yx_renderer_scene_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "yx_renderer_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "yx_renderer_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "yx_renderer_autotune": (ctypes.c_int, [_P, ctypes.POINTER(Outputs), ctypes.c_int32, ctypes.POINTER(TuneTrial),
                                            ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _P]),
    "yx_brick_pool_create": (ctypes.c_int, [ctypes.POINTER(VolumeDev), _P, ctypes.POINTER(_P)]),
    "yx_brick_pool_destroy": (None, [_P]),
    "yx_brick_pool_alloc": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P, ctypes.POINTER(ctypes.c_uint32), _P]),
    "yx_brick_pool_release": (ctypes.c_int, [_P, _P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), _P]),
    "yx_brick_pool_update": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P, _P]),
    "yx_brick_pool_stats": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "yx_voxtree_calc_density": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P]),
    "yx_voxtree_mark_bricks": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _P]),
    "yx_voxtree_pack_bricks": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, ctypes.c_uint32, _P]),
    "yx_voxtree_classify": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _P, _P, _P, ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_uint32), _P]),
    "yx_scene_fill_interior": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "yx_scene_build_range_volume": (ctypes.c_int, [_P, ctypes.POINTER(VolumeSource), ctypes.c_int32, _P]),
    "yx_scene_build_range": (ctypes.c_int, [_P, ctypes.POINTER(SphereEdit), ctypes.c_uint32, ctypes.c_int32, _P]),
    "yx_scene_write_nodes": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32]),
    "yx_scene_set_root": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "yx_renderer_update_scene": (ctypes.c_int, [_P, _P, _P]),
    "yx_renderer_count_changed_pages": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_uint32),
                                                       ctypes.POINTER(ctypes.c_uint64)]),
    "yx_renderer_last_update": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)]),
    "yx_renderer_layout": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "yx_renderer_set_detail_coef": (ctypes.c_int, [_P, ctypes.c_float]),
    "yx_renderer_get_detail_coef": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float)]),
    "yx_renderer_set_ssna": (ctypes.c_int, [_P, ctypes.c_int32]),
    "yx_renderer_get_ssna": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "yx_renderer_set_show_normals": (ctypes.c_int, [_P, ctypes.c_int32]),
    "yx_renderer_set_point_light": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(LightParams)]),
    "yx_render_pipeline": (ctypes.c_int, [_P, ctypes.POINTER(PipelineOutputs), _P]),
    "yx_batch_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "yx_batch_destroy": (None, [_P]),
    "yx_batch_set_scenes": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.c_int32]),
    "yx_batch_set_jobs": (ctypes.c_int, [_P, ctypes.POINTER(BatchJob), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "yx_batch_job_offset": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64)]),
    "yx_batch_set_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "yx_batch_get_tuning": (ctypes.c_int, [_P, ctypes.POINTER(Tuning)]),
    "yx_batch_set_graph": (ctypes.c_int, [_P, ctypes.c_int32]),
    "yx_batch_set_payload": (ctypes.c_int, [_P, ctypes.c_int32]),
    "yx_render_batch": (ctypes.c_int, [_P, ctypes.POINTER(Outputs), _P]),
    "yx_batch_autotune": (ctypes.c_int, [_P, ctypes.POINTER(Outputs), ctypes.c_int32, ctypes.POINTER(TuneTrial),
                                         ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _P]),
    "yx_batch_scene_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "yx_volume_build_synthetic": (ctypes.c_int, [ctypes.POINTER(VolumeDesc), ctypes.POINTER(_P)]),
    "yx_volume_get_info": (ctypes.c_int, [_P, ctypes.POINTER(VolumeInfo)]),
    "yx_volume_arrays": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "yx_volume_free": (None, [_P]),
    "yx_camera_basis": (ctypes.c_int, [_F3, _F3, _F3, ctypes.c_float, ctypes.c_int32, ctypes.c_int32, _F3]),
    "yx_compositor_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "yx_compositor_destroy": (None, [_P]),
    "yx_composite_forward": (ctypes.c_int, [_P, ctypes.POINTER(VolumeDev), _P, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32, _P, _P, _P, _P]),
    "yx_composite_backward": (ctypes.c_int, [_P, ctypes.POINTER(VolumeDev), _P, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, _P, _P, _P, _P, _P]),
    "yx_compositor_last_stats": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64),
                                                ctypes.POINTER(ctypes.c_uint64)]),
    "yx_compositor_reserve": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "yx_ntree_build_fbm": (ctypes.c_int, [ctypes.POINTER(NTreeDesc), ctypes.POINTER(_P), _P]),
    "yx_ntree_from_dense": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_P)]),
    "yx_ntree_get_info": (ctypes.c_int, [_P, ctypes.POINTER(NTreeInfo)]),
    "yx_ntree_download": (ctypes.c_int, [_P, _P, _P, _P]),
    "yx_ntree_free": (None, [_P]),
    "yx_ntree_set_launch": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "yx_prim_sort_u32": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P]),
    "yx_prim_sort_u64": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P]),
    "yx_prim_scan_u32": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, _P]),
    "yx_prim_scan_u8": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, _P]),
    "yx_prim_scan_u32_u64": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, _P]),
    "yx_prim_reduce_u32": (ctypes.c_int, [_P, ctypes.c_uint64, _P, _P]),
    "yx_prim_distinct_u64": (ctypes.c_int, [_P, _P, ctypes.c_uint64, ctypes.c_int32, _P, _P]),
    "yx_ntree_render": (ctypes.c_int, [_P, _F3, ctypes.c_int32, ctypes.c_int32, _P, ctypes.POINTER(NTreeOutputs),
                                       _P]),
    "yx_last_error": (ctypes.c_char_p, []),
    "yx_version": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """The loaded libyoxel (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python yoxel-voxel_amd/build.py` "
                "(or __graft_entry__.build()); there is no fallback path")
        l = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(rc):
    if rc != YX_OK:
        msg = lib().yx_last_error()
        raise YoxelError(rc, msg.decode() if msg else "")
    return rc


def f3(v):
    a = (ctypes.c_float * 3)(*[float(x) for x in v])
    return a
