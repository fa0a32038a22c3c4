"""ctypes wrapper of the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / baseline. Never a product path.
"""
import ctypes
import os
import pathlib
import subprocess
import threading

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
# YORACLE_LIB names another build of the restatement (the sanitized build of
# tests/sanitize/run.py); the default is the in-tree one.
LIB = pathlib.Path(os.environ["YORACLE_LIB"]) if os.environ.get("YORACLE_LIB") else HERE / "liboracle.so"

_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build(plant=0, out=None):
    """Compile the restatement with gcc (PLANT != 0 builds a planted-mistake
    variant for the pin tests). Returns the library path."""
    out = pathlib.Path(out) if out else LIB
    subprocess.run(["make", "-s", "-C", str(HERE), f"OUT={out}", f"PLANT={plant}", "-B"], check=True)
    return out


class Oracle:
    def __init__(self, path=None):
        path = pathlib.Path(path) if path else LIB
        if not path.exists():
            build(out=path)
        l = ctypes.CDLL(str(path))
        l.oracle_init_ray_dir.argtypes = [_f32p, _f32p, ctypes.c_float, ctypes.c_int, ctypes.c_int, _f32p]
        l.oracle_init_ray_dir.restype = None
        l.oracle_render_rows.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _f32p, _u32p, _i32p, _f32p, _u32p, _u8p, _u32p]
        l.oracle_render_rows.restype = ctypes.c_longlong
        l.oracle_render_rows_fmt.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, _f32p, ctypes.c_int, _u32p, _i32p, _f32p, _u32p, _u8p,
                                             _u32p]
        l.oracle_render_rows_fmt.restype = ctypes.c_longlong
        l.oracle_decode_c16n16.argtypes = [ctypes.c_uint32, _u32p, _f32p]
        l.oracle_decode_c16n16.restype = None
        l.oracle_trace_rays.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _f32p,
                                        _f32p, ctypes.c_int, _u32p, _i32p, _f32p, _i32p, _f32p, _u32p]
        l.oracle_trace_rays.restype = ctypes.c_int
        l.oracle_pixel_dirs.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _f32p]
        l.oracle_pixel_dirs.restype = None
        l.oracle_shade_rays.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _f32p, _f32p,
                                        ctypes.c_int, _f32p, ctypes.c_int, _u32p, _u32p, _u8p, _f32p, _u32p]
        l.oracle_shade_rays.restype = ctypes.c_int
        l.oracle_plant.restype = ctypes.c_int
        self.lib = l

    @staticmethod
    def _p(a, t):
        return a.ctypes.data_as(t) if a is not None else None

    def init_ray_dir(self, direction, up, fov, w, h):
        d = np.asarray(direction, np.float32)
        u = np.asarray(up, np.float32)
        out = np.zeros(9, np.float32)
        self.lib.oracle_init_ray_dir(self._p(d, _f32p), self._p(u, _f32p), float(fov), int(w), int(h),
                                     self._p(out, _f32p))
        return out

    def pixel_dirs(self, basis, w, h):
        b = np.asarray(basis, np.float32)
        out = np.zeros((h, w, 3), np.float32)
        self.lib.oracle_pixel_dirs(self._p(b, _f32p), w, h, self._p(out, _f32p))
        return out

    def render(self, nodes, root, pos, direction, up=(0, 0, 1), fov=70.0, w=64, h=64, light_mode=0,
               light=(0, 0, 1), rows=None, threads=1, payload=0):
        """RenderRect over rows (default all) into full-frame arrays. `threads`
        splits the rows into bands, as TreadedRenderer (cell/ppu_renderer.cpp:
        121-144) does. payload 1 reads leaves as colour16 + normal16. Returns a
        dict of outputs plus 'rays' traced."""
        nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1)
        nn = nodes.size // 10
        basis = self.init_ray_dir(direction, up, fov, w, h)
        p = np.asarray(pos, np.float32)
        lt = np.asarray(light, np.float32)
        out = {"hit_id": np.full((h, w), 0xFFFFFFFF, np.uint32),
               "voxel": np.full((h, w, 3), -1, np.int32),
               "t": np.full((h, w), -1.0, np.float32),
               "rgba": np.zeros((h, w), np.uint32),
               "shadow": np.full((h, w), 2, np.uint8),
               "visits": np.zeros((h, w), np.uint32)}
        y0, y1 = (0, h) if rows is None else rows
        bands = []
        step = -(-(y1 - y0) // threads)
        for i in range(threads):
            a, b = y0 + i * step, min(y1, y0 + (i + 1) * step)
            if a < b:
                bands.append((a, b))
        rays = [0] * len(bands)

        def run(k):
            a, b = bands[k]
            rays[k] = self.lib.oracle_render_rows_fmt(
                self._p(nodes, _u32p), nn, int(root), self._p(p, _f32p), self._p(basis, _f32p), w, h, a, b,
                int(light_mode), self._p(lt, _f32p), int(payload), self._p(out["hit_id"], _u32p),
                self._p(out["voxel"], _i32p), self._p(out["t"], _f32p), self._p(out["rgba"], _u32p),
                self._p(out["shadow"], _u8p), self._p(out["visits"], _u32p))

        if len(bands) == 1:
            run(0)
        else:
            ts = [threading.Thread(target=run, args=(k,)) for k in range(len(bands))]
            for t in ts:
                t.start()
            for t in ts:
                t.join()
        if any(r < 0 for r in rays):
            raise ValueError("oracle followed a child id out of range")
        out["rays"] = int(sum(rays))
        return out

    def trace_rays(self, nodes, root, pos, dirs, occlusion=False):
        nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1)
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = dirs.shape[0]
        if pos.shape[0] == 1 and n > 1:
            pos = np.ascontiguousarray(np.repeat(pos, n, axis=0))
        hit = np.zeros(n, np.uint32)
        vox = np.zeros((n, 3), np.int32)
        t = np.zeros(n, np.float32)
        lvl = np.zeros(n, np.int32)
        t1 = np.zeros((n, 3), np.float32)
        vis = np.zeros(n, np.uint32)
        rc = self.lib.oracle_trace_rays(self._p(nodes, _u32p), nodes.size // 10, int(root), n,
                                        self._p(pos, _f32p), self._p(dirs, _f32p), 1 if occlusion else 0,
                                        self._p(hit, _u32p), self._p(vox, _i32p), self._p(t, _f32p),
                                        self._p(lvl, _i32p), self._p(t1, _f32p), self._p(vis, _u32p))
        if rc < 0:
            raise ValueError("oracle followed a child id out of range")
        return {"hit_id": hit, "voxel": vox, "t": t, "level": lvl, "t1": t1, "visits": vis}

    def decode_c16n16(self, leaf):
        """A colour16 + normal16 payload: (RGBA8, unit normal float32[3])."""
        rgba = np.zeros(1, np.uint32)
        n = np.zeros(3, np.float32)
        self.lib.oracle_decode_c16n16(int(leaf) & 0xFFFFFFFF, self._p(rgba, _u32p), self._p(n, _f32p))
        return int(rgba[0]), n

    def shade_rays(self, nodes, root, pos, dirs, light_mode=0, light=(0, 0, 1), payload=0):
        """RenderRect's per-pixel work (trace, Lambert, shadow ray) for explicit
        rays: dict of hit_id, rgba, shadow, t, visits."""
        nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1)
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = dirs.shape[0]
        if pos.shape[0] == 1 and n > 1:
            pos = np.ascontiguousarray(np.repeat(pos, n, axis=0))
        lt = np.ascontiguousarray(light, np.float32).reshape(3)
        out = {"hit_id": np.zeros(n, np.uint32), "rgba": np.zeros(n, np.uint32),
               "shadow": np.zeros(n, np.uint8), "t": np.zeros(n, np.float32),
               "visits": np.zeros(n, np.uint32)}
        rc = self.lib.oracle_shade_rays(self._p(nodes, _u32p), nodes.size // 10, int(root), n,
                                        self._p(pos, _f32p), self._p(dirs, _f32p), int(light_mode),
                                        self._p(lt, _f32p), int(payload), self._p(out["hit_id"], _u32p),
                                        self._p(out["rgba"], _u32p), self._p(out["shadow"], _u8p),
                                        self._p(out["t"], _f32p), self._p(out["visits"], _u32p))
        if rc < 0:
            raise ValueError("oracle followed a child id out of range")
        return out

    def plant(self):
        return self.lib.oracle_plant()


class CompositeOracle:
    """FP64 reference of the brick-voxel compositing and its gradient
    (composite_oracle.c). Arrays: table int32 [G^3], params float32 [slots*B^3, 4],
    views float32 [n, 12]."""

    def __init__(self, path=None):
        path = pathlib.Path(path) if path else LIB
        if not path.exists():
            build(out=path)
        l = ctypes.CDLL(str(path))
        vp = ctypes.c_void_p
        l.comp_oracle_forward.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp]
        l.comp_oracle_forward.restype = None
        l.comp_oracle_grad.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp, vp,
                                       vp, vp, ctypes.c_int]
        l.comp_oracle_grad.restype = None
        l.comp_oracle_loss.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int64,
                                       ctypes.c_int, ctypes.c_double]
        l.comp_oracle_loss.restype = ctypes.c_double
        l.comp_oracle_ray_voxels.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_uint32, vp, ctypes.c_int]
        l.comp_oracle_ray_voxels.restype = ctypes.c_int
        l.oracle_plant.restype = ctypes.c_int
        self.lib = l

    @staticmethod
    def _prep(table, params, views):
        t = np.ascontiguousarray(table, np.int32)
        p = np.ascontiguousarray(params, np.float32).reshape(-1, 4)
        v = np.ascontiguousarray(views, np.float32).reshape(-1, 12)
        return t, p, v

    def _common(self, t, p, v, grid, brick, w, h):
        nslots = p.shape[0] // brick ** 3
        return (t.ctypes.data, p.ctypes.data, grid, brick, nslots, v.ctypes.data, v.shape[0], w, h)

    def forward(self, table, params, grid, brick, views, w, h, threads=1):
        t, p, v = self._prep(table, params, views)
        n = v.shape[0] * w * h
        rgb = np.zeros((n, 3)), np.zeros(n), np.zeros(n, np.uint32)
        rgb, trans, ns = rgb
        c = self._common(t, p, v, grid, brick, w, h)
        bands = _bands(n, threads)

        def run(a, b):
            self.lib.comp_oracle_forward(*c, a, b, rgb[a:].ctypes.data, trans[a:].ctypes.data, ns[a:].ctypes.data)
        _run_threads(run, bands)
        shp = (v.shape[0], h, w)
        return rgb.reshape(shp + (3,)), trans.reshape(shp), ns.reshape(shp)

    def grad(self, table, params, grid, brick, views, w, h, target, selected, max_rays=256, threads=1):
        """Loss sum |C - target|^2 and its FP64 gradient at the `selected` voxel
        slot indices (rows [d raw, d r, d g, d b]); also the rays that sample
        each selected voxel."""
        t, p, v = self._prep(table, params, views)
        n = v.shape[0] * w * h
        tgt = np.ascontiguousarray(target, np.float32).reshape(-1)
        sel = np.asarray(selected, np.int64)
        sel_map = np.full(p.shape[0], -1, np.int32)
        sel_map[sel] = np.arange(sel.size, dtype=np.int32)
        c = self._common(t, p, v, grid, brick, w, h)
        bands = _bands(n, threads)
        parts = [(np.zeros((sel.size, 4)), np.zeros(1), np.zeros((sel.size, max_rays), np.uint32),
                  np.zeros(sel.size, np.uint32)) for _ in bands]

        def run(a, b, k):
            g, L, rl, rc = parts[k]
            self.lib.comp_oracle_grad(*c, a, b, tgt.ctypes.data, sel_map.ctypes.data, g.ctypes.data, L.ctypes.data,
                                      rl.ctypes.data, rc.ctypes.data, max_rays)
        _run_threads(run, bands, with_index=True)
        grad = sum(pp[0] for pp in parts)
        loss = float(sum(pp[1][0] for pp in parts))
        rays = []
        for i in range(sel.size):
            lst = []
            for pp in parts:
                lst.extend(pp[2][i, :pp[3][i]].tolist())
            rays.append(np.array(sorted(set(lst)), np.uint32))
        return loss, grad, rays

    def loss_rays(self, table, params, grid, brick, views, w, h, target, rays, p_idx, p_comp, delta):
        t, p, v = self._prep(table, params, views)
        tgt = np.ascontiguousarray(target, np.float32).reshape(-1)
        r = np.ascontiguousarray(rays, np.uint32)
        c = self._common(t, p, v, grid, brick, w, h)
        return self.lib.comp_oracle_loss(*c, r.ctypes.data, r.size, tgt.ctypes.data, int(p_idx), int(p_comp),
                                         float(delta))

    def ray_voxels(self, table, grid, brick, num_slots, views, w, h, ray, max_out=4096):
        t = np.ascontiguousarray(table, np.int32)
        v = np.ascontiguousarray(views, np.float32).reshape(-1, 12)
        out = np.zeros(max_out, np.int64)
        n = self.lib.comp_oracle_ray_voxels(t.ctypes.data, grid, brick, num_slots, v.ctypes.data, v.shape[0], w,
                                            h, ray, out.ctypes.data, max_out)
        return out[:n]

    def plant(self):
        return self.lib.oracle_plant()


def _bands(n, threads):
    step = -(-n // max(1, threads))
    return [(a, min(n, a + step)) for a in range(0, n, step)]


def _run_threads(fn, bands, with_index=False):
    if len(bands) == 1:
        fn(*bands[0], 0) if with_index else fn(*bands[0])
        return
    ts = [threading.Thread(target=fn, args=(a, b, k) if with_index else (a, b)) for k, (a, b) in enumerate(bands)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()


class NTreeOracle:
    """Brick n-tree volume and ray cast restated on the CPU (ntree_oracle.c)."""

    class Desc(ctypes.Structure):
        _fields_ = [("size", ctypes.c_int32), ("levels", ctypes.c_int32), ("octaves", ctypes.c_int32),
                    ("frequency", ctypes.c_float), ("threshold", ctypes.c_float), ("scale", ctypes.c_float),
                    ("seed", ctypes.c_uint64)]

    def __init__(self, path=None):
        path = pathlib.Path(path) if path else LIB
        if not path.exists():
            build(out=path)
        l = ctypes.CDLL(str(path))
        vp = ctypes.c_void_p
        D = ctypes.POINTER(NTreeOracle.Desc)
        l.nt_perm.argtypes = [ctypes.c_uint64, vp]
        l.nt_density_box.argtypes = [D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, vp]
        l.nt_occupancy_sample.argtypes = [D, ctypes.c_int64, ctypes.c_uint64]
        l.nt_occupancy_sample.restype = ctypes.c_double
        l.nt_build_dense.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
        l.nt_render.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp]
        l.oracle_plant.restype = ctypes.c_int
        self.lib = l

    @staticmethod
    def desc(size, levels, octaves, frequency, threshold, scale, seed):
        return NTreeOracle.Desc(size, levels, octaves, frequency, threshold, scale, seed)

    def perm(self, seed):
        p = np.zeros(512, np.int32)
        self.lib.nt_perm(seed, p.ctypes.data)
        return p

    def density_box(self, d, x0, y0, z0, nx, ny, nz):
        out = np.zeros((nz, ny, nx), np.uint8)
        self.lib.nt_density_box(ctypes.byref(d), x0, y0, z0, nx, ny, nz, out.ctypes.data)
        return out

    def occupancy(self, d, n=200000, seed=1):
        return self.lib.nt_occupancy_sample(ctypes.byref(d), n, seed)

    def build_dense(self, dense, levels):
        """dense[z, y, x] u8 -> (root[R^3] u32, nodes[n, 4] u32, bricks[m, 512] u8)."""
        dense = np.ascontiguousarray(dense, np.uint8)
        size = dense.shape[0]
        nn, nb = ctypes.c_uint64(), ctypes.c_uint64()
        rc = self.lib.nt_build_dense(dense.ctypes.data, size, levels, None, None, None, 0, 0, ctypes.byref(nn),
                                     ctypes.byref(nb))
        if rc != 0:
            raise ValueError(f"bad n-tree shape: size {size}, levels {levels}")
        R = (size // 8) >> (2 * levels)
        root = np.zeros(R ** 3, np.uint32)
        nodes = np.zeros((max(1, nn.value), 4), np.uint32)
        bricks = np.zeros((max(1, nb.value), 512), np.uint8)
        self.lib.nt_build_dense(dense.ctypes.data, size, levels, root.ctypes.data, nodes.ctypes.data,
                                bricks.ctypes.data, nn.value, nb.value, ctypes.byref(nn), ctypes.byref(nb))
        return root, nodes[:nn.value], bricks[:nb.value]

    def render(self, root, nodes, bricks, size, levels, view, palette, w, h, ystep=1, threads=8):
        """Outputs for rows 0, ystep, 2 ystep, ... (other rows stay zero)."""
        root = np.ascontiguousarray(root, np.uint32)
        nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1)
        bricks = np.ascontiguousarray(bricks, np.uint8).reshape(-1)
        view = np.ascontiguousarray(view, np.float32)
        palette = np.ascontiguousarray(palette, np.uint32)
        out = {"rgba": np.zeros((h, w), np.uint32), "t": np.zeros((h, w), np.float32),
               "voxel": np.zeros((h, w, 3), np.int32), "nodes": np.zeros((h, w), np.uint32),
               "voxels": np.zeros((h, w), np.uint32)}
        self.lib.nt_render(root.ctypes.data, nodes.ctypes.data if nodes.size else None,
                           bricks.ctypes.data if bricks.size else None, size, levels, view.ctypes.data,
                           palette.ctypes.data, w, h, ystep, threads, out["rgba"].ctypes.data,
                           out["t"].ctypes.data, out["voxel"].ctypes.data, out["nodes"].ctypes.data,
                           out["voxels"].ctypes.data)
        return out

    def plant(self):
        return self.lib.oracle_plant()


LIGHT_DTYPE = np.dtype([("enabled", "<i4"), ("pos", "<f4", 3), ("diffuse", "<f4", 3), ("specular", "<f4", 3),
                        ("attenuation", "<f4", 3)])  # 52 bytes, as o_light / yx_light_params
MAX_LIGHTS = 4


def make_lights(specs):
    """specs: list of dicts with pos, diffuse, specular, attenuation (enabled
    unless enabled=False); returns the MAX_LIGHTS-entry light table."""
    lt = np.zeros(MAX_LIGHTS, LIGHT_DTYPE)
    for i, s in enumerate(specs):
        lt[i]["enabled"] = 1 if s.get("enabled", True) else 0
        for k in ("pos", "diffuse", "specular", "attenuation"):
            lt[i][k] = np.asarray(s[k], np.float32)
    return lt


class PipelineOracle:
    """The reference GPU pipeline (SVORenderer::Render, demo/SVORenderer.cpp:
    95-149) restated on the CPU: LOD colours, LOD trace (svo_oracle.c), BlurZ x5
    and ShadeSimple (ssna_oracle.c)."""

    def __init__(self, path=None):
        path = pathlib.Path(path) if path else LIB
        if not path.exists():
            build(out=path)
        l = ctypes.CDLL(str(path))
        vp, ci, cf = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        l.oracle_lod_colors.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp]
        l.oracle_lod_colors.restype = ci
        l.oracle_lod_render_rows.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp, ci, ci, ci, ci, cf,
                                             vp, vp, vp, vp, vp]
        l.oracle_lod_render_rows.restype = ctypes.c_longlong
        l.oracle_lod_trace_rays.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ci, vp, vp, cf, vp, vp, vp, vp]
        l.oracle_lod_trace_rays.restype = ci
        l.oracle_blur_kernel.argtypes = [vp]
        l.oracle_blur_kernel.restype = None
        l.oracle_pipeline_params.argtypes = [cf, ci, ci, cf, vp]
        l.oracle_pipeline_params.restype = None
        l.oracle_blur_z.argtypes = [ci, ci, vp, vp, cf, cf, vp, ci, ci]
        l.oracle_blur_z.restype = None
        l.oracle_view_frame.argtypes = [vp, vp, vp]
        l.oracle_view_frame.restype = None
        l.oracle_shade_simple.argtypes = [ci, ci, vp, vp, vp, vp, vp, cf, vp, ci, cf, ci, vp, ci, ci]
        l.oracle_shade_simple.restype = None
        l.oracle_init_ray_dir.argtypes = [vp, vp, cf, ci, ci, vp]
        l.oracle_init_ray_dir.restype = None
        l.oracle_plant.restype = ci
        self.lib = l

    def lod_colors(self, nodes, root):
        """Mean RGBA8 of the leaves under every node reachable from `root`,
        indexed by VoxNode id (0 elsewhere)."""
        nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1)
        nn = nodes.size // 10
        out = np.zeros(max(nn, 1), np.uint32)
        scratch = np.zeros(5 * max(nn, 1), np.uint64)
        stack = np.zeros(2 * max(nn, 1) + 2, np.uint32)
        rc = self.lib.oracle_lod_colors(nodes.ctypes.data, nn, int(root), out.ctypes.data, scratch.ctypes.data,
                                        stack.ctypes.data)
        if rc < 0:
            raise ValueError("bad child id or shared node")
        return out

    def params(self, fov, w, depth, coef):
        out = np.zeros(10, np.float32)
        self.lib.oracle_pipeline_params(float(fov), int(w), int(depth), float(coef), out.ctypes.data)
        return {"detail": out[0], "pixel_ang": out[1], "pix": out[2], "vox": out[3], "edge": out[4],
                "zlimits": out[5:10].copy()}

    def basis(self, direction, up, fov, w, h):
        d = np.asarray(direction, np.float32)
        u = np.asarray(up, np.float32)
        out = np.zeros(9, np.float32)
        self.lib.oracle_init_ray_dir(d.ctypes.data, u.ctypes.data, float(fov), int(w), int(h), out.ctypes.data)
        return out

    def view_frame(self, direction, up):
        d = np.asarray(direction, np.float32)
        u = np.asarray(up, np.float32)
        out = np.zeros(9, np.float32)
        self.lib.oracle_view_frame(d.ctypes.data, u.ctypes.data, out.ctypes.data)
        return out

    def lod_trace_rays(self, nodes, root, pos, dirs, detail):
        nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = dirs.shape[0]
        pos = np.ascontiguousarray(np.broadcast_to(np.asarray(pos, np.float32).reshape(-1, 3), (n, 3)))
        node = np.zeros(n, np.uint32)
        child = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        lvl = np.zeros(n, np.int32)
        rc = self.lib.oracle_lod_trace_rays(nodes.ctypes.data, nodes.size // 10, int(root), n, pos.ctypes.data,
                                            dirs.ctypes.data, float(detail), node.ctypes.data, child.ctypes.data,
                                            t.ctypes.data, lvl.ctypes.data)
        if rc < 0:
            raise ValueError("oracle followed a child id out of range")
        return {"node": node, "child": child, "t": t, "level": lvl}

    def lod_render(self, nodes, root, lod, pos, basis, w, h, detail, threads=8, rows=None):
        nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1)
        lod = np.ascontiguousarray(lod, np.uint32)
        p = np.asarray(pos, np.float32)
        b = np.asarray(basis, np.float32)
        out = {"t": np.full((h, w), -1.0, np.float32), "z": np.full((h, w), np.inf, np.float32),
               "albedo": np.zeros((h, w), np.uint32), "level": np.zeros((h, w), np.uint8),
               "visits": np.zeros((h, w), np.uint32)}
        y0, y1 = (0, h) if rows is None else rows
        rays = []

        def run(a, c):
            rays.append(self.lib.oracle_lod_render_rows(
                nodes.ctypes.data, nodes.size // 10, int(root), lod.ctypes.data, p.ctypes.data, b.ctypes.data, w, h,
                a, c, float(detail), out["t"].ctypes.data, out["z"].ctypes.data, out["albedo"].ctypes.data,
                out["level"].ctypes.data, out["visits"].ctypes.data))

        _run_threads(run, [(y0 + a, y0 + c) for a, c in _bands(y1 - y0, threads)])
        if any(r < 0 for r in rays):
            raise ValueError("oracle followed a child id out of range")
        return out

    def blur_kernel(self):
        k = np.zeros(25, np.float32)
        self.lib.oracle_blur_kernel(k.ctypes.data)
        return k

    def blur_z(self, z, zlimit, edge, kern=None, threads=8):
        z = np.ascontiguousarray(z, np.float32)
        h, w = z.shape
        kern = self.blur_kernel() if kern is None else np.ascontiguousarray(kern, np.float32)
        out = np.empty_like(z)
        _run_threads(lambda a, c: self.lib.oracle_blur_z(w, h, z.ctypes.data, out.ctypes.data, float(zlimit),
                                                         float(edge), kern.ctypes.data, a, c), _bands(h, threads))
        return out

    def shade(self, z, albedo, basis, eye, frame, pix, lights, ambient=0.1, show_normals=False, threads=8):
        z = np.ascontiguousarray(z, np.float32)
        albedo = np.ascontiguousarray(albedo, np.uint32)
        h, w = z.shape
        b = np.ascontiguousarray(basis, np.float32)
        e = np.ascontiguousarray(eye, np.float32)
        f = np.ascontiguousarray(frame, np.float32)
        lt = np.ascontiguousarray(lights, LIGHT_DTYPE)
        out = np.zeros((h, w), np.uint32)
        _run_threads(lambda a, c: self.lib.oracle_shade_simple(
            w, h, z.ctypes.data, albedo.ctypes.data, b.ctypes.data, e.ctypes.data, f.ctypes.data, float(pix),
            lt.ctypes.data, lt.size, float(ambient), 1 if show_normals else 0, out.ctypes.data, a, c),
            _bands(h, threads))
        return out

    def pipeline(self, nodes, root, depth, pos, direction, up=(0, 0, 1), fov=70.0, w=64, h=64, detail_coef=1.0,
                 ssna=True, lights=None, show_normals=False, threads=8, lod=None):
        """SVORenderer::Render restated: returns t, z (raw), zb (blurred, or z
        when ssna is off), albedo, level, visits, rgba."""
        if lod is None:
            lod = self.lod_colors(nodes, root)
        prm = self.params(fov, w, depth, detail_coef)
        basis = self.basis(direction, up, fov, w, h)
        out = self.lod_render(nodes, root, lod, pos, basis, w, h, prm["detail"], threads)
        zb = out["z"]
        if ssna:
            k = self.blur_kernel()
            for zl in prm["zlimits"]:
                zb = self.blur_z(zb, zl, prm["edge"], k, threads)
        out["zb"] = zb
        lights = make_lights([]) if lights is None else lights
        out["rgba"] = self.shade(zb, out["albedo"], basis, pos, self.view_frame(direction, up), prm["pix"],
                                 lights, 0.1, show_normals, threads)
        return out

    def plant(self):
        return self.lib.oracle_plant()


class BrickPoolOracle:
    """The brick pool's allocation policy restated in numpy (allocMap of
    reaction/voxtree/brickpool/BrickPool.cpp:21-62 enumerates the mapped
    bricks in request order; UpdateBrickPool of reaction/voxtree/density.cu:
    240-284 copies packed bricks into their slots). Policy: the listed bricks
    without a slot, in order of first appearance, get the lowest free slots in
    increasing order. Arrays: table int32 [(grid/brick)^3], params float32
    [slots * brick^3, 4]."""

    def __init__(self, table, params, brick):
        self.table = np.array(table, np.int32)
        self.params = np.array(params, np.float32).reshape(-1, 4)
        self.vox = brick ** 3
        self.cap = self.params.shape[0] // self.vox
        self.used = np.zeros(self.cap, bool)
        live = self.table[self.table >= 0]
        if live.size and (live.max() >= self.cap or np.unique(live).size != live.size):
            raise ValueError("slot used twice or beyond capacity")
        self.used[live] = True

    def alloc(self, bricks):
        bricks = np.asarray(bricks, np.int64)
        if bricks.size and (bricks.min() < 0 or bricks.max() >= self.table.size):
            raise ValueError("brick index out of range")
        _, first = np.unique(bricks, return_index=True)
        order = np.sort(first)
        new = [int(bricks[i]) for i in order if self.table[bricks[i]] < 0]
        free = np.nonzero(~self.used)[0]
        if len(new) > free.size:
            raise ValueError("not enough free slots")
        for b, s in zip(new, free[:len(new)]):
            self.table[b] = s
            self.used[s] = True
        return self.table[bricks].copy(), len(new)

    def release(self, bricks):
        n = 0
        for b in np.unique(np.asarray(bricks, np.int64)):
            s = self.table[b]
            if s >= 0:
                self.used[s] = False
                self.table[b] = -1
                n += 1
        return n

    def update(self, bricks, packed):
        packed = np.asarray(packed, np.float32).reshape(-1, self.vox, 4)
        for i, b in enumerate(np.asarray(bricks, np.int64)):
            s = self.table[b]
            if s >= 0:
                self.params[s * self.vox:(s + 1) * self.vox] = packed[i]


# ---- brick classification of a bit volume (reaction/voxtree) ---------------

VT_SOLID = 0x80000000
VT_EMPTY = 0x40000000


class VoxtreeOracle:
    """numpy restatement of the voxtree pipeline (reaction/voxtree/voxtree.py:80-95):
    CalcDensity (reaction/voxtree/density.cu:32-60), MarkBricks (:131-176),
    the host cumsum_ex of the column counts (voxtree.py:12-18, 88-89) and
    PackBricks (:210-221). It works on dense arrays: bits are unpacked to a
    boolean volume, summed over 2x2x2 blocks, and bricks are classified from
    min / max over their 5^3 windows. `window` and `nibble_order` exist so the
    tests can plant mistakes."""

    window = 5            # a brick's counts per axis: 4 + one of overlap
    nibble_order = None   # None: nibble n of a density word is z = 8w + n

    @staticmethod
    def pack_bits(vol):
        """bool [x, y, z] of side B -> uint32 words [B/128][B (y)][B (x)][4]."""
        B = vol.shape[0]
        v = np.ascontiguousarray(np.transpose(vol.reshape(B, B, B // 128, 4, 32), (2, 1, 0, 3, 4)))
        return np.packbits(v, axis=-1, bitorder="little").view(np.uint32).reshape(-1)

    @staticmethod
    def unpack_bits(words, B):
        w = np.ascontiguousarray(words, dtype=np.uint32).reshape(B // 128, B, B, 4)
        bits = np.unpackbits(w.view(np.uint8).reshape(B // 128, B, B, 4, 4), axis=-1, bitorder="little")
        bits = bits.reshape(B // 128, B, B, 4, 32)  # [s][y][x][layer][bit]
        return np.transpose(bits, (2, 1, 0, 3, 4)).reshape(B, B, B).astype(bool)

    def counts(self, words, B):
        """2x2x2 block sums (CalcDensity) as uint8 [x, y, z] of side B / 2."""
        v = self.unpack_bits(words, B).astype(np.uint8)
        V = B // 2
        return v.reshape(V, 2, V, 2, V, 2).sum(axis=(1, 3, 5), dtype=np.uint8)

    def pack_density(self, c):
        """uint8 counts [x, y, z] -> density words [V (y)][V (x)][V/8]."""
        V = c.shape[0]
        n = np.transpose(c, (1, 0, 2)).reshape(V, V, V // 8, 8).astype(np.uint32)
        order = self.nibble_order if self.nibble_order is not None else range(8)
        out = np.zeros((V, V, V // 8), np.uint32)
        for k, src in enumerate(order):
            out |= n[..., src] << np.uint32(4 * k)
        return out.reshape(-1)

    @staticmethod
    def unpack_density(dens, V):
        d = np.ascontiguousarray(dens, dtype=np.uint32).reshape(V, V, V // 8, 1)
        nib = (d >> (np.arange(8, dtype=np.uint32) * 4)) & 0xF  # [y][x][w][n]
        return np.transpose(nib.reshape(V, V, V).astype(np.uint8), (1, 0, 2))  # [x, y, z]

    def density(self, words, B):
        return self.pack_density(self.counts(words, B))

    def mark(self, dens, V):
        """(brick_state uint32 [G*G*G], column_count uint32 [G*G]) from density words."""
        c = self.unpack_density(dens, V)
        G = V // 4
        k = self.window
        pad = np.zeros((V + k, V + k, V + k), np.uint8)  # counts past the volume are 0
        pad[:V, :V, :V] = c
        win = np.lib.stride_tricks.sliding_window_view(pad, (k, k, k))[::4, ::4, ::4][:G, :G, :G]
        lo = win.min(axis=(3, 4, 5))
        hi = win.max(axis=(3, 4, 5))
        st = np.where(lo == 8, VT_SOLID, np.where(hi == 0, VT_EMPTY, 0)).astype(np.uint32)  # [bx, by, bz]
        st = np.transpose(st, (1, 0, 2)).reshape(G * G, G)  # [col = bx + by G][bz]
        mixed = (st == 0).astype(np.uint32)
        below = np.cumsum(mixed, axis=1, dtype=np.uint32) - mixed
        return (st | below).reshape(-1), mixed.sum(axis=1, dtype=np.uint32)

    @staticmethod
    def column_start(col_count):
        """Exclusive prefix sum and total (cumsum_ex, voxtree.py:12-18)."""
        c = np.asarray(col_count, np.uint64)
        s = np.cumsum(c) - c
        return s.astype(np.uint32), int(c.sum())

    @staticmethod
    def pack(state, col_start, V):
        G = V // 4
        st = np.asarray(state, np.uint32).reshape(G * G, G)
        col, z = np.nonzero((st & np.uint32(VT_SOLID | VT_EMPTY)) == 0)
        idx = col_start[col] + st[col, z]
        out = np.zeros(len(col), np.uint32)
        out[idx] = (col % G).astype(np.uint32) | ((col // G).astype(np.uint32) << 8) | (z.astype(np.uint32) << 16)
        return out

    def classify(self, words, B):
        V = B // 2
        dens = self.density(words, B)
        state, cnt = self.mark(dens, V)
        start, total = self.column_start(cnt)
        return dict(density=dens, brick_state=state, column_count=cnt, column_start=start, total=total,
                    packed=self.pack(state, start, V))
