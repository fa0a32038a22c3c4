"""Benchmark: SVO ray cast throughput on one MI355X.

Workload (default): large_1080p - depth-11 octree (2048^3; 512 spheres of radius
16-128 voxels united with a 5-octave fBm heightfield, ~32M surface leaves),
1920x1080 primary rays plus one shadow ray per lit hit toward a fixed
directional light, 32 orbit frames. One step = one frame (primary + shadow
rays, shading, all per-pixel outputs written to HBM).

Prints one JSON line. `value` = (primary + shadow rays traced) / wall time of
the timed steps. `roofline` covers the frame kernel: algorithmic bytes per
frame = 8 B per node record read (exact visit counts from the kernel) + 4 B per
hit leaf payload + 25 B of outputs per pixel, divided by the kernel's average
duration from HIP events on its stream.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import __graft_entry__ as entry

HEADLINE_METRIC = "Mrays/s (1080p, depth-11 SVO, primary+shadow) + %HBM peak; fwd+bwd Mrays/s"
HBM_PEAK_GBS = 8000.0
OUT_BYTES_PER_PIXEL = 4 + 12 + 4 + 4 + 1  # hit_id, voxel, t, rgba, shadow
NTREE_NODE_BYTES = 12  # minimal 64-ary node: 64-bit child mask + 32-bit first-child index


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--workload", default="large_1080p", choices=["large_1080p", "plain_512"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=8, help="frames in the CPU baseline sample")
    ap.add_argument("--no-composite", action="store_true", help="skip the fwd+bwd compositing measurement")
    ap.add_argument("--composite-steps", type=int, default=8)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-launch measurement")
    ap.add_argument("--no-brick", action="store_true", help="skip the brick n-tree measurement")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain_512 measurement")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="skip the shading-pipeline measurement (LOD trace, SSNA blur, Phong shade)")
    ap.add_argument("--no-update", action="store_true", help="skip the scene-sync measurement")
    ap.add_argument("--no-voxtree", action="store_true", help="skip the brick-classification measurement")
    ap.add_argument("--only", choices=["composite", "batch", "brick", "plain", "pipeline", "update", "pool",
                                       "voxtree"],
                    default=None,
                    help="run one secondary workload alone and print its object (profiling)")
    ap.add_argument("--brick-frames", type=int, default=8)
    ap.add_argument("--autotune", action="store_true",
                    help="time the headline workload with the autotuner's launch shape instead of the "
                         "library default (the committed traffic profile is for the default)")
    ap.add_argument("--batch-frames", type=int, default=4)
    ap.add_argument("--batch-tune-cache", default=None,
                    help="JSON file for the batch autotuner's result (default: tune every run)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, copy 256 MiB device-to-device 3 times (a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X and no multi-GPU run is "
              f"made (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    import torch

    yx = entry.load_package()
    wl = yx.workloads
    if args.only:
        # one secondary workload alone (for profiling); prints that leg's object
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        if args.only == "composite":
            leg = composite_bench(yx, torch, dev, args.composite_steps, not args.no_cpu_baseline)
        elif args.only == "plain":
            leg = plain_bench(yx, torch, not args.no_cpu_baseline)
        elif args.only == "batch":
            leg = batch_bench(yx, torch, args.batch_frames, args.batch_tune_cache)
        elif args.only == "pool":
            leg = pool_bench(yx, torch, dev)
        elif args.only == "voxtree":
            leg = voxtree_bench(yx, torch, not args.no_cpu_baseline)
        elif args.only == "update":
            leg = update_bench(yx, torch, yx.SVOData.build_synthetic_gpu(**wl.LARGE_1080P))
        elif args.only == "pipeline":
            leg = pipeline_bench(yx, torch, yx.SVOData.build_synthetic(**wl.LARGE_1080P), wl.orbit_cameras(32, seed=0),
                                 not args.no_cpu_baseline)
        else:
            leg = brick_bench(yx, torch, args.brick_frames, not args.no_cpu_baseline)
        if args.copy_calibration:
            copy_calibration(torch, dev)
        print(json.dumps(leg), flush=True)
        return
    if args.workload == "large_1080p":
        desc, w, h, nframes, mode = wl.LARGE_1080P, 1920, 1080, 32, yx.YX_LIGHT_SUN_SHADOW
    else:
        desc, w, h, nframes, mode = wl.PLAIN_512, 512, 512, 8, yx.YX_LIGHT_HEADLIGHT
    t0 = time.time()
    scene = yx.SVOData.build_synthetic(**desc)
    info = scene.info()
    build_s = time.time() - t0
    print(f"[bench] scene {args.workload}: {info['num_nodes']} nodes, {info['num_leaves']} leaves, "
          f"built in {build_s:.1f} s", file=sys.stderr, flush=True)
    cams = wl.orbit_cameras(nframes, seed=0)

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    r = yx.SVORenderer()
    r.SetScene(scene)
    r.SetFOV(60.0)
    r.SetResolution(w, h)
    r.SetLight(wl.SUN, mode)
    outs = {"hit_id": torch.empty((h, w), dtype=torch.int32, device=dev),
            "voxel": torch.empty((h, w, 3), dtype=torch.int32, device=dev),
            "t": torch.empty((h, w), dtype=torch.float32, device=dev),
            "rgba": torch.empty((h, w), dtype=torch.int32, device=dev),
            "shadow": torch.empty((h, w), dtype=torch.uint8, device=dev)}
    visits = torch.empty((h, w), dtype=torch.int32, device=dev)

    # untimed counting pass: rays and algorithmic bytes of every frame
    rays_f, bytes_f, hits_f, unlit_f = [], [], [], []
    for pos, d in cams:
        r.SetViewPos(pos)
        r.SetViewDir(d)
        r.Render(visits=visits, **outs)
        torch.cuda.synchronize()
        sh = outs["shadow"]
        nshadow = int((sh < 2).sum().item())
        nhit = int((outs["hit_id"] != -1).sum().item())
        nvis = int(visits.to(torch.int64).sum().item())
        rays_f.append(w * h + nshadow)
        bytes_f.append(8 * nvis + 4 * nhit + OUT_BYTES_PER_PIXEL * w * h)
        hits_f.append(nhit)
        unlit_f.append(nhit - nshadow if mode == yx.YX_LIGHT_SUN_SHADOW else 0)
    print(f"[bench] counting pass done: mean rays/frame {np.mean(rays_f):.0f}", file=sys.stderr, flush=True)

    stream = torch.cuda.current_stream()
    tuning = r.get_tuning()
    if args.autotune:
        r.SetViewPos(cams[0][0])
        r.SetViewDir(cams[0][1])
        t0 = time.time()
        tuning, trials = r.autotune(outs, reps=3)
        print(f"[bench] renderer autotune: {len(trials)} trials in {time.time() - t0:.1f} s, best {tuning}",
              file=sys.stderr, flush=True)

    def step(k):
        pos, d = cams[k % nframes]
        r.SetViewPos(pos)
        r.SetViewDir(d)
        r.Render(stream=stream, **outs)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        step(k)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    t_end = time.perf_counter()
    elapsed = t_end - t_start
    kern_ms = [a.elapsed_time(b) for a, b in ev]
    total_rays = sum(rays_f[k % nframes] for k in range(args.steps))
    total_bytes = sum(bytes_f[k % nframes] for k in range(args.steps))
    value = total_rays / elapsed / 1e6
    # the roofline covers exactly the timed frames: their mean algorithmic
    # bytes over their mean kernel time
    avg_kernel_s = float(np.mean(kern_ms)) / 1e3
    timed_bytes = total_bytes / args.steps
    achieved = timed_bytes / avg_kernel_s / 1e9

    # every orbit frame once more, whatever --steps is: the per-frame kernel
    # times and the all-frames mean
    ev_all = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(nframes)]
    for k in range(nframes):
        ev_all[k][0].record(stream)
        step(k)
        ev_all[k][1].record(stream)
    torch.cuda.synchronize()
    all_ms = [a.elapsed_time(b) for a, b in ev_all]
    all_kernel_s = float(np.mean(all_ms)) / 1e3
    all_bytes = float(np.mean(bytes_f))
    all_frames = {"frames": nframes, "kernel_ms": round(all_kernel_s * 1e3, 4),
                  "algorithmic_bytes_per_launch": int(all_bytes),
                  "achieved": round(all_bytes / all_kernel_s / 1e9, 2),
                  "frac": round(all_bytes / all_kernel_s / 1e9 / HBM_PEAK_GBS, 5),
                  "rays_per_frame": round(float(np.mean(rays_f)), 1),
                  "value": round(float(np.sum(rays_f)) / (all_kernel_s * nframes) / 1e6, 3)}

    if args.copy_calibration:
        copy_calibration(torch, dev)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, cams, w, h, mode, wl.SUN, args.cpu_frames)
    fwd_bwd = None if args.no_composite else composite_bench(yx, torch, dev, args.composite_steps,
                                                            not args.no_cpu_baseline)
    plain = None if args.no_plain else plain_bench(yx, torch, not args.no_cpu_baseline)
    batched = None if args.no_batch else batch_bench(yx, torch, args.batch_frames, args.batch_tune_cache)
    brick = None if args.no_brick else brick_bench(yx, torch, args.brick_frames, not args.no_cpu_baseline)
    pipeline = (None if args.no_pipeline or args.workload != "large_1080p"
                else pipeline_bench(yx, torch, scene, cams, not args.no_cpu_baseline))
    update = None if args.no_update or args.workload != "large_1080p" else update_bench(yx, torch, scene)
    pool = None if args.no_composite else pool_bench(yx, torch, dev)
    voxtree = None if args.no_voxtree else voxtree_bench(yx, torch, not args.no_cpu_baseline)

    line = {
        "metric": HEADLINE_METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "depth": desc["depth"], "resolution": [w, h],
                   "frames": nframes, "rays": "primary+shadow" if mode == 2 else "primary",
                   "nodes": info["num_nodes"], "leaves": info["num_leaves"],
                   "mean_rays_per_frame": round(float(np.mean(rays_f)), 1),
                   "mean_hits_per_frame": round(float(np.mean(hits_f)), 1),
                   # hits whose face turns away from the light (Lambert term 0) get no shadow ray
                   "mean_hits_without_shadow_ray": round(float(np.mean(unlit_f)), 1), "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(args.workload),
                     "kernel": f"stack_kernel<{tuning['block']}, {tuning['window']}>",
                     "kernel_ms": round(avg_kernel_s * 1e3, 4),
                     "algorithmic_bytes_per_launch": int(timed_bytes),
                     "frames": f"the timed steps (orbit frame k mod {nframes})",
                     "all_frames": all_frames},
        "frame_ms": {"median": round(float(np.median(kern_ms)), 4), "min": round(float(np.min(kern_ms)), 4),
                     "max": round(float(np.max(kern_ms)), 4)},
        "cpu_baseline": cpu,
        "plain": plain,
        "fwd_bwd": fwd_bwd,
        "batched": batched,
        "brick": brick,
        "pipeline": pipeline,
        "update": update,
        "brick_pool": pool,
        "voxtree": voxtree,
    }
    print(json.dumps(line), flush=True)


def copy_calibration(torch, dev):
    """Three 256 MiB device-to-device copies: a kernel of known traffic for
    scaling rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def composite_bench(yx, torch, dev, steps, with_cpu=False):
    """fwd+bwd of the differentiable compositing workload composite_512: 512^3
    grid as 8^3 bricks (~10 % allocated), 16 views of 512x512 rays, L2 loss
    against targets rendered with raw densities perturbed by N(0, 0.1). One
    step = forward, dL/dC, adjoint with the sorted-key gradient reduction.
    value = rays per second (each ray gets one forward and one backward pass)."""
    comp_mod = yx.composite
    cfg = comp_mod.COMPOSITE_512
    grid, brick, w, h, nv = cfg["grid"], cfg["brick"], 512, 512, 16
    vol = comp_mod.Volume(**cfg)
    table, params = vol.to(dev)
    views = torch.from_numpy(comp_mod.orbit_views(nv, 60.0, w, h, seed=0)).to(dev)
    rng = np.random.default_rng(1)
    p2 = vol.params.copy()
    p2[:, 0] += rng.normal(0.0, 0.1, size=p2.shape[0]).astype(np.float32)
    comp = comp_mod.Compositor()
    target, _, _ = comp.forward(table, torch.from_numpy(p2).to(dev), grid, brick, views, w, h)
    stream = torch.cuda.current_stream()

    def step():
        rgb, trans, ns = comp.forward(table, params, grid, brick, views, w, h)
        g = 2.0 * (rgb - target)
        return comp.backward(table, params, grid, brick, views, w, h, rgb, ns, g), ns

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    ev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a.record(stream)
        rgb, trans, ns = comp.forward(table, params, grid, brick, views, w, h)
        b.record(stream)
        g = 2.0 * (rgb - target)
        comp.backward(table, params, grid, brick, views, w, h, rgb, ns, g)
        c.record(stream)
        ev.append((a, b, c))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    fwd_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    bwd_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))
    records, unique = comp.last_stats()
    rays = nv * w * h
    nvox = params.shape[0]
    # algorithmic bytes: forward 16 B per sample + 20 B per ray out; backward
    # 16 B per sample recomputed, 36 B per ray in (dL/dC, C, offset, count), a
    # 20 B record per sample written (key, float4), 2 radix passes over the 15
    # brick-slot bits reading and writing the 8 B (key, id) pairs, the per-slot
    # reduction (24 B in per record: key, id, float4) writing every voxel of
    # every allocated brick (16 B per voxel)
    fwd_bytes = 16 * records + 20 * rays
    bwd_bytes = 16 * records + 36 * rays + 20 * records + 2 * 16 * records + 24 * records + 16 * nvox
    res = {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "workload": "composite_512",
           "rays_per_step": rays, "ms_per_step": round(dt * 1e3, 3), "fwd_ms": round(fwd_ms, 3),
           "bwd_ms": round(bwd_ms, 3), "records": records, "unique_voxels": unique,
           "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                        "fwd_achieved": round(fwd_bytes / (fwd_ms / 1e3) / 1e9, 1),
                        "bwd_achieved": round(bwd_bytes / (bwd_ms / 1e3) / 1e9, 1)},
           "cpu_baseline": None}
    if with_cpu:
        # the FP64 restatement's forward on one of the 16 views
        co = entry.load_oracle().CompositeOracle()
        threads = min(16, os.cpu_count() or 1)
        v1 = comp_mod.orbit_views(nv, 60.0, w, h, seed=0)[:1]
        t0 = time.perf_counter()
        co.forward(vol.table, vol.params, grid, brick, v1, w, h, threads=threads)
        cdt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(w * h / cdt / 1e6, 3), "unit": "Mrays/s (forward only, FP64)",
                               "cores": threads, "kind": "port", "sample": f"view 0 of 16, {w}x{h} rays"}
    return res


def plain_bench(yx, torch, with_cpu, reps=8):
    """plain_512: depth-8 octree (64 spheres), 8 orbit views of 512x512
    primary rays, light at the eye; value = rays / mean kernel time of a view
    (HIP events), CPU baseline = the restatement on all 8 views at 1, 4 and up
    to 16 threads (SimpleRenderer / TreadedRenderer's 4 threads / all)."""
    wl = yx.workloads
    scene = yx.SVOData.build_synthetic(**wl.PLAIN_512)
    cams = wl.orbit_cameras(8, seed=0)
    r = yx.SVORenderer()
    r.SetScene(scene)
    r.SetFOV(60.0)
    r.SetResolution(512, 512)
    r.SetLight(wl.SUN, yx.YX_LIGHT_HEADLIGHT)
    out = r.render_tensors(("hit_id", "voxel", "t", "rgba", "shadow"))
    stream = torch.cuda.current_stream()
    for pos, d in cams:
        r.SetViewPos(pos)
        r.SetViewDir(d)
        r.Render(stream=stream, **out)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        for pos, d in cams:
            r.SetViewPos(pos)
            r.SetViewDir(d)
            r.Render(stream=stream, **out)
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / (reps * len(cams))
    res = {"value": round(512 * 512 / (ms / 1e3) / 1e6, 3), "unit": "Mrays/s", "workload": "plain_512",
           "ms_per_view": round(ms, 4), "views": len(cams), "cpu_baseline": None}
    if with_cpu:
        orc = entry.load_oracle().Oracle()
        nodes, root = scene.nodes(), scene.GetRoot()
        by = {}
        for th in (1, 4, min(16, os.cpu_count() or 1)):
            t0 = time.perf_counter()
            for pos, d in cams:
                orc.render(nodes, root, pos, d, fov=60.0, w=512, h=512, threads=th)
            by[str(th)] = round(len(cams) * 512 * 512 / (time.perf_counter() - t0) / 1e6, 4)
        top = max(by, key=lambda k: int(k))
        res["cpu_baseline"] = {"value": by[top], "unit": "Mrays/s", "cores": int(top), "kind": "port",
                               "sample": "all 8 views, 512x512", "by_threads": by}
    return res


def pipeline_bench(yx, torch, scene, cams, with_cpu, reps=2, cpu_frames=4):
    """SVORenderer::Render's GPU pipeline (demo/SVORenderer.cpp:95-149) on
    large_1080p: the 32 orbit frames at 1920x1080, fov 60, detail coef 1 and
    SSNA on (SVORenderer's defaults), one point light at the eye with the
    demo's parameters (demo/Demo.cpp:141-146). A frame is the LOD trace, five
    blur passes and the shade; value = primary rays / mean frame time (HIP
    events on the stream). The blur's share is the difference to the same
    frames with SSNA off. CPU baseline: the restatement on `cpu_frames` frames."""
    w, h = 1920, 1080
    depth = scene.info()["depth"]
    r = yx.SVORenderer()
    r.SetScene(scene)
    r.SetFOV(60.0)
    r.SetResolution(w, h)
    r.SetDetailCoef(1.0)
    out = r.pipeline_tensors(("rgba", "z"))
    extra = {"visits": torch.empty((h, w), dtype=torch.int32, device=out["rgba"].device),
             "level": torch.empty((h, w), dtype=torch.uint8, device=out["rgba"].device)}
    stream = torch.cuda.current_stream()

    def frame(k, **o):
        pos, d = cams[k % len(cams)]
        r.SetViewPos(pos)
        r.SetViewDir(d)
        r.SetLigth(0, tuple(float(v) for v in pos))
        r.RenderPipeline(stream=stream, **o)

    # untimed counting pass: visits, hits and LOD stops of every frame
    visits = hits = stops = 0
    for k in range(len(cams)):
        frame(k, **out, **extra)
        torch.cuda.synchronize()
        lv = extra["level"]
        visits += int(extra["visits"].to(torch.int64).sum().item())
        hits += int((lv > 0).sum().item())
        stops += int(((lv > 0) & (lv < depth)).sum().item())
    nf = len(cams)
    res = {"unit": "Mrays/s", "workload": "large_1080p pipeline", "resolution": [w, h], "frames": nf,
           "detail_coef": 1.0, "lod_stop_fraction_of_hits": round(stops / max(1, hits), 5)}
    for ssna in (True, False):
        r.SetSSNA(ssna)
        for k in range(2):
            frame(k, **out)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for k in range(reps * nf):
            frame(k, **out)
        b.record(stream)
        torch.cuda.synchronize()
        res["ms_per_frame" if ssna else "ms_per_frame_no_ssna"] = round(a.elapsed_time(b) / (reps * nf), 4)
    r.SetSSNA(True)
    # a coarser level of detail (more LOD stops): coef 4
    r.SetDetailCoef(4.0)
    stops4 = hits4 = 0
    for k in range(nf):
        frame(k, **out, **extra)
        torch.cuda.synchronize()
        lv = extra["level"]
        hits4 += int((lv > 0).sum().item())
        stops4 += int(((lv > 0) & (lv < depth)).sum().item())
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for k in range(reps * nf):
        frame(k, **out)
    b.record(stream)
    torch.cuda.synchronize()
    res["coef4"] = {"ms_per_frame": round(a.elapsed_time(b) / (reps * nf), 4),
                    "lod_stop_fraction_of_hits": round(stops4 / max(1, hits4), 5)}
    r.SetDetailCoef(1.0)
    ms = res["ms_per_frame"]
    res["value"] = round(w * h / (ms / 1e3) / 1e6, 3)
    blur_ms = max(1e-6, ms - res["ms_per_frame_no_ssna"])
    res["blur_ms_5_passes"] = round(blur_ms, 4)
    # per blur pass: z read + z written per pixel
    pass_bytes = 8 * w * h
    res["blur_roofline"] = {"bound": "hbm", "achieved": round(pass_bytes / (blur_ms / 5 / 1e3) / 1e9, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(pass_bytes / (blur_ms / 5 / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                            "bytes_per_pass": pass_bytes}
    # algorithmic bytes of a frame: 8 B per node record read, 4 B per hit
    # colour, trace z + colour writes 8 B, five passes 8 B, shade reads z and
    # colour and writes RGBA8 and z 16 B per pixel
    fb = (8 * visits + 4 * hits) / nf + (8 + 5 * 8 + 16) * w * h
    res["algorithmic_bytes_per_frame"] = int(fb)
    res["achieved_gbs"] = round(fb / (ms / 1e3) / 1e9, 1)
    res["cpu_baseline"] = None
    if with_cpu:
        po = entry.load_oracle().PipelineOracle()
        nodes, root = scene.nodes(), scene.GetRoot()
        lod = po.lod_colors(nodes, root)
        mk = entry.load_oracle().make_lights
        th = min(16, os.cpu_count() or 1)
        sel = list(range(0, nf, nf // cpu_frames))[:cpu_frames]
        t0 = time.perf_counter()
        for k in sel:
            pos, d = cams[k]
            lt = mk([dict(pos=pos, diffuse=(0.7,) * 3, specular=(0.3,) * 3, attenuation=(1.0, 0.0, 0.5))])
            po.pipeline(nodes, root, depth, pos, d, fov=60.0, w=w, h=h, detail_coef=1.0, lights=lt, lod=lod,
                        threads=th)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(len(sel) * w * h / dt / 1e6, 4), "unit": "Mrays/s", "cores": th,
                               "kind": "port", "sample": f"orbit frames {sel}, full 1920x1080 pipeline "
                                                         f"(trace, 5 blurs, shade), {dt:.1f} s"}
    return res


def update_bench(yx, torch, scene, reps=3, with_edit=True):
    """Scene sync on large_1080p (14.2 M VoxNodes, 46.4 M records):
    SetSceneHost (host breadth-first re-layout + LOD colours + upload of the
    records) and SetScene (upload of the VoxNodes, re-layout on the GPU)
    against UpdateScene (CudaSVO::Update, demo/SVORenderer.cpp:39-53: upload
    of the changed 256-node pages to the device node copy, re-layout and LOD
    colours on the GPU). Wall time per call, synchronised; the edit recolours
    the leaves of 64 nodes spread over 4 pages."""
    nodes = scene.nodes()
    cand = np.nonzero(nodes[:, 0] & 0xFF)[0]
    pages = sorted({int(cand[int(k * len(cand) / 4)]) // 256 for k in range(4)})
    idx = [int(i) for p in pages for i in cand[(cand // 256) == p][:16]]
    orig = nodes[idx].copy()
    edited = orig.copy()
    for row in edited:
        for sl in range(8):
            if (int(row[0]) >> sl) & 1:
                row[2 + sl] = 0xFF0000FF

    def wall(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    r_host = yx.SVORenderer()
    r_gpu = yx.SVORenderer()
    host_ms = min(wall(lambda: r_host.SetSceneHost(scene)) for _ in range(reps))
    set_ms = min(wall(lambda: r_host.SetScene(scene)) for _ in range(reps))
    full_ms = wall(lambda: r_gpu.UpdateScene(scene))
    full_pages, full_bytes = r_gpu.last_update()
    nop_ms = min(wall(lambda: r_gpu.UpdateScene(scene)) for _ in range(reps))
    edit_ms = []
    for k in range(reps):
        for i, row in zip(idx, edited if k % 2 == 0 else orig):
            scene.write_nodes(i, row)
        edit_ms.append(wall(lambda: r_gpu.UpdateScene(scene)))
        edit_pages, edit_bytes = r_gpu.last_update()
    for i, row in zip(idx, orig):
        scene.write_nodes(i, row)
    r_gpu.UpdateScene(scene)
    edit = edit_bench(yx, torch, wall, host_ms) if with_edit else None
    return {"workload": "large_1080p scene sync", "nodes": int(nodes.shape[0]),
            "set_scene_host_ms": round(host_ms, 1), "set_scene_ms": round(set_ms, 1),
            "update_full_ms": round(full_ms, 1), "update_full_pages": full_pages,
            "update_full_mb": round(full_bytes / 1e6, 1),
            "update_unchanged_ms": round(nop_ms, 2),
            "update_edit_ms": round(min(edit_ms), 2), "update_edit_pages": edit_pages,
            "update_edit_kb": round(edit_bytes / 1e3, 1),
            "speedup_edit_vs_set_scene_host": round(host_ms / min(edit_ms), 1), "sphere_edits": edit}


def edit_bench(yx, torch, wall, set_scene_ms, reps=5):
    """Demo::DoEdit on large_1080p (demo/Demo.cpp:82-114): 100 sphere edits of
    radius 4 at level 11, centred on voxels that primary rays of an orbit view
    hit, applied with BuildRange in place on the host, then UpdateScene
    (CudaSVO::Update). Alternate batches GROW and CLEAR. Against: the same
    batch applied by rebuilding the scene (on the host, or on the GPU) and
    SetSceneHost's full host re-layout. Wall time, synchronised."""
    wl = yx.workloads
    s = yx.SVOData.build_synthetic_gpu(**wl.LARGE_1080P)
    r = yx.SVORenderer()
    r.UpdateScene(s)
    pos, d = wl.orbit_cameras(1, seed=0)[0]
    r.SetViewPos(pos)
    r.SetViewDir(d)
    r.SetFOV(60.0)
    r.SetResolution(480, 270)
    out = r.render_tensors(("hit_id", "voxel"))
    torch.cuda.synchronize()
    vox = out["voxel"].reshape(-1, 3).cpu().numpy()
    hit = out["hit_id"].reshape(-1).cpu().numpy() != -1
    rng = np.random.default_rng(4)
    cand = vox[hit]
    batches = []
    for k in range(reps + 1):
        pick = cand[rng.choice(cand.shape[0], 100, replace=False)]
        mode = k % 2
        batches.append([(mode, 11, tuple(float(v) + 0.5 for v in p), 4.0, 0xFF80C0FF) for p in pick])
    edit_ms, upd_ms, pages = [], [], []
    for k in range(reps):
        edit_ms.append(wall(lambda: s.BuildRanges(batches[k])))
        upd_ms.append(wall(lambda: r.UpdateScene(s)))
        pages.append(r.last_update()[0])
    b = batches[reps]
    gpu_rebuild_ms = wall(lambda: s.BuildRanges(b, how=yx.SVOData.REBUILD_GPU))
    gpu_update_ms = wall(lambda: r.UpdateScene(s))
    gpu_pages = r.last_update()[0]
    cs = yx.SVOData.build_synthetic(**wl.LARGE_1080P)
    cpu_rebuild_ms = wall(lambda: cs.BuildRanges(b, how=yx.SVOData.REBUILD))
    # an IsoSource of 128^3 values (scene_gen.py:111-130's bonsai is 256^3) grown at level 11
    blob = wl.iso_blob((128, 128, 128), 7)
    iso_place = wall(lambda: s.BuildRangeIso(11, (900, 900, 1100), 0, blob, 90, color=0xFF30A050))
    iso_upd = wall(lambda: r.UpdateScene(s))
    iso_pages = r.last_update()[0]
    iso_gpu = wall(lambda: s.BuildRangeIso(11, (300, 1400, 1150), 0, blob, 90, color=0xFF30A050,
                                           how=yx.SVOData.REBUILD_GPU))
    iso_cpu = wall(lambda: cs.BuildRangeIso(11, (900, 900, 1100), 0, blob, 90, color=0xFF30A050,
                                            how=yx.SVOData.REBUILD))
    k = int(np.argmin([a + u for a, u in zip(edit_ms, upd_ms)]))
    total = edit_ms[k] + upd_ms[k]
    return {"workload": "large_1080p, 100 sphere edits (radius 4, level 11) per batch",
            "in_place_edit_ms": round(edit_ms[k], 2), "update_ms": round(upd_ms[k], 2),
            "update_pages": pages[k], "in_place_total_ms": round(total, 2),
            "gpu_rebuild_ms": round(gpu_rebuild_ms, 1), "gpu_rebuild_update_ms": round(gpu_update_ms, 1),
            "gpu_rebuild_update_pages": gpu_pages,
            "cpu_rebuild_ms": round(cpu_rebuild_ms, 1),
            "cpu_rebuild_set_scene_ms": round(cpu_rebuild_ms + set_scene_ms, 1),
            "speedup_vs_cpu_rebuild_set_scene": round((cpu_rebuild_ms + set_scene_ms) / total, 1),
            "iso_source_128": {"in_place_ms": round(iso_place, 1), "update_ms": round(iso_upd, 2),
                               "update_pages": iso_pages, "gpu_rebuild_ms": round(iso_gpu, 1),
                               "cpu_rebuild_ms": round(iso_cpu, 1)},
            "nodes": s.info()["num_nodes"]}


def voxtree_bench(yx, torch, with_cpu=True, reps=20):
    """Brick classification of the reference's 1024^3 bit volume
    (reaction/voxtree/voxtree.py:80-95): CalcDensity, MarkBricks and PackBricks
    each timed with HIP events over `reps` launches on the current stream, and
    the whole classify call (with the column scan and the total read back) by
    wall clock. Algorithmic bytes: density reads B^3/8 and writes V^3/2 bytes;
    marking reads the V^3/2 density bytes once and writes 4 G^3 + 4 G^2; packing
    reads 4 G^3 + 4 G^2 and writes 4 per mixed brick."""
    cfg = yx.workloads.VOXTREE_1024
    B = cfg["bit_size"]
    V, G = B // 2, B // 8
    vol = yx.workloads.voxtree_volume(B, cfg["seed"])
    bits = yx.voxtree.pack_bits(vol)
    del vol
    stream = torch.cuda.current_stream()
    r = yx.voxtree.classify(bits, B)
    start = r["column_start"][:-1].contiguous()
    total = r["total"]

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    d_out = torch.empty_like(r["density"])
    m_out = (torch.empty_like(r["brick_state"]), torch.empty_like(r["column_count"]))
    p_out = torch.empty(max(1, total), dtype=torch.int32, device=bits.device)
    d_ms = timed(lambda: yx.voxtree.calc_density(bits, B, out=d_out))
    m_ms = timed(lambda: yx.voxtree.mark_bricks(r["density"], V, out=m_out))
    p_ms = timed(lambda: yx.voxtree.pack_bricks(r["brick_state"], start, V, total, out=p_out))
    walls = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        yx.voxtree.classify(bits, B)
        torch.cuda.synchronize()
        walls.append((time.perf_counter() - t0) * 1e3)
    d_bytes = B ** 3 // 8 + V ** 3 // 2
    m_bytes = V ** 3 // 2 + 4 * G ** 3 + 4 * G * G
    p_bytes = 4 * G ** 3 + 4 * G * G + 4 * total

    def rf(nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"ms": round(ms, 4), "bytes": nbytes, "achieved_gbs": round(gbs, 1),
                "frac": round(gbs / HBM_PEAK_GBS, 4)}

    out = {"workload": f"voxtree_1024: {B}^3 bit volume, {G}^3 bricks of 4^3 (5^3 with overlap)",
           "mixed_bricks": total, "calc_density": rf(d_bytes, d_ms), "mark_bricks": rf(m_bytes, m_ms),
           "pack_bricks": rf(p_bytes, p_ms), "classify_wall_ms": round(min(walls), 3),
           "bricks_per_s": round(G ** 3 / (min(walls) * 1e-3), 1)}
    if with_cpu:
        orc = entry.load_oracle().VoxtreeOracle()
        words = bits.cpu().numpy().view(np.uint32)
        t0 = time.perf_counter()
        exp = orc.classify(words, B)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(G ** 3 / cpu_s, 1), "unit": "bricks/s", "cores": 1, "kind": "port",
                               "sample": f"the whole {B}^3 volume through the numpy restatement, {cpu_s:.1f} s"}
        out["cpu_total_matches"] = exp["total"] == total
    return out


def pool_bench(yx, torch, dev, reps=5):
    """Brick pool churn on the composite_512 volume (26 k of 32 768 bricks
    allocated, 8^3 float4 per brick): per round, release 10 % of the bricks,
    allocate as many others and upload their packed contents
    (CuBrickPoolManager::allocMap + UpdateBrickPool), against repacking the
    whole volume on the host and uploading it. Wall time, synchronised."""
    comp_mod = yx.composite
    cfg = comp_mod.COMPOSITE_512
    grid, brick = cfg["grid"], cfg["brick"]
    vox = brick ** 3
    vol = comp_mod.Volume(**cfg)
    table, params = vol.to(dev)
    spare = vol.num_slots // 5
    params = torch.cat([params, torch.zeros((spare * vox, 4), dtype=torch.float32, device=dev)])
    pool = comp_mod.BrickPool(table, params, grid, brick)
    rng = np.random.default_rng(2)
    live = np.nonzero(vol.table >= 0)[0].astype(np.int32)
    empty = np.nonzero(vol.table < 0)[0].astype(np.int32)
    k = live.size // 10
    packed = torch.from_numpy(rng.standard_normal((k * vox, 4)).astype(np.float32)).to(dev)
    times = []
    for r in range(reps):
        drop = torch.from_numpy(rng.choice(live, k, replace=False)).to(dev)
        add = torch.from_numpy(rng.choice(empty, k, replace=False)).to(dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pool.release(drop)
        pool.alloc(add)
        pool.update(add, packed)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        pool.release(add)  # back to the original occupancy (timing excluded)
        pool.alloc(drop)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    v2 = comp_mod.Volume(**cfg)
    t2, p2 = v2.to(dev)
    torch.cuda.synchronize()
    rebuild_ms = (time.perf_counter() - t0) * 1e3
    return {"workload": "composite_512 brick pool churn", "bricks": int(live.size), "churn_bricks": int(k),
            "churn_ms": round(float(np.min(times)), 3), "churn_upload_mb": round(k * vox * 16 / 1e6, 1),
            "host_rebuild_ms": round(rebuild_ms, 1), "speedup_vs_rebuild": round(rebuild_ms / np.min(times), 1)}


def batch_bench(yx, torch, frames, tune_cache):
    """batched_48: 48 octrees of depth 6-11 (sphere and sphere + heightfield
    recipes), 416 cameras at 256x256 .. 1920x1080 (348.5 M primary rays, plus
    shadow rays on the heightfield scenes). One frame of every job in one
    persistent launch replayed from a hipGraph, with the launch shape from the
    library's autotuner, against per-scene plain launches (one SVORenderer
    per scene, one Render per camera, default launch shape) writing the same
    output buffers. Both are timed with HIP events on the same stream over
    `frames` frames after one untimed frame."""
    wl = yx.workloads
    recipes = wl.batched_recipes()
    jobs = wl.batch_jobs(recipes)
    t0 = time.time()
    scenes = [yx.SVOData.build_synthetic(**d) for d, _ in recipes]
    print(f"[bench] batched_48: {len(scenes)} scenes built in {time.time() - t0:.1f} s", file=sys.stderr,
          flush=True)
    b = yx.BatchRenderer()
    b.set_scenes(scenes)
    total = b.set_jobs(jobs)
    keys = ("hit_id", "voxel", "t", "rgba", "shadow")
    out = b.alloc_outputs(keys)
    t0 = time.time()
    best, trials = b.autotune(out, reps=1, cache=tune_cache)
    tune_s = time.time() - t0
    print(f"[bench] batch autotune: {len(trials)} trials in {tune_s:.1f} s, best {best}", file=sys.stderr,
          flush=True)
    stream = torch.cuda.current_stream()
    b.render(stream=stream, **out)
    torch.cuda.synchronize()
    rays = total + int((out["shadow"] < 2).sum().item())

    def timed(frame):
        frame()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(frames):
            frame()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / frames

    batch_ms = timed(lambda: b.render(stream=stream, **out))
    rs = []
    for s in scenes:
        r = yx.SVORenderer()
        r.SetScene(s)
        rs.append(r)
    views = []
    for (off, w, h), j in zip(b.job_slices(), jobs):
        n = w * h
        views.append((j, {k: out[k][off:off + n] for k in keys}))

    def per_scene():
        for j, o in views:
            r = rs[j["scene"]]
            r.SetViewPos(j["pos"])
            r.SetViewDir(j["dir"])
            r.SetFOV(j["fov"])
            r.SetResolution(j["width"], j["height"])
            r.SetLight(j["light"], j["light_mode"])
            r.Render(stream=stream, **o)

    plain_ms = timed(per_scene)
    # the same per-scene launches with the batch's tuned launch shape, to
    # separate what the shape and what the single launch contribute
    for r in rs:
        r.set_tuning(**best)
    plain_tuned_ms = timed(per_scene)
    del rs
    return {"value": round(rays / (batch_ms / 1e3) / 1e6, 3), "unit": "Mrays/s", "workload": "batched_48",
            "scenes": len(scenes), "jobs": len(jobs), "primary_rays": total, "rays_per_frame": rays,
            "ms_per_frame": round(batch_ms, 3), "per_scene_ms_per_frame": round(plain_ms, 3),
            "per_scene_value": round(rays / (plain_ms / 1e3) / 1e6, 3),
            "speedup_vs_per_scene": round(plain_ms / batch_ms, 3),
            "per_scene_tuned_ms_per_frame": round(plain_tuned_ms, 3),
            "speedup_vs_per_scene_tuned": round(plain_tuned_ms / batch_ms, 3), "frames": frames,
            "graph": True, "tuning": best, "autotune_trials": len(trials), "autotune_s": round(tune_s, 1),
            "scene_bytes": b.scene_bytes()}


def brick_bench(yx, torch, frames, with_cpu):
    """brick_4096: 4096^3 voxels as an 8^3 root grid over three levels of 4^3
    nodes and 8^3 u8 bricks (fBm density, ~3 % occupancy), built on the GPU;
    1920x1080 primary rays per frame from `frames` orbit cameras, cast by the
    n-tree DDA with front-to-back compositing. value = rays / mean kernel time
    (HIP events on the launch stream). Algorithmic bytes per frame: 12 B per
    node record read + 1 B per brick voxel read + 20 B of outputs per pixel
    (RGBA8, t, voxel), the read counts taken from the kernel's own counters in
    an untimed pass. 12 B is the least a 4^3 node can hold - a 64-bit child
    mask and a 32-bit first-child index - independent of the 16-byte padded
    record the builder stores."""
    wl = yx.workloads
    cfg = wl.BRICK_4096
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    t0 = time.time()
    tree = yx.BrickTree.build_fbm(stream=stream, **cfg)
    torch.cuda.synchronize()
    build_s = time.time() - t0
    info = tree.info()
    print(f"[bench] brick_4096 built in {build_s:.2f} s: {info}", file=sys.stderr, flush=True)
    w, h = 1920, 1080
    pal_np = wl.brick_palette()
    pal = torch.from_numpy(pal_np.view(np.int32)).cuda()
    views = yx.composite.camera_views(wl.orbit_cameras(frames, seed=2), 60.0, w, h)
    out = tree.render_tensors(views[0], w, h, pal, outputs=("rgba", "t", "voxel", "nodes", "voxels"))
    algo = []
    reads = []
    for v in views:
        tree.render(v, w, h, pal, stream=stream, **out)
        torch.cuda.synchronize()
        nn = int(out["nodes"].to(torch.int64).sum().item())
        nv = int(out["voxels"].to(torch.int64).sum().item())
        algo.append(NTREE_NODE_BYTES * nn + nv + 20 * w * h)
        reads.append((nn, nv))
    timed = {k: out[k] for k in ("rgba", "t", "voxel")}
    tree.render(views[0], w, h, pal, stream=stream, **timed)
    ev = []
    for v in views:
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        tree.render(v, w, h, pal, stream=stream, **timed)
        b.record(stream)
        ev.append((a, b))
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b in ev]
    kern_ms = float(np.mean(ms))
    achieved = float(np.mean(algo)) / (kern_ms / 1e3) / 1e9
    res = {"value": round(w * h / (kern_ms / 1e3) / 1e6, 3), "unit": "Mrays/s", "workload": "brick_4096",
           "resolution": [w, h], "frames": frames, "ms_per_frame": round(kern_ms, 4),
           "build_s": round(build_s, 3), "bricks": info["num_bricks"], "nodes": info["num_nodes"],
           "tree_bytes": info["bytes"],
           "mean_node_reads_per_ray": round(float(np.mean([r[0] for r in reads])) / (w * h), 3),
           "mean_voxel_reads_per_ray": round(float(np.mean([r[1] for r in reads])) / (w * h), 3),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5), "kernel": "nt_trace_kernel<3, false>",
                        "kernel_ms": round(kern_ms, 4), "node_bytes": NTREE_NODE_BYTES,
                        "algorithmic_bytes_per_launch": int(np.mean(algo))},
           "cpu_baseline": None}
    if with_cpu:
        oracle = entry.load_oracle()
        o = oracle.NTreeOracle()
        root, nodes, bricks = tree.download()
        threads = min(16, os.cpu_count() or 1)
        ystep = 8
        t0 = time.perf_counter()
        o.render(root, nodes, bricks, cfg["size"], cfg["levels"], views[0], pal_np, w, h, ystep=ystep,
                 threads=threads)
        dt = time.perf_counter() - t0
        rays = w * len(range(0, h, ystep))
        res["cpu_baseline"] = {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                               "kind": "port", "sample": f"every {ystep}th row of orbit frame 0 ({rays} rays)"}
    return res


def measured_traffic(workload):
    """HBM bytes per launch of the frame kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-kernel calibration), or
    None when no measurement for this workload is recorded."""
    import pathlib
    f = pathlib.Path(__file__).resolve().parent / "profiles" / f"traffic_{workload}.json"
    if not f.exists():
        return None
    rec = json.loads(f.read_text())
    return rec.get("traffic_bytes_per_launch")


def cpu_baseline(scene, cams, w, h, mode, sun, nframes):
    """The oracle (line-faithful CPU restatement) on a bounded sample: `nframes`
    of the orbit frames, row bands over the host threads (TreadedRenderer's
    split, cell/ppu_renderer.cpp:121-144)."""
    orc = entry.load_oracle().Oracle()
    nodes, root = scene.nodes(), scene.GetRoot()
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    idx = np.linspace(0, len(cams) - 1, nframes).astype(int)
    rays = 0
    t0 = time.perf_counter()
    for i in idx:
        pos, d = cams[i]
        o = orc.render(nodes, root, pos, d, fov=60.0, w=w, h=h, light_mode=mode, light=sun, threads=threads)
        rays += o["rays"]
    dt = time.perf_counter() - t0
    # SimpleRenderer (1 thread) and TreadedRenderer's hard-coded 4 threads
    # (cell/ppu_renderer.cpp:129) on one frame each
    by_threads = {}
    for th in (1, 4):
        pos, d = cams[int(idx[0])]
        t1 = time.perf_counter()
        o = orc.render(nodes, root, pos, d, fov=60.0, w=w, h=h, light_mode=mode, light=sun, threads=th)
        by_threads[str(th)] = round(o["rays"] / (time.perf_counter() - t1) / 1e6, 4)
    by_threads[str(threads)] = round(rays / dt / 1e6, 4)
    return {"value": round(rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{len(idx)} of the {len(cams)} orbit frames (frames {idx.tolist()}), full "
                      f"{w}x{h}, {rays} rays, {dt:.1f} s wall on {threads} threads; 1 and 4 threads on "
                      f"frame {int(idx[0])}",
            "by_threads": by_threads, "cpu": _cpu_model()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
