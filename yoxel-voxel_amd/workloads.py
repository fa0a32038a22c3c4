"""Synthetic workloads: scene recipes and orbit cameras.

plain_512: depth-8 octree (256^3) from 64 spheres, radii 8-32 voxels, 512x512
           primary rays, 60 deg FOV, 8 seeded orbit angles at radius 1.5.
large_1080p: depth-11 octree (2048^3) from 512 spheres, radii 16-128 voxels,
           united with a 5-octave fBm-Perlin heightfield of amplitude 0.2;
           1920x1080 primary rays plus one shadow ray per lit hit toward a fixed
           directional light; 32 orbit frames.
"""
import math

import numpy as np

CENTER = (0.5, 0.5, 0.5)
ORBIT_RADIUS = 1.5
ELEVATION_DEG = 30.0
UP = (0.0, 0.0, 1.0)

PLAIN_512 = dict(depth=8, num_spheres=64, radius_min=8.0, radius_max=32.0, heightfield=False,
                 hf_octaves=5, hf_amplitude=0.2, hf_base=0.3, seed=1)
LARGE_1080P = dict(depth=11, num_spheres=512, radius_min=16.0, radius_max=128.0, heightfield=True,
                   hf_octaves=5, hf_amplitude=0.2, hf_base=0.3, seed=11)

# fixed directional light of the large scene: unit vector toward the light
SUN = tuple(np.float32(v) for v in (np.array([0.4, 0.3, 0.85]) / np.linalg.norm([0.4, 0.3, 0.85])))


def orbit_cameras(n, seed=0, radius=ORBIT_RADIUS, elevation_deg=ELEVATION_DEG):
    """n cameras on a circle around the cube centre looking at it: azimuth
    2*pi*k/n plus a seeded offset in [0, 2*pi/n). Returns [(pos, dir)] as
    float32 triples."""
    rng = np.random.default_rng(seed)
    off = rng.uniform(0.0, 2.0 * math.pi / n)
    el = math.radians(elevation_deg)
    cams = []
    for k in range(n):
        az = 2.0 * math.pi * k / n + off
        p = np.array([CENTER[0] + radius * math.cos(el) * math.cos(az),
                      CENTER[1] + radius * math.cos(el) * math.sin(az),
                      CENTER[2] + radius * math.sin(el)])
        d = np.array(CENTER) - p
        cams.append((tuple(np.float32(p)), tuple(np.float32(d))))
    return cams


# batched_48: 48 octrees of depth 6..11, sphere-only (plain_512-like, headlight)
# and sphere + heightfield (large_1080p-like, sun with shadow rays) recipes,
# seeded per scene, each with 1..16 orbit cameras at resolutions drawn from
# RESOLUTIONS.
RESOLUTIONS = ((256, 256), (640, 480), (1280, 720), (1920, 1080))


def batched_recipes(n_scenes=48, seed=48):
    """Scene i: depth 6 + i % 6; recipe plain-like for (i // 6) even, else
    large-like; sphere radii scale with the leaf grid (2^(depth-8) for the plain
    recipe, 2^(depth-11) for the large one, at least one voxel); seed 1000 + i.
    Returns [(scene desc dict, [camera dicts])], cameras as
    {width, height, pos, dir, light_mode, light}."""
    out = []
    for i in range(n_scenes):
        depth = 6 + i % 6
        large = (i // 6) % 2 == 1
        base = LARGE_1080P if large else PLAIN_512
        s = 2.0 ** (depth - (11 if large else 8))
        desc = dict(base)
        desc.update(depth=depth, seed=1000 + i, radius_min=max(1.0, base["radius_min"] * s),
                    radius_max=max(1.5, base["radius_max"] * s))
        rng = np.random.default_rng(seed * 1000 + i)
        n_cams = int(rng.integers(1, 17))
        cams = []
        for (pos, d) in orbit_cameras(n_cams, seed=seed * 1000 + i):
            w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
            cams.append(dict(width=w, height=h, pos=pos, dir=d,
                             light_mode=2 if large else 0, light=SUN if large else (0.0, 0.0, 1.0)))
        out.append((desc, cams))
    return out


def batch_jobs(recipes, fov=60.0):
    """Flatten recipes into BatchRenderer jobs (scene index = recipe index)."""
    jobs = []
    for si, (_, cams) in enumerate(recipes):
        for c in cams:
            jobs.append(dict(scene=si, width=c["width"], height=c["height"], pos=c["pos"], dir=c["dir"],
                             up=UP, fov=fov, light_mode=c["light_mode"], light=c["light"]))
    return jobs


# brick_4096: 4096^3 voxels as an 8^3 root grid over three levels of 4^3 nodes
# and 8^3 u8 bricks; density from a 5-octave fBm thresholded to ~3 % voxel
# occupancy (~3.8 % of the bricks allocated); 1920x1080 primary rays.
BRICK_4096 = dict(size=4096, levels=3, octaves=5, frequency=8.0, threshold=0.58, scale=400.0, seed=77)


def brick_palette():
    """256-entry RGBA8 transfer table (entry 0 unused): a blue-to-orange ramp
    with opacity rising from ~0.1 to 1 over the first ~90 density steps."""
    d = np.arange(256, dtype=np.float64) / 255.0
    r = np.clip(255.0 * np.sqrt(d), 0, 255)
    g = np.clip(255.0 * (1.0 - np.abs(2.0 * d - 0.6)), 0, 255)
    b = np.clip(255.0 * (1.0 - d) ** 2, 0, 255)
    a = np.clip(25.0 + 2.5 * np.arange(256), 0, 255)
    p = (r.astype(np.uint32) | (g.astype(np.uint32) << 8) | (b.astype(np.uint32) << 16)
         | (a.astype(np.uint32) << 24))
    p[0] = 0
    return p.astype(np.uint32)


# brick classification workload: the reference's 1024^3 bit volume (voxtree.py:75)
VOXTREE_1024 = dict(bit_size=1024, seed=1024)


def voxtree_volume(B, seed, dev="cuda", flip=2e-5):
    """Synthetic bit volume (bool [x, y, z] of side B on `dev`): 16 spheres of
    radius B/32..B/32+B/6 united with a wavy floor z < B (0.3 + 0.1 sin cos),
    then `flip` of the voxels flipped (scattered voxels and holes)."""
    import torch
    g = torch.Generator(device=dev).manual_seed(seed)
    c = torch.arange(B, device=dev, dtype=torch.float32) + 0.5
    X, Y, Z = c[:, None, None], c[None, :, None], c[None, None, :]
    vol = Z < B * (0.3 + 0.1 * torch.sin(X / B * 6.0) * torch.cos(Y / B * 5.0))
    for _ in range(16):
        p = torch.rand(3, generator=g, device=dev) * B
        r = float(torch.rand(1, generator=g, device=dev)) * B / 6 + B / 32
        vol |= (X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2 <= r * r
    vol ^= torch.rand((B, B, B), generator=g, device=dev) < flip
    return vol


def iso_blob(shape, seed):
    """uint8 values [z][y][x] for an IsoSource: a sum of 6 Gaussians of radius
    shape/16..shape/4 at uniform centres, scaled to 0..255."""
    import numpy as np
    rng = np.random.default_rng(seed)
    axes = [np.arange(k, dtype=np.float32) for k in shape]
    v = np.zeros(shape, np.float32)
    for _ in range(6):
        c = rng.uniform(0, 1, 3) * np.array(shape)
        r = rng.uniform(max(shape) / 16, max(shape) / 4)
        g = [np.exp(-((axes[a] - c[a]) ** 2) / (2 * r * r)).astype(np.float32) for a in range(3)]
        v += g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    return np.clip(v / v.max() * 255.0, 0, 255).astype(np.uint8)
