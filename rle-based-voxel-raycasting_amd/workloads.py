"""Synthetic workloads: seeded scenes and camera poses.  Scene presets:

c1  512x512 columns x 256, mips 0-3, one 512x512 view (CPU-sized case).
c2  4096x4096 x 1024, mips 0-5, 100-frame flythrough at 1920x1080.
c3  16384x16384 x 1024, mips 0-7, grazing 3840x2160 views.
View mixes: 64 views over 4 c2 scenes with resolutions uniform over 4 sizes
(c4_views), and request streams over resident scenes with resolutions
weighted 0.5/0.35/0.15.

A camera at height `alt` above the terrain surface at column (x, z) has
pos[1] = alt - surface_row(x, z): rows grow downward inside a column and the
eye sits at row -pos[1].
"""
import math

import numpy as np

SCENES = {
    "c1": dict(size_xz=512, size_y=256, nmips=4, octaves=6, persistence=0.5, hmin=32, hmax=224,
               cave_layers=1.0, cave_max_layers=31),
    "c2": dict(size_xz=4096, size_y=1024, nmips=6, octaves=8, persistence=0.5, hmin=128, hmax=896,
               cave_layers=2.0, cave_max_layers=31),
    "c3": dict(size_xz=16384, size_y=1024, nmips=8, octaves=8, persistence=0.5, hmin=128,
               hmax=896, cave_layers=3.0, cave_max_layers=31),
    # small scene for fast tests
    "tiny": dict(size_xz=64, size_y=64, nmips=3, octaves=4, persistence=0.5, hmin=8, hmax=48,
                 cave_layers=1.0, cave_max_layers=8),
}


def gen(name, seed=1, threads=0):
    from rlevox import gen_scene
    p = SCENES[name]
    return gen_scene(p["size_xz"], p["size_y"], p["nmips"], p["octaves"], seed, p["persistence"],
                     p["hmin"], p["hmax"], p["cave_layers"], p["cave_max_layers"], threads)


def surface_row(mip0, x, z):
    """Row of the first stored voxel of column (x, z), or sy if empty."""
    sx, sy, sz = mip0["sx"], mip0["sy"], mip0["sz"]
    c = (int(z) % sz) * sx + (int(x) % sx)
    ofs = int(mip0["ptrmap"][2 * c]) << int(mip0.get("shift", 0))  # widened offsets
    slabs = mip0["slabs"]
    count = int(slabs[ofs])
    y = 0
    for e in range(count):
        w = int(slabs[ofs + 2 + e])
        y += w & 1023
        if w >> 10:
            return y
    return sy


def pose(x, y, z, pitch, yaw):
    from rlevox import Pose
    return Pose.make((x, y, z), (pitch, yaw, 0.0))


def flythrough(mip0, frames=100, seed=7, alt=(10.0, 200.0), pitch=(0.05, 0.45), speed=12.0):
    """Seeded flythrough: heading drifts, altitude and pitch oscillate."""
    rng = np.random.default_rng(seed)
    sx, sz = mip0["sx"], mip0["sz"]
    x = float(rng.uniform(0.25, 0.75) * sx) + 0.37
    z = float(rng.uniform(0.25, 0.75) * sz) + 0.61
    yaw0 = float(rng.uniform(0, 2 * math.pi))
    ph = rng.uniform(0, 2 * math.pi, 3)
    out = []
    for f in range(frames):
        t = f / max(frames - 1, 1)
        yaw = yaw0 + 0.6 * math.sin(2 * math.pi * t + ph[0])
        a = alt[0] + (alt[1] - alt[0]) * (0.5 + 0.5 * math.sin(4 * math.pi * t + ph[1]))
        p = pitch[0] + (pitch[1] - pitch[0]) * (0.5 + 0.5 * math.sin(6 * math.pi * t + ph[2]))
        srow = surface_row(mip0, x, z)
        out.append(pose(x % sx + sx, a - srow, z % sz + sz, p, yaw))
        # forward on the ground plane (direction_rot of Cuda_Render.h:271-272 at ray_x=0, ray_z=1)
        x += speed * math.sin(yaw)
        z -= speed * math.cos(yaw)
    return out


def grazing(mip0, n=8, seed=11, alt=(10.0, 200.0)):
    """Views with |pitch| in [0.6, 5] degrees, either sign."""
    rng = np.random.default_rng(seed)
    sx, sz = mip0["sx"], mip0["sz"]
    out = []
    for _ in range(n):
        x = float(rng.uniform(0, sx)) + 0.29
        z = float(rng.uniform(0, sz)) + 0.53
        deg = float(rng.uniform(0.6, 5.0)) * (1 if rng.random() < 0.5 else -1)
        a = float(rng.uniform(*alt))
        out.append(pose(x + sx, a - surface_row(mip0, x, z), z + sz, math.radians(deg),
                        float(rng.uniform(0, 2 * math.pi))))
    return out


def random_views(mip0, n, seed, alt=(10.0, 200.0), pitch=(0.05, 0.6)):
    rng = np.random.default_rng(seed)
    sx, sz = mip0["sx"], mip0["sz"]
    out = []
    for _ in range(n):
        x = float(rng.uniform(0, sx)) + 0.13
        z = float(rng.uniform(0, sz)) + 0.71
        out.append(pose(x + sx, float(rng.uniform(*alt)) - surface_row(mip0, x, z), z + sz,
                        float(rng.uniform(*pitch)), float(rng.uniform(0, 2 * math.pi))))
    return out


C4_RESOLUTIONS = [(320, 240), (1280, 720), (1920, 1080), (3840, 2160)]
C5_RESOLUTIONS = [((640, 480), 0.5), ((1920, 1080), 0.35), ((3840, 2160), 0.15)]


def c4_views(scene_mip0s, n=64, seed=4):
    """(scene index, pose, W, H) for the batched config: uniform resolutions."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        si = i % len(scene_mip0s)
        W, H = C4_RESOLUTIONS[int(rng.integers(0, len(C4_RESOLUTIONS)))]
        p = random_views(scene_mip0s[si], 1, int(rng.integers(1 << 30)))[0]
        out.append((si, p, W, H))
    return out


def to_ppm(path, colour, depth=None):
    """Writes a false-colour image of a colour16 framebuffer (debug aid)."""
    c = colour.astype(np.uint32)
    shade = (c & 0xFF).astype(np.float32)
    band = (c >> 8) & 3
    tint = np.array([[1.0, 0.85, 0.6], [0.6, 1.0, 0.6], [0.8, 0.8, 1.0], [1.0, 0.7, 0.7]], np.float32)
    rgb = shade[..., None] * tint[band]
    sky = colour == 0x8844
    if depth is not None:
        sky = depth == 0x7FFFFFFF
    rgb[sky] = (140, 180, 255)
    rgb = np.clip(rgb, 0, 255).astype(np.uint8)
    H, W = colour.shape
    with open(path, "wb") as f:
        f.write(b"P6 %d %d 255\n" % (W, H))
        f.write(rgb.tobytes())
