"""Bit-exact parity of the HIP path with the oracle (GPU).

Every stage is compared on the same inputs: the ray buffer (all texels,
including the ones a line must leave untouched), the per-line cleared range,
the composed framebuffer, the gap-filled framebuffer, and the whole-frame
entry point.  Integer outputs must match exactly.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import rlevox  # noqa: E402
import workloads  # noqa: E402
from oracle import UNTOUCHED  # noqa: E402


def gpu_stages(ctx, scene, rm):
    lines = max(rm.map_line_count, 1)
    d_rb = ctx.alloc(lines * rm.rs * 4)
    d_clip = ctx.alloc(lines * 8)
    d_rb.fill_u32(UNTOUCHED)
    d_clip.fill_u32(0xFFFFFFFF)
    ctx.render_raybuf(scene, rm, d_rb, d_clip)
    d_col = ctx.alloc(rm.W * rm.H * 2)
    d_dep = ctx.alloc(rm.W * rm.H * 4)
    d_col2 = ctx.alloc(rm.W * rm.H * 2)
    ctx.compose(rm, d_rb, d_clip, d_col, d_dep)
    ctx.gapfill(rm.W, rm.H, d_col, d_dep, d_col2)
    ctx.sync()
    out = {
        "raybuf": d_rb.download(np.uint32, lines * rm.rs).reshape(lines, rm.rs),
        "clip": d_clip.download(np.int32, lines * 2).reshape(lines, 2),
        "compose": d_col.download(np.uint16, rm.W * rm.H).reshape(rm.H, rm.W),
        "depth": d_dep.download(np.int32, rm.W * rm.H).reshape(rm.H, rm.W),
        "colour": d_col2.download(np.uint16, rm.W * rm.H).reshape(rm.H, rm.W),
    }
    for b in (d_rb, d_clip, d_col, d_dep, d_col2):
        b.free()
    return out


def assert_same(name, got, want):
    if not np.array_equal(got, want):
        diff = np.argwhere(got != want)
        i = tuple(diff[0])
        raise AssertionError("%s: %d of %d differ; first at %s: gpu=%r oracle=%r"
                             % (name, len(diff), got.size, i, got[i], want[i]))


# Raycast kernel variants (lines per wavefront, waves per block, LDS run-cache
# words): one lane per line in blocks of 64 or 16 lanes, one wavefront per
# line, two wavefronts per line, and the run-cache depths the tuning offers.
# All must agree bit-exactly.
KERNELS = [(64, 1, 0), (16, 1, 0), (1, 1, 40), (1, 2, 40), (1, 1, 24), (1, 2, 56)]


def check_view(ctx, oracle, dscene, mips, pose, W, H, mip_distance=0, far_distance=0):
    rm = rlevox.raymap(pose, W, H, mip_distance, far_distance)
    orm = oracle.raymap(pose, W, H, mip_distance, far_distance)
    assert rm.to_bytes() == orm.to_bytes(), "ray map differs from the oracle's"
    ref_rb, ref_clip, _ = oracle.raycast(mips, orm)
    ref_col, ref_dep = oracle.compose(orm, ref_rb, ref_clip)
    ref_fin = oracle.gapfill(ref_col, ref_dep)
    n = max(rm.map_line_count, 1)
    for lpw, wpb, rc in KERNELS:
        ctx.set_tuning(lines_per_wave=lpw, waves_per_block=wpb, run_cache=rc)
        got = gpu_stages(ctx, dscene, rm)
        tag = " [lines/wave %d, waves/block %d, run cache %d]" % (lpw, wpb, rc)
        assert_same("line clip" + tag, got["clip"][:n], ref_clip[:n])
        assert_same("ray buffer" + tag, got["raybuf"], ref_rb)
        assert_same("compose colour" + tag, got["compose"], ref_col)
        assert_same("depth" + tag, got["depth"], ref_dep)
        assert_same("gap-filled colour" + tag, got["colour"], ref_fin)
    ctx.set_tuning()
    return rm


@pytest.fixture(scope="module")
def c1_dev(ctx, c1_scene):
    s = ctx.upload(c1_scene.mips)
    yield s
    s.free()


@pytest.fixture(scope="module")
def tiny_dev(ctx, tiny_scene):
    s = ctx.upload(tiny_scene.mips)
    yield s
    s.free()


def test_c1_flythrough_512(ctx, oracle, c1_scene, c1_dev):
    poses = workloads.flythrough(c1_scene.mips[0], frames=6, seed=3)
    for p in poses:
        check_view(ctx, oracle, c1_dev, c1_scene.mips, p, 512, 512)


@pytest.mark.parametrize("W,H", [(640, 480), (1280, 720), (1920, 1080), (320, 240), (2560, 1440)])
def test_c1_resolutions(ctx, oracle, c1_scene, c1_dev, W, H):
    for p in workloads.random_views(c1_scene.mips[0], 2, seed=W + H):
        check_view(ctx, oracle, c1_dev, c1_scene.mips, p, W, H)


def test_c1_pitch_and_altitude_edges(ctx, oracle, c1_scene, c1_dev):
    m0 = c1_scene.mips[0]
    cases = [
        (0.01, 0.3, 40.0), (-0.01, 1.0, 40.0), (-0.4, 2.0, 30.0),      # level and looking up
        (1.2, 0.7, 80.0), (-1.2, 4.0, 20.0), (2.5, 5.5, 60.0),          # steep, over the top
        (0.3, 0.0, 900.0), (0.3, math.pi / 2, 3000.0),                  # high: y mip switch
        (0.2, 3.0, -30.0),                                               # inside the terrain
    ]
    for pitch, yaw, alt in cases:
        x, z = 700.3, 800.6
        p = workloads.pose(x, alt - workloads.surface_row(m0, x, z), z, pitch, yaw)
        check_view(ctx, oracle, c1_dev, c1_scene.mips, p, 512, 384)


def test_c1_views_among_the_hills(ctx, oracle, c1_scene, c1_dev):
    """Eye 2-20 voxels above the ground, below the highest runs: most lines
    end their walk early (no later step can draw); every kernel still gives
    the oracle's full walk bit for bit."""
    for k, p in enumerate(workloads.random_views(c1_scene.mips[0], 4, seed=12, alt=(2.0, 20.0))):
        W, H = ((640, 480), (1280, 720))[k % 2]
        check_view(ctx, oracle, c1_dev, c1_scene.mips, p, W, H)


def test_c1_mip_and_far_distance(ctx, oracle, c1_scene, c1_dev):
    p = workloads.random_views(c1_scene.mips[0], 1, seed=99)[0]
    for mipd, far in [(64, 5000), (300, 80000), (4096, 20000), (4, 100)]:
        check_view(ctx, oracle, c1_dev, c1_scene.mips, p, 640, 480, mipd, far)


def test_c1_random_pose_sweep(ctx, oracle, c1_scene, c1_dev):
    """96 seeded random cameras: anywhere over the map (edges and wrap
    included), pitch -0.6..1.5 (mostly looking down, the method's lower
    limit |pitch| = 0.01 approached), any heading, eye from inside the
    terrain to 1500 voxels above it, screens of several aspects, and the
    mip and far distances varied; every kernel against the oracle."""
    m0 = c1_scene.mips[0]
    rng = np.random.default_rng(2024)
    sizes = [(96, 64), (160, 120), (320, 240), (400, 400), (640, 360), (256, 64), (1024, 576)]
    dists = [(0, 0), (0, 0), (64, 5000), (512, 30000), (16, 800)]
    for k in range(96):
        x = float(rng.uniform(0, m0["sx"])) + 0.17
        z = float(rng.uniform(0, m0["sz"])) + 0.29
        alt = float(rng.uniform(-30.0, 1500.0) if k % 3 else rng.uniform(1.0, 40.0))
        pitch = float(rng.uniform(-0.6, 1.5))
        pitch = math.copysign(max(abs(pitch), 0.02), pitch)  # the method needs |pitch| >= 0.01
        yaw = float(rng.uniform(0, 2 * math.pi))
        p = workloads.pose(x, alt - workloads.surface_row(m0, x, z), z, pitch, yaw)
        W, H = sizes[k % len(sizes)]
        mipd, far = dists[k % len(dists)]
        check_view(ctx, oracle, c1_dev, c1_scene.mips, p, W, H, mipd, far)


def test_tiny_wrapping_world(ctx, oracle, tiny_scene, tiny_dev):
    for p in workloads.random_views(tiny_scene.mips[0], 4, seed=5, alt=(2.0, 40.0)):
        check_view(ctx, oracle, tiny_dev, tiny_scene.mips, p, 400, 300)


def test_render_entry_matches_stages(ctx, oracle, c1_scene, c1_dev):
    p = workloads.random_views(c1_scene.mips[0], 1, seed=1234)[0]
    W, H = 800, 600
    d_col = ctx.alloc(W * H * 2)
    d_dep = ctx.alloc(W * H * 4)
    ctx.render(c1_dev, p, W, H, d_col, d_dep)
    col = d_col.download(np.uint16, W * H).reshape(H, W)
    dep = d_dep.download(np.int32, W * H).reshape(H, W)
    ref = oracle.render(c1_scene.mips, p, W, H)
    assert_same("render colour", col, ref["colour"])
    assert_same("render depth", dep, ref["depth"])
    # compose only
    ctx.render(c1_dev, p, W, H, d_col, d_dep, params=rlevox.RenderParams(0, 0, rlevox.POST_COMPOSE, 0))
    col = d_col.download(np.uint16, W * H).reshape(H, W)
    assert_same("render compose-only", col, ref["compose"])


def test_work_counters_match_oracle(ctx, oracle, c1_scene, c1_dev):
    """The counting kernel's exact work counts equal the oracle's.  The
    flythrough's eye stays above the terrain's highest run, where no line can
    end before the far distance; the low views (eye among the hills) end
    most of their lines early."""
    ended = 0
    low = workloads.random_views(c1_scene.mips[0], 3, seed=12, alt=(2.0, 20.0))
    for p in workloads.flythrough(c1_scene.mips[0], frames=3, seed=11) + low:
        rm = rlevox.raymap(p, 1280, 720)
        lines = max(rm.map_line_count, 1)
        d_rb = ctx.alloc(lines * rm.rs * 4)
        d_clip = ctx.alloc(lines * 8)
        _, _, want = oracle.raycast(c1_scene.mips, oracle.raymap(p, 1280, 720))
        for lpw, wpb, rc in KERNELS:
            ctx.set_tuning(lines_per_wave=lpw, waves_per_block=wpb, run_cache=rc)
            got = ctx.raycast_counters(c1_dev, rm, d_rb, d_clip)
            assert got == want, (lpw, got, want)
            # the kernels' own walk (lines end once nothing more can draw):
            # the same drawing, fewer steps and loads
            walk = ctx.raycast_counters(c1_dev, rm, d_rb, d_clip, walk=True)
            for k in ("lines", "spans", "col_loads", "cleared"):
                assert walk[k] == want[k], (lpw, k)
            for k in ("dda_steps", "ptr_loads", "rle_loads", "elems"):
                assert walk[k] <= want[k], (lpw, k)
            ended += want["dda_steps"] - walk["dda_steps"]
        ctx.set_tuning()
        d_rb.free()
        d_clip.free()
    assert ended > 0  # some lines of these views end before the far distance


def test_golden_digests(ctx, c1_scene, tiny_scene, c1_dev, tiny_dev):
    """Device renders reproduce the committed golden digests (independent of
    the oracle build on this machine)."""
    import json
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_golden as G
    golden = json.load(open(os.path.join(here, "golden", "golden.json")))
    dev = {"c1": c1_dev, "tiny": tiny_dev}
    for case in golden["renders"]:
        pose = rlevox.Pose.make(case["pose"][:3], case["pose"][3:])
        rm = rlevox.raymap(pose, case["W"], case["H"], case["mip_distance"], case["far_distance"])
        got = gpu_stages(ctx, dev[case["scene"]], rm)
        n = max(rm.map_line_count, 1)
        assert G.raybuf_digest(got["raybuf"][:n], got["clip"][:n]) == case["raybuf"], case["name"]
        assert G.digest(got["compose"], got["depth"]) == case["compose"], case["name"]
        assert G.digest(got["colour"]) == case["colour"], case["name"]


@pytest.mark.parametrize("shift", [3, 4])
def test_offset_shift_scene(ctx, oracle, c1_scene, shift):
    """Columns aligned to 2^shift words (widened offsets) render identically."""
    rp = rlevox.repack(c1_scene.mips, shift)
    dev = ctx.upload(rp.mips)
    try:
        poses = workloads.flythrough(c1_scene.mips[0], frames=5, seed=3)
        for p in poses[1:3]:
            check_view(ctx, oracle, dev, c1_scene.mips, p, 512, 384)
    finally:
        dev.free()


def test_tuning_rejects_other_run_cache_depths(ctx):
    for bad in (8, 32, 48, 64):
        with pytest.raises(rlevox.RlevoxError):
            ctx.set_tuning(run_cache=bad)
    ctx.set_tuning(run_cache=56)
    assert ctx.get_tuning().run_cache == 56
    ctx.set_tuning()


@pytest.mark.parametrize("W,H", [(333, 97), (64, 16), (1000, 17), (1920, 1080)])
def test_gapfill_arbitrary_depths(ctx, oracle, W, H):
    """Gap-fill on images no render produces: depths around every decision
    boundary of the filter (radius cut-off near 1874, tap limit 10*dmin/7),
    negative and zero depths, depths above 16 bits and the sky value, on
    sizes that do not divide into whole tiles."""
    rng = np.random.default_rng(W * 7 + H)
    near = rng.integers(1, 2800, size=(H, W))
    edge = rng.integers(1860, 1890, size=(H, W))
    big = rng.integers(60000, 1 << 20, size=(H, W))
    neg = rng.integers(-5, 2, size=(H, W))
    sky = np.full((H, W), 0x7FFFFFFF)
    pick = rng.integers(0, 10, size=(H, W))
    dep = np.where(pick < 5, near, np.where(pick < 7, edge, np.where(pick < 8, big,
                   np.where(pick < 9, neg, sky)))).astype(np.int32)
    col = rng.integers(0, 1 << 16, size=(H, W)).astype(np.uint16)
    want = oracle.gapfill(col, dep)
    d_col, d_dep, d_out = ctx.alloc(W * H * 2), ctx.alloc(W * H * 4), ctx.alloc(W * H * 2)
    try:
        d_col.upload(col)
        d_dep.upload(dep)
        ctx.gapfill(W, H, d_col, d_dep, d_out)
        ctx.sync()
        got = d_out.download(np.uint16, W * H).reshape(H, W)
    finally:
        for b in (d_col, d_dep, d_out):
            b.free()
    assert_same("gap-fill %dx%d" % (W, H), got, want)


# Poses inside a 64x64x1024 map whose columns hold 60-150 elements each
# (150 cave layers on average): the oracle examines 20-70 elements per DDA
# step on these views, so the ordered replay walks past its first 64
# remaining elements of a column (the element preparation runs in groups
# of 64 lanes, and this is where the groups' running height and texture
# coordinate carry over).
LAYERED_POSES = [(-350.0, -0.2, 0.7), (-350.0, -0.6, 2.0), (-600.0, 0.1, 1.3),
                 (-800.0, -0.3, 4.0), (-500.0, 0.4, 5.5), (-900.0, 0.05, 3.1)]


@pytest.fixture(scope="module")
def layered_scene():
    return rlevox.gen_scene(64, 1024, 3, 4, 9, 0.5, 600, 1000, 150.0, 400, 0)


def test_many_elements_per_column(ctx, oracle, layered_scene):
    m0 = layered_scene.mips[0]
    counts = np.array([int(m0["slabs"][int(m0["ptrmap"][2 * c])]) for c in range(64 * 64)])
    assert (counts > 64).mean() > 0.9 and counts.max() > 128, "scene lost its layers"
    dev = ctx.upload(layered_scene.mips)
    views = []
    try:
        for y, pitch, yaw in LAYERED_POSES:
            p = workloads.pose(20.3, y, 30.7, pitch, yaw)
            check_view(ctx, oracle, dev, layered_scene.mips, p, 320, 240)
        # the persistent batch kernel on the same views
        for y, pitch, yaw in LAYERED_POSES:
            p = workloads.pose(20.3, y, 30.7, pitch, yaw)
            dc, dd = ctx.alloc(320 * 240 * 2), ctx.alloc(320 * 240 * 4)
            views.append((p, dc, dd, rlevox.make_view(dev, p, 320, 240, dc, dd)))
        ctx.render_batch([v[3] for v in views])
        for k, (p, dc, dd, _) in enumerate(views):
            want = oracle.render(layered_scene.mips, p, 320, 240)
            assert_same("batch view %d depth" % k, dd.download(np.int32, 320 * 240).reshape(240, 320),
                        want["depth"])
            assert_same("batch view %d colour" % k,
                        dc.download(np.uint16, 320 * 240).reshape(240, 320), want["colour"])
    finally:
        for _, dc, dd, _ in views:
            dc.free()
            dd.free()
        dev.free()


def test_wide_views_fit_the_workgroup_lds(ctx, oracle, tiny_scene, tiny_dev):
    """7680x4320: a lane-per-line block keeps ceil(7680/32) bitmap words per
    lane in LDS, more than one workgroup may hold at 256 lanes; the library
    shrinks such blocks, so every tuning (and the autotuner over them)
    renders the oracle's frame."""
    W, H = 7680, 4320
    pose = workloads.random_views(tiny_scene.mips[0], 1, seed=77, alt=(4.0, 30.0))[0]
    ref = oracle.render(tiny_scene.mips, pose, W, H)
    d_col, d_dep = ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)
    try:
        for lpw, wpb in [(1, 0), (64, 4), (64, 2), (32, 2), (16, 4)]:
            ctx.set_tuning(lines_per_wave=lpw, waves_per_block=wpb)
            d_col.fill_u32(0)
            ctx.render(tiny_dev, pose, W, H, d_col, d_dep)
            tag = " [lines/wave %d, waves/block %d]" % (lpw, wpb)
            assert_same("colour" + tag, d_col.download(np.uint16, W * H).reshape(H, W),
                        ref["colour"])
            assert_same("depth" + tag, d_dep.download(np.int32, W * H).reshape(H, W),
                        ref["depth"])
        ctx.set_tuning()
        best, ms = ctx.autotune([rlevox.make_view(tiny_dev, pose, W, H, d_col, d_dep)])
        assert ms > 0
        d_col.fill_u32(0)
        ctx.render(tiny_dev, pose, W, H, d_col, d_dep)
        assert_same("colour [autotuned]", d_col.download(np.uint16, W * H).reshape(H, W),
                    ref["colour"])
    finally:
        ctx.set_tuning()
        d_col.free()
        d_dep.free()
