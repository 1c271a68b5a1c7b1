"""Out-of-memory paths, forced from the host (GPU).

`rlevox_debug_fail_alloc(n)` makes the n-th device allocation of the library
fail as out of memory.  Every entry point must report
RLEVOX_ERR_OUT_OF_MEMORY, leave nothing half-built, and the next call must
succeed with the same framebuffer as an undisturbed render.  No kernel is
made to fail: the failing allocation happens on the host before any launch
that would use it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import rlevox  # noqa: E402
import workloads  # noqa: E402


def arm(n):
    rlevox.lib.rlevox_debug_fail_alloc(n)


def frame(ctx, scene, pose, W, H):
    d_col, d_dep = ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)
    try:
        ctx.render(scene, pose, W, H, d_col, d_dep)
        ctx.sync()
        return (d_col.download(np.uint16, W * H).reshape(H, W),
                d_dep.download(np.int32, W * H).reshape(H, W))
    finally:
        d_col.free()
        d_dep.free()


@pytest.fixture
def fresh_ctx():
    ctx = rlevox.Context(0)  # empty scratch: the next render must allocate
    yield ctx
    arm(0)
    ctx.close()


def test_scene_upload_out_of_memory(fresh_ctx, c1_scene):
    # level 0 pointer map, level 0 slabs, level 1 pointer map
    for n in (1, 2, 3):
        arm(n)
        with pytest.raises(rlevox.RlevoxError) as e:
            fresh_ctx.upload(c1_scene.mips)
        assert e.value.name == "OUT_OF_MEMORY", n
    arm(0)
    s = fresh_ctx.upload(c1_scene.mips)
    s.free()


def test_render_scratch_out_of_memory(fresh_ctx, ctx, c1_scene):
    scene = fresh_ctx.upload(c1_scene.mips)
    ref_scene = ctx.upload(c1_scene.mips)
    try:
        pose = workloads.flythrough(c1_scene.mips[0], frames=3, seed=5)[2]
        W, H = 640, 480
        want = frame(ctx, ref_scene, pose, W, H)
        d_col, d_dep = fresh_ctx.alloc(W * H * 2), fresh_ctx.alloc(W * H * 4)
        arm(1)  # the render's ray-buffer scratch
        with pytest.raises(rlevox.RlevoxError) as e:
            fresh_ctx.render(scene, pose, W, H, d_col, d_dep)
        assert e.value.name == "OUT_OF_MEMORY"
        d_col.free()
        d_dep.free()
        got = frame(fresh_ctx, scene, pose, W, H)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    finally:
        scene.free()
        ref_scene.free()


def test_server_request_out_of_memory(fresh_ctx, ctx, c1_scene):
    scene = fresh_ctx.upload(c1_scene.mips)
    ref_scene = ctx.upload(c1_scene.mips)
    srv = rlevox.Server(fresh_ctx, max_batch_views=4, streams=1)
    try:
        poses = workloads.random_views(c1_scene.mips[0], 2, seed=21)
        arm(1)  # the lane's buffers for the first batch
        t = srv.submit(scene, poses[0], 320, 240)
        with pytest.raises(rlevox.RlevoxError) as e:
            srv.wait(t)
        assert e.value.name == "OUT_OF_MEMORY"
        t = srv.submit(scene, poses[1], 320, 240)
        col, dep, _ = srv.wait(t)
        want = frame(ctx, ref_scene, poses[1], 320, 240)
        assert np.array_equal(col, want[0]) and np.array_equal(dep, want[1])
    finally:
        srv.close()
        scene.free()
        ref_scene.free()
