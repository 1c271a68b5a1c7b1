"""rlevox_render_noalloc (GPU): the whole-frame render that allocates
nothing.  Called through the C ABI on two frames in a row (the second
smaller) with the library's next device allocation set to fail, each frame
must equal rlevox_render's bit for bit; a view beyond the fixed scratch is
rejected with OVER_BUDGET."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import rlevox  # noqa: E402
import workloads  # noqa: E402


def test_noalloc_render_equals_render(ctx, c1_scene):
    scene = ctx.upload(c1_scene.mips)
    bufs = []
    try:
        poses = workloads.random_views(c1_scene.mips[0], 2, seed=77)
        for p, (W, H) in zip(poses, [(1920, 1080), (640, 480)]):
            print("noalloc %dx%d" % (W, H), flush=True)
            want_c, want_d, got_c, got_d = (ctx.alloc(W * H * n) for n in (2, 4, 2, 4))
            bufs += [want_c, want_d, got_c, got_d]
            got_c.fill_u32(0)
            got_d.fill_u32(0)
            ctx.render(scene, p, W, H, want_c, want_d)
            rlevox.lib.rlevox_debug_fail_alloc(1)  # the library's next allocation fails
            try:
                ctx.render_noalloc(scene, p, W, H, got_c, got_d)
                ctx.sync()
            finally:
                rlevox.lib.rlevox_debug_fail_alloc(0)
            for name, a, b, t in (("colour", got_c, want_c, np.uint16), ("depth", got_d, want_d, np.int32)):
                x, y = a.download(t, W * H), b.download(t, W * H)
                assert np.array_equal(x, y), "%dx%d %s: %d pixels differ" % (W, H, name,
                                                                         int((x != y).sum()))
            assert (want_d.download(np.int32, W * H) != rlevox.SKY_DEPTH).any()
        # a view whose ray buffer exceeds the fixed scratch: a clear error, nothing launched
        with pytest.raises(rlevox.RlevoxError) as e:
            ctx.render_noalloc(scene, poses[0], 3840, 2160, bufs[2], bufs[3])
        assert e.value.name == "OVER_BUDGET"
    finally:
        for b in bufs:
            b.free()
        scene.free()
