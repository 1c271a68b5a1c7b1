"""Oracle parity on the measured workloads (GPU).

* the plain 1920x1080 flythrough over the 4096x4096x1024 terrain (scene
  seed 1, pose seed 7), on four of its hundred poses: ray buffer, line clip
  ranges, compose, depth and gap-fill against the oracle, for the library's
  default kernel choice and with one and two wavefronts per line forced;
  and the whole-frame entry point;
* the 64-view batch over four such terrains (seeds 1-4) with resolutions
  from 320x240 to 3840x2160: the persistent launch and its graph replay,
  a sample of views (every 3840x2160 one among them) against the oracle and
  every view against a plain render.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import rlevox  # noqa: E402
import workloads  # noqa: E402
from test_gpu_parity import assert_same, gpu_stages  # noqa: E402


@pytest.fixture(scope="module")
def c2_hosts():
    hs = [workloads.gen("c2", seed=s) for s in (1, 2, 3, 4)]
    yield hs
    for h in hs:
        h.close()


@pytest.fixture(scope="module")
def c2_devs(ctx, c2_hosts):
    ds = [ctx.upload(h.mips) for h in c2_hosts]
    yield ds
    for d in ds:
        d.free()


def test_c2_flythrough_1080p(ctx, oracle, c2_hosts, c2_devs):
    hs, dev = c2_hosts[0], c2_devs[0]
    poses = workloads.flythrough(hs.mips[0], frames=100, seed=7)
    W, H = 1920, 1080
    for k in (0, 33, 66, 99):
        pose = poses[k]
        print("c2 flythrough pose %d" % k, flush=True)
        rm = rlevox.raymap(pose, W, H)
        orm = oracle.raymap(pose, W, H)
        assert rm.to_bytes() == orm.to_bytes()
        ref_rb, ref_clip, _ = oracle.raycast(hs.mips, orm)
        ref_col, ref_dep = oracle.compose(orm, ref_rb, ref_clip)
        ref_fin = oracle.gapfill(ref_col, ref_dep)
        assert np.mean(ref_dep < rlevox.SKY_DEPTH) > 0.2
        n = max(rm.map_line_count, 1)
        for wpb in (0, 1, 2):  # library's choice, one and two wavefronts per line
            ctx.set_tuning(waves_per_block=wpb)
            try:
                got = gpu_stages(ctx, dev, rm)
            finally:
                ctx.set_tuning()
            tag = " [pose %d, waves per line %d]" % (k, wpb)
            assert_same("line clip" + tag, got["clip"][:n], ref_clip[:n])
            assert_same("ray buffer" + tag, got["raybuf"], ref_rb)
            assert_same("compose colour" + tag, got["compose"], ref_col)
            assert_same("depth" + tag, got["depth"], ref_dep)
            assert_same("gap-filled colour" + tag, got["colour"], ref_fin)
        d_col, d_dep = ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)
        ctx.render(dev, pose, W, H, d_col, d_dep)
        assert_same("frame colour", d_col.download(np.uint16, W * H).reshape(H, W), ref_fin)
        assert_same("frame depth", d_dep.download(np.int32, W * H).reshape(H, W), ref_dep)
        d_col.free()
        d_dep.free()


def _c4_views(ctx, c2_hosts, c2_devs):
    spec = workloads.c4_views([h.mips[0] for h in c2_hosts], n=64, seed=4)
    views, bufs = [], []
    for si, p, W, H in spec:
        dc, dd = ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)
        bufs.append((dc, dd))
        views.append(rlevox.make_view(c2_devs[si], p, W, H, dc, dd))
    return spec, views, bufs


def _plain(ctx, dev, pose, W, H):
    dc, dd = ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)
    ctx.render(dev, pose, W, H, dc, dd)
    out = (dc.download(np.uint16, W * H).reshape(H, W), dd.download(np.int32, W * H).reshape(H, W))
    dc.free()
    dd.free()
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_c4_batch_64_views(ctx, oracle, c2_hosts, c2_devs, graph):
    spec, views, bufs = _c4_views(ctx, c2_hosts, c2_devs)
    try:
        assert {(W, H) for _, _, W, H in spec} == set(workloads.C4_RESOLUTIONS)
        for rep in range(2 if graph else 1):  # a graph is captured, then replayed
            for dc, dd in bufs:
                dc.fill_u32(0)
                dd.fill_u32(0)
            ctx.render_batch(views, graph=graph)
            ctx.sync()
        # oracle: every 4K view and every fourth of the others
        sample = [k for k, (_, _, W, _) in enumerate(spec) if W == 3840 or k % 4 == 0]
        for k, (si, p, W, H) in enumerate(spec):
            col = bufs[k][0].download(np.uint16, W * H).reshape(H, W)
            dep = bufs[k][1].download(np.int32, W * H).reshape(H, W)
            tag = "view %d (%dx%d, scene %d%s)" % (k, W, H, si + 1, ", graph" if graph else "")
            if k in sample:
                print("c4 oracle " + tag, flush=True)
                ref = oracle.render(c2_hosts[si].mips, p, W, H)
                assert_same(tag + " colour", col, ref["colour"])
                assert_same(tag + " depth", dep, ref["depth"])
            else:
                rc, rd = _plain(ctx, c2_devs[si], p, W, H)
                assert_same(tag + " colour", col, rc)
                assert_same(tag + " depth", dep, rd)
    finally:
        for dc, dd in bufs:
            dc.free()
            dd.free()
