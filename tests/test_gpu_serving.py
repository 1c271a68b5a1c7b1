"""Served request loop and autotuner (GPU).

Each served framebuffer must equal rlevox_render's for the same request,
whether the request ran alone or packed into batches with others; admission
by memory, queue limits, cancellation and time-outs behave as documented.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import rlevox  # noqa: E402
import workloads  # noqa: E402

SIZES = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]


def plain_render(ctx, scene, pose, W, H):
    d_col = ctx.alloc(W * H * 2)
    d_dep = ctx.alloc(W * H * 4)
    ctx.render(scene, pose, W, H, d_col, d_dep)
    ctx.sync()
    col = d_col.download(np.uint16, W * H).reshape(H, W)
    dep = d_dep.download(np.int32, W * H).reshape(H, W)
    d_col.free()
    d_dep.free()
    return col, dep


@pytest.fixture(scope="module")
def scenes(ctx, c1_scene, tiny_scene):
    s = [ctx.upload(c1_scene.mips), ctx.upload(tiny_scene.mips)]
    yield s
    for x in s:
        x.free()


def requests(c1_scene, tiny_scene, n, seed):
    rng = np.random.default_rng(seed)
    c1 = workloads.random_views(c1_scene.mips[0], n, seed)
    tn = workloads.random_views(tiny_scene.mips[0], n, seed + 1, alt=(4.0, 30.0))
    out = []
    for i in range(n):
        W, H = SIZES[int(rng.integers(len(SIZES)))]
        si = int(rng.integers(2))
        out.append((si, (c1 if si == 0 else tn)[i], W, H))
    return out


def test_served_results_equal_plain_renders(ctx, scenes, c1_scene, tiny_scene, oracle):
    reqs = requests(c1_scene, tiny_scene, 40, seed=21)
    hosts = [c1_scene, tiny_scene]
    want = [plain_render(ctx, scenes[si], p, W, H) for si, p, W, H in reqs]
    srv = rlevox.Server(ctx, memory_budget=2 << 30, max_batch_views=8, streams=3)
    try:
        tickets = []
        for k, (si, p, W, H) in enumerate(reqs):
            tickets.append(srv.submit(scenes[si], p, W, H))
            if k % 7 == 3:
                time.sleep(0.002)  # later arrivals join while earlier batches run
        for k, t in enumerate(tickets):
            col, dep, lat = srv.wait(t, timeout_ms=60000)
            assert np.array_equal(col, want[k][0]), "request %d colour" % k
            assert np.array_equal(dep, want[k][1]), "request %d depth" % k
            assert lat > 0
            if k % 8 == 0:  # a sample straight against the oracle
                si, p, W, H = reqs[k]
                ref = oracle.render(hosts[si].mips, p, W, H)
                assert np.array_equal(col, ref["colour"]), "request %d colour vs oracle" % k
                assert np.array_equal(dep, ref["depth"]), "request %d depth vs oracle" % k
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["submitted"] == len(reqs)
        assert st["batches"] >= 5  # at most 8 views per batch
        assert st["rays"] == sum(W * H for _, _, W, H in reqs)
        assert st["busy_seconds"] > 0
        assert st["buffer_growths"] >= 1  # lanes grew their buffers for larger batches
    finally:
        srv.close()


def test_large_lanes_and_batch_pixel_cap(ctx, scenes, c1_scene, tiny_scene, oracle):
    """Large requests on a lane of their own (fewer persistent wavefronts),
    the others on the remaining lanes with a pixel cap per batch: every
    framebuffer equals the plain render, a sample equals the oracle, and no
    batch mixes the two classes or exceeds the cap (beyond its first view)."""
    reqs = requests(c1_scene, tiny_scene, 30, seed=33)
    hosts = [c1_scene, tiny_scene]
    big = 1280 * 720
    srv = rlevox.Server(ctx, memory_budget=2 << 30, max_batch_views=16, streams=3,
                        large_lanes=1, large_pixels=big, large_waves=8,
                        batch_pixels=1920 * 1080)
    try:
        ts = [srv.submit(scenes[si], p, W, H) for si, p, W, H in reqs]
        for k, ((si, p, W, H), t) in enumerate(zip(reqs, ts)):
            col, dep, _ = srv.wait(t, timeout_ms=60000)
            want = plain_render(ctx, scenes[si], p, W, H)
            assert np.array_equal(col, want[0]) and np.array_equal(dep, want[1]), k
            if k % 6 == 1:
                ref = oracle.render(hosts[si].mips, p, W, H)
                assert np.array_equal(col, ref["colour"]) and np.array_equal(dep, ref["depth"]), k
        st = srv.stats()
        assert st["completed"] == len(reqs)
        # the pixel cap alone forces at least this many batches
        small = sum(W * H for _, _, W, H in reqs if W * H < big)
        assert st["batches"] >= -(-small // (1920 * 1080))
    finally:
        srv.close()
    for bad in (dict(large_lanes=3), dict(large_lanes=1), dict(large_waves=2)):
        kw = dict(streams=3, large_pixels=0)
        kw.update(bad)
        with pytest.raises(rlevox.RlevoxError) as e:
            rlevox.Server(ctx, **kw)
        assert e.value.name == "INVALID_ARG"


def test_alone_and_in_a_mix_are_identical(ctx, scenes, c1_scene, tiny_scene):
    reqs = requests(c1_scene, tiny_scene, 24, seed=5)
    probe = reqs[7]
    srv = rlevox.Server(ctx, max_batch_views=16, streams=2)
    try:
        alone = srv.wait(srv.submit(scenes[probe[0]], probe[1], probe[2], probe[3]))
        ts = [srv.submit(scenes[si], p, W, H) for si, p, W, H in reqs]
        mixed = [srv.wait(t) for t in ts]
        assert np.array_equal(alone[0], mixed[7][0]) and np.array_equal(alone[1], mixed[7][1])
    finally:
        srv.close()


def test_memory_admission_serialises_but_completes(ctx, scenes, c1_scene):
    poses = workloads.random_views(c1_scene.mips[0], 6, seed=8)
    # device bytes of one request: ray buffer + clip + gap-fill temp + framebuffers
    need = []
    for p in poses:
        rm = rlevox.raymap(p, 1920, 1080)
        need.append(rm.map_line_count * (rm.rs * 4 + 8) + 1920 * 1080 * 8)
    assert 2 * min(need) > max(need) + (2 << 20)  # no two requests fit together
    srv = rlevox.Server(ctx, memory_budget=max(need) + (2 << 20), max_batch_views=8, streams=4)
    try:
        ts = [srv.submit(scenes[0], p, 1920, 1080) for p in poses]
        outs = [srv.wait(t) for t in ts]
        st = srv.stats()
        assert st["batches"] == len(poses)  # one request in flight at a time
        for p, (col, dep, _) in zip(poses, outs):
            ref = plain_render(ctx, scenes[0], p, 1920, 1080)
            assert np.array_equal(col, ref[0]) and np.array_equal(dep, ref[1])
        with pytest.raises(rlevox.RlevoxError) as e:
            srv.submit(scenes[0], poses[0], 3840, 2160)
        assert e.value.name == "OVER_BUDGET"
        assert srv.stats()["rejected"] == 1
    finally:
        srv.close()


def test_queue_limit_cancel_and_timeout(ctx, scenes, c1_scene):
    poses = workloads.random_views(c1_scene.mips[0], 12, seed=9)
    srv = rlevox.Server(ctx, max_batch_views=1, max_queue=3, streams=1)
    try:
        ts, full = [], 0
        for p in poses:
            try:
                ts.append(srv.submit(scenes[0], p, 3840, 2160))
            except rlevox.RlevoxError as e:
                assert e.name == "QUEUE_FULL"
                full += 1
        assert full >= 1 and len(ts) >= 3
        srv.cancel(ts[-1])
        with pytest.raises(rlevox.RlevoxError) as e:
            srv.wait(ts[-1])
        assert e.value.name == "CANCELLED"
        for t in ts[:-1]:
            srv.wait(t)
        t = srv.submit(scenes[0], poses[0], 3840, 2160)
        try:
            srv.wait(t, timeout_ms=0)
        except rlevox.RlevoxError as e:
            assert e.name == "NOT_READY"
        srv.wait(t)  # the ticket is still valid after a time-out
        with pytest.raises(rlevox.RlevoxError):
            srv.wait(t)  # consumed
    finally:
        srv.close()


def test_server_rejects_bad_config_and_requests(ctx, scenes, c1_scene):
    for kw in [dict(memory_budget=1000), dict(max_batch_views=0), dict(max_queue=0),
               dict(streams=9)]:
        with pytest.raises(rlevox.RlevoxError):
            rlevox.Server(ctx, **kw)
    srv = rlevox.Server(ctx)
    try:
        p = workloads.random_views(c1_scene.mips[0], 1, seed=2)[0]
        with pytest.raises(rlevox.RlevoxError):
            srv.submit(scenes[0], p, 100, 200)  # H > W
        assert srv.stats()["rejected"] == 1
    finally:
        srv.close()


def test_destroy_with_requests_in_flight(ctx, scenes, c1_scene):
    poses = workloads.random_views(c1_scene.mips[0], 10, seed=3)
    srv = rlevox.Server(ctx, max_batch_views=2, streams=2)
    for p in poses:
        srv.submit(scenes[0], p, 1920, 1080)
    srv.close()  # drains the batches on the device, drops the queue


def test_autotune_keeps_results(ctx, scenes, c1_scene, oracle):
    poses = workloads.random_views(c1_scene.mips[0], 3, seed=12)
    bufs = [(ctx.alloc(1280 * 720 * 2), ctx.alloc(1280 * 720 * 4)) for _ in poses]
    views = [rlevox.make_view(scenes[0], p, 1280, 720, c, d) for p, (c, d) in zip(poses, bufs)]
    want = [plain_render(ctx, scenes[0], p, 1280, 720) for p in poses]
    best, ms = ctx.autotune(views)
    assert ms > 0
    assert ctx.get_tuning().as_tuple() == best.as_tuple()
    for p, w in zip(poses, want):
        got = plain_render(ctx, scenes[0], p, 1280, 720)
        assert np.array_equal(got[0], w[0]) and np.array_equal(got[1], w[1])
        ref = oracle.render(c1_scene.mips, p, 1280, 720)  # the tuned render vs the oracle
        assert np.array_equal(got[0], ref["colour"]) and np.array_equal(got[1], ref["depth"])
    ctx.set_tuning()
    for c, d in bufs:
        c.free()
        d.free()


def test_autotune_per_scene_and_resolution(ctx, scenes, c1_scene, tiny_scene):
    """Each (scene, W, H) among the sample views keeps its own winner; the
    first one also becomes the context's tuning; setting a tuning forgets
    them; renders stay identical throughout."""
    spec = [(0, 1280, 720), (1, 640, 480), (0, 1280, 720), (1, 640, 480)]
    hosts = [c1_scene, tiny_scene]
    poses = [workloads.random_views(hosts[si].mips[0], 1, seed=40 + i,
                                    alt=(2.0, 40.0) if si else (10.0, 120.0))[0]
             for i, (si, W, H) in enumerate(spec)]
    bufs = [(ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)) for _, W, H in spec]
    views = [rlevox.make_view(scenes[si], p, W, H, c, d)
             for (si, W, H), p, (c, d) in zip(spec, poses, bufs)]
    want = [plain_render(ctx, scenes[si], p, W, H) for (si, W, H), p in zip(spec, poses)]
    try:
        best, ms = ctx.autotune(views)
        assert ms > 0
        t0 = ctx.get_tuning_for(scenes[0], 1280, 720).as_tuple()
        t1 = ctx.get_tuning_for(scenes[1], 640, 480).as_tuple()
        assert t0 == best.as_tuple() == ctx.get_tuning().as_tuple()
        assert t1[0] in (1, 16, 32, 64) and t1[1] in (1, 2, 4)
        for (si, W, H), p, w in zip(spec, poses, want):
            got = plain_render(ctx, scenes[si], p, W, H)
            assert np.array_equal(got[0], w[0]) and np.array_equal(got[1], w[1])
        ctx.set_tuning(lines_per_wave=64, waves_per_block=1)
        assert ctx.get_tuning_for(scenes[1], 640, 480).as_tuple()[:2] == (64, 1)
    finally:
        ctx.set_tuning()
        for c, d in bufs:
            c.free()
            d.free()


def test_device_delivery(ctx, scenes, c1_scene, tiny_scene):
    reqs = requests(c1_scene, tiny_scene, 16, seed=31)
    bufs = [(ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)) for _, _, W, H in reqs]
    srv = rlevox.Server(ctx, max_batch_views=6, streams=2)
    try:
        ts = [srv.submit(scenes[si], p, W, H, c, d) for (si, p, W, H), (c, d) in zip(reqs, bufs)]
        for t in ts:
            col, dep, lat = srv.wait(t)
            assert col is None and lat > 0
        for (si, p, W, H), (c, d) in zip(reqs, bufs):
            ref = plain_render(ctx, scenes[si], p, W, H)
            assert np.array_equal(c.download(np.uint16, W * H).reshape(H, W), ref[0])
            assert np.array_equal(d.download(np.int32, W * H).reshape(H, W), ref[1])
        with pytest.raises(rlevox.RlevoxError):
            srv.submit(scenes[0], reqs[0][1], reqs[0][2], reqs[0][3], bufs[0][0], None)
    finally:
        srv.close()
        for c, d in bufs:
            c.free()
            d.free()
