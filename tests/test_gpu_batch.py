"""Batched rendering (persistent work-queue launch, and its captured graph)
gives every view the same framebuffer as a plain per-view render, which in
turn matches the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import rlevox  # noqa: E402
import workloads  # noqa: E402

RES = [(320, 240), (640, 480), (512, 512), (1280, 720), (1920, 1080)]


@pytest.fixture(scope="module")
def scenes(ctx, c1_scene, tiny_scene):
    hs2 = workloads.gen("c1", seed=2)
    hosts = [c1_scene, tiny_scene, hs2]
    devs = [ctx.upload(h.mips) for h in hosts]
    yield hosts, devs
    for d in devs:
        d.free()


def make_views(ctx, hosts, devs, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        si = int(rng.integers(0, len(devs)))
        W, H = RES[int(rng.integers(0, len(RES)))]
        alt = (2.0, 40.0) if hosts[si].mips[0]["sx"] <= 64 else (10.0, 120.0)
        p = workloads.random_views(hosts[si].mips[0], 1, seed=int(rng.integers(1 << 30)), alt=alt)[0]
        dc, dd = ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)
        out.append({"si": si, "pose": p, "W": W, "H": H, "dc": dc, "dd": dd,
                    "view": rlevox.make_view(devs[si], p, W, H, dc, dd)})
    return out


def plain_reference(ctx, devs, v, params=None):
    W, H = v["W"], v["H"]
    dc, dd = ctx.alloc(W * H * 2), ctx.alloc(W * H * 4)
    ctx.render(devs[v["si"]], v["pose"], W, H, dc, dd, params=params)
    c = dc.download(np.uint16, W * H)
    d = dd.download(np.int32, W * H)
    dc.free()
    dd.free()
    return c, d


def check_views(ctx, devs, views, params=None, tag=""):
    for k, v in enumerate(views):
        W, H = v["W"], v["H"]
        c = v["dc"].download(np.uint16, W * H)
        d = v["dd"].download(np.int32, W * H)
        rc, rd = plain_reference(ctx, devs, v, params)
        assert np.array_equal(d, rd), "%s view %d (%dx%d): depth differs" % (tag, k, W, H)
        assert np.array_equal(c, rc), "%s view %d (%dx%d): colour differs" % (tag, k, W, H)


def free_views(views):
    for v in views:
        v["dc"].free()
        v["dd"].free()


def test_batch_matches_plain(ctx, scenes):
    hosts, devs = scenes
    views = make_views(ctx, hosts, devs, 24, seed=1)
    ctx.render_batch([v["view"] for v in views])
    check_views(ctx, devs, views, tag="batch")
    free_views(views)


def test_batch_matches_oracle(ctx, scenes, oracle):
    """24 views over three scenes, sizes 320x240 to 1920x1080, in one batch:
    every view's framebuffer equals the oracle's render."""
    hosts, devs = scenes
    views = make_views(ctx, hosts, devs, 24, seed=2)
    ctx.render_batch([v["view"] for v in views])
    for v in views:
        W, H = v["W"], v["H"]
        ref = oracle.render(hosts[v["si"]].mips, v["pose"], W, H)
        assert np.array_equal(v["dc"].download(np.uint16, W * H).reshape(H, W), ref["colour"])
        assert np.array_equal(v["dd"].download(np.int32, W * H).reshape(H, W), ref["depth"])
    free_views(views)


def test_batch_compose_only_and_view_order(ctx, scenes):
    hosts, devs = scenes
    prm = rlevox.RenderParams(0, 0, rlevox.POST_COMPOSE, 0)
    for order in (0, 1):
        ctx.set_tuning(order=order)
        views = make_views(ctx, hosts, devs, 10, seed=3 + order)
        ctx.render_batch([v["view"] for v in views], params=prm)
        check_views(ctx, devs, views, params=prm, tag="compose-only order %d" % order)
        free_views(views)
    ctx.set_tuning()


def test_batch_graph_replays(ctx, scenes, oracle):
    """The captured graph is replayed on three different batches, with eager
    renders in between; every replay is checked against plain renders, and
    its first two views against the oracle."""
    hosts, devs = scenes
    for rep, n in enumerate((16, 5, 30)):
        views = make_views(ctx, hosts, devs, n, seed=10 + rep)
        ctx.render_batch([v["view"] for v in views], graph=True)
        for v in views[:2]:
            W, H = v["W"], v["H"]
            ref = oracle.render(hosts[v["si"]].mips, v["pose"], W, H)
            assert np.array_equal(v["dc"].download(np.uint16, W * H).reshape(H, W), ref["colour"])
            assert np.array_equal(v["dd"].download(np.int32, W * H).reshape(H, W), ref["depth"])
        check_views(ctx, devs, views, tag="graph replay %d" % rep)
        free_views(views)


def test_batch_shapes_equal_plain(ctx, scenes):
    """Every batch shape the batch autotuner can choose (run-cache depth,
    element replay, persistent wavefronts per CU, view order) renders the
    same framebuffers as plain renders."""
    hosts, devs = scenes
    views = make_views(ctx, hosts, devs, 12, seed=41)
    for rc, par, waves, order in ((24, 1, 8, 0), (56, 2, 24, 1), (40, 2, 12, 0), (56, 1, 16, 1)):
        ctx.set_tuning(persistent_waves=waves, order=order, batch_run_cache=rc, batch_par=par)
        for v in views:
            v["dc"].fill_u32(0)
            v["dd"].fill_u32(0)
        ctx.render_batch([v["view"] for v in views])
        check_views(ctx, devs, views, tag="batch rc %d par %d waves %d" % (rc, par, waves))
    ctx.set_tuning()
    free_views(views)


def test_batch_autotune(ctx, scenes, oracle):
    """The batch autotuner picks a batch shape, makes it the context's, and
    batches rendered with it (eager and graph) equal plain renders and, on a
    sample, the oracle."""
    hosts, devs = scenes
    views = make_views(ctx, hosts, devs, 16, seed=43)
    try:
        best, ms = ctx.autotune_batch([v["view"] for v in views])
        assert ms > 0
        t = ctx.get_tuning()
        assert t.as_tuple() == best.as_tuple()
        assert best.persistent_waves in (8, 12, 16, 24) and best.order in (0, 1)
        assert best.batch_run_cache in (24, 40, 56) and best.batch_par in (1, 2)
        for graph in (False, True):
            for v in views:
                v["dc"].fill_u32(0)
                v["dd"].fill_u32(0)
            ctx.render_batch([v["view"] for v in views], graph=graph)
            check_views(ctx, devs, views, tag="tuned batch graph=%s" % graph)
        for v in views[:3]:
            W, H = v["W"], v["H"]
            ref = oracle.render(hosts[v["si"]].mips, v["pose"], W, H)
            assert np.array_equal(v["dc"].download(np.uint16, W * H).reshape(H, W), ref["colour"])
            assert np.array_equal(v["dd"].download(np.int32, W * H).reshape(H, W), ref["depth"])
    finally:
        ctx.set_tuning()
        free_views(views)


def test_batch_rejects_bad_views(ctx, scenes):
    hosts, devs = scenes
    d = ctx.alloc(64)
    bad = rlevox.make_view(devs[0], workloads.pose(10, -10, 10, 0.3, 0.1), 200, 400, d, d)
    with pytest.raises(rlevox.RlevoxError) as e:
        ctx.render_batch([bad])
    assert e.value.name == "INVALID_ARG"
    flat = rlevox.make_view(devs[0], workloads.pose(10, -10, 10, 0.0, 0.1), 320, 240, d, d)
    with pytest.raises(rlevox.RlevoxError):
        ctx.render_batch([flat])
    d.free()
