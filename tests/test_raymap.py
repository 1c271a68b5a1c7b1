"""The per-frame ray map (RayMap.h:64-402 restated in csrc/raymap.cpp): the
product equals the oracle byte for byte, its geometry follows the camera
model's closed forms, and unsupported poses are rejected (CPU)."""
import ctypes as C
import math

import numpy as np
import pytest

import rlevox
import workloads as W


def random_pose(rng):
    pitch = float(rng.uniform(0.01, 1.55)) * (1 if rng.random() < 0.5 else -1)
    return W.pose(float(rng.uniform(-5000, 5000)), float(rng.uniform(-2000, 200)),
                  float(rng.uniform(-5000, 5000)), pitch, float(rng.uniform(-7, 7)))


SIZES = [(16, 16), (320, 240), (640, 480), (1280, 720), (1920, 1080), (3840, 2160), (8192, 8192),
         (1000, 999)]


def test_product_raymap_equals_oracle(oracle):
    rng = np.random.default_rng(11)
    n = 0
    for i in range(1500):
        pose = random_pose(rng)
        Wd, H = SIZES[i % len(SIZES)]
        mipd = int(rng.choice([0, 64, 256, 4096]))
        far = int(rng.choice([0, 512, 1 << 22]))
        try:
            a = rlevox.raymap(pose, Wd, H, mipd, far)
        except rlevox.RlevoxError:
            with pytest.raises(ValueError):
                oracle.raymap(pose, Wd, H, mipd, far)
            continue
        b = oracle.raymap(pose, Wd, H, mipd, far)
        assert a.to_bytes() == b.to_bytes(), (i, Wd, H)
        n += 1
    assert n > 1200


def test_vanishing_point_closed_form():
    """Vertical lines meet where the camera ray is parallel to the y axis:
    screen y = (1 - cot(pitch) * W/H) / 2 on the centre column."""
    rng = np.random.default_rng(3)
    for _ in range(300):
        pose = random_pose(rng)
        Wd, H = SIZES[int(rng.integers(len(SIZES)))]
        rm = rlevox.raymap(pose, Wd, H)
        pitch = pose.rot[0]
        expect = (1.0 - math.tan(math.pi / 2 - pitch) * Wd / H) / 2
        assert abs(rm.vanish[0] - 0.5) < 1e-6
        assert abs(rm.vanish[1] - expect) <= 1e-5 * max(1.0, abs(expect))  # fp32 matrices


def screen_segments(rm):
    """Which of the compose shader's four screen segments hold pixel centres
    (colorize_buddha_soft.frag:27-34)."""
    Wd, H = rm.W, rm.H
    ys, xs = np.mgrid[0:H, 0:Wd]
    scx = (xs + 0.5) / Wd - rm.vanish[0]
    scy = (H - ys - 0.5) / H - rm.vanish[1]
    side = np.abs(scy) - np.abs(scx) * Wd / H <= 0
    up = ~side & (scy > 0)
    dn = ~side & (scy <= 0)
    return [int(up.sum()), int(dn.sum()), int((side & (scx > 0)).sum()), int((side & (scx <= 0)).sum())]


def test_segment_line_counts_follow_the_screen_partition():
    rng = np.random.default_rng(5)
    for _ in range(200):
        pose = random_pose(rng)
        Wd, H = SIZES[int(rng.integers(1, 5))]
        rm = rlevox.raymap(pose, Wd, H)
        px = screen_segments(rm)
        res = list(rm.res)
        assert sum(res) == rm.map_line_count <= 4 * rm.rs
        for k in (0, 1):
            if px[k] == 0:
                assert res[k] == 0, (res, px)
            elif px[k] > 2 * Wd:
                assert res[k] > 0, (res, px)
        sides = px[2] + px[3]
        if sides == 0:
            assert res[2] == 0 and res[3] == 0
        elif min(px[2], px[3]) > 2 * H:
            assert res[2] > 0 and res[3] > 0


def test_map_line_count_and_sizes():
    rm = rlevox.raymap(W.pose(10, -50, 10, 0.3, 0.2), 1920, 1080)
    assert rm.rs == rm.maxres == 1920
    assert all(0 <= r <= 3 * rm.rs for r in rm.res)  # each segment holds at most 3*rs lines
    assert rm.map_line_count == sum(rm.res)
    assert 1 <= rm.mapswitch and rm.far_distance >= 1


@pytest.mark.parametrize("pose,W,H,params,why", [
    (W.pose(0, -10, 0, 0.3, 0.0), 640, 800, (0, 0), "H > W"),
    (W.pose(0, -10, 0, 0.3, 0.0), 8200, 600, (0, 0), "W too large"),
    (W.pose(0, -10, 0, 0.3, 0.0), 15, 15, (0, 0), "too small"),
    (W.pose(0, -10, 0, 0.005, 0.0), 640, 480, (0, 0), "pitch too flat"),
    (W.pose(0, -10, 0, 3.14, 0.0), 640, 480, (0, 0), "pitch beyond the range"),
    (W.pose(float("nan"), -10, 0, 0.3, 0.0), 640, 480, (0, 0), "non-finite position"),
    (W.pose(2e6, -10, 0, 0.3, 0.0), 640, 480, (0, 0), "position too far out"),
    (W.pose(0, -10, 0, 1.5, 0.0), 640, 480, (1, 0), "mip distance gives mapswitch < 1"),
    (W.pose(0, -10, 0, 0.3, 0.0), 640, 480, (0, 1 << 23), "far distance too large"),
    (W.pose(0, 9e5, 0, 0.3, 0.0), 640, 480, (1 << 20, 0), "mip-switch doublings would overflow"),
])
def test_rejected_poses(pose, W, H, params, why):
    with pytest.raises(rlevox.RlevoxError) as e:
        rlevox.raymap(pose, W, H, *params)
    assert e.value.name == "INVALID_ARG", why


def test_roll_is_rejected():
    p = rlevox.Pose.make((0, -10, 0), (0.3, 0.1, 0.2))
    with pytest.raises(rlevox.RlevoxError):
        rlevox.raymap(p, 640, 480)
