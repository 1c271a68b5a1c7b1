"""Bit-level pins of the oracle's fine arithmetic (CPU).

The image-level pins (tests/test_oracle_pins.py) bound the oracle's render
within a tolerance of an independent raycaster; they cannot see an off-by-
one pixel at a span end or a depth rounded instead of truncated.  These
checks compare the oracle bit for bit with float64 models built from the
geometry (tests/fine_pins.py), away from rounding boundaries:

* compose: a ray buffer whose texels encode their own (line, position) is
  composed, and every pixel must show the texel the shader's texture
  coordinate selects with GL_NEAREST (floor), from a float64 evaluation of
  colorize_buddha_soft.frag:12-77;
* line clip ranges: the cleared range of every line from its 2D end points
  by the rule of Cuda_Render.h:203-250;
* spans: a scene holding one voxel column whose colours number its voxels;
  on every ray line through its cell, the drawn texels, each texel's colour
  index and depth must equal what the projection of the column's edges and
  the perspective-correct interpolation along its near edge give
  (Cuda_Render.h:456-733): span [trunc(bottom - 1), trunc(top)), colour
  index clamped to [texture, tex - 1], depth truncated with bit 0 cleared.

Planted mistakes (ORC_PLANT 6-10: span bottom without the -1, colour index
clamped at tex, p_add swapped, depth rounded, compose row rounded) and the
older plants 1-4 must each fail at least one of these checks.
"""
import numpy as np
import pytest

import fine_pins as F
import rlevox
import workloads as W
from oracle import UNTOUCHED, Oracle

SKY_DEPTH = 0x7FFFFFFF
CLEAR_TEXEL = 0x00FF8844
SIZES = [(512, 384), (640, 480), (320, 240), (800, 450)]


def _view_poses(n, seed, pos):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        pitch = float(rng.choice([-1, 1]) * rng.uniform(0.02, 1.4))
        yaw = float(rng.uniform(0, 2 * np.pi))
        out.append((W.pose(pos[0], pos[1], pos[2], pitch, yaw), SIZES[i % len(SIZES)]))
    return out


def compose_violations(orc, n=8):
    """Pixels (away from texel edges) whose composed texel is not the one
    the shader's texture coordinate selects; and the fraction of pixels
    skipped as ambiguous."""
    bad, near_frac = 0, 0.0
    for pose, (Wd, Hd) in _view_poses(n, 1, (100.3, -50.0, 200.7)):
        rm = orc.raymap(pose, Wd, Hd)
        lines, rs = rm.map_line_count, rm.rs
        ln = np.arange(lines, dtype=np.uint32)[:, None]
        pos = np.arange(rs, dtype=np.uint32)[None, :]
        raybuf = (pos | ((2 * ln) << 16)).astype(np.uint32)  # colour = position, depth = 2*line
        clip = np.zeros((lines, 2), np.int32)
        clip[:, 1] = rs - 1
        col, dep = orc.compose(rm, raybuf, clip)
        tr, tc, near, finite = F.shader_texel(rm)
        inside = finite & (tr >= 0) & (tr < lines)
        drawn = dep < SKY_DEPTH
        got_tc = np.where(drawn, col.astype(np.int64), -1)
        got_tr = np.where(drawn, dep.astype(np.int64) // 2, -1)
        bad += int((inside & ~near & ((got_tc != tc) | (got_tr != tr))).sum())
        bad += int((~inside & drawn).sum())  # outside the ray buffer: sky
        near_frac = max(near_frac, float(near.mean()))
    return bad, near_frac


def clip_violations(orc, n=8):
    solid = np.zeros((32, 32, 32), np.uint8)
    sc = rlevox.encode_dense(solid, np.zeros_like(solid, np.uint16), 1)
    bad = lines = 0
    for pose, (Wd, Hd) in _view_poses(n, 2, (20.3, -20.0, 10.7)):
        rm = orc.raymap(pose, Wd, Hd)
        _, clip, _ = orc.raycast(sc.mips, rm)
        want = F.clip_ranges(rm, orc.line_planes(rm))
        bad += int(np.any(want != clip[:rm.map_line_count], axis=1).sum())
        lines += rm.map_line_count
    return bad, lines


# the lone column: cell (150, 140), voxels at rows 40..99, colour 0x1000 + index
COL_CELL, COL_TOP, COL_H = (150, 140), 40, 60


@pytest.fixture(scope="module")
def lone_column():
    S, SY = 256, 128
    cx, cz = COL_CELL
    solid = np.zeros((S, SY, S), np.uint8)
    colour = np.zeros((S, SY, S), np.uint16)
    solid[cz, COL_TOP:COL_TOP + COL_H, cx] = 1
    colour[cz, COL_TOP:COL_TOP + COL_H, cx] = 0x1000 + np.arange(COL_H)
    return rlevox.encode_dense(solid, colour, 1)


def column_poses(n=15, seed=5):
    """Cameras 4-40 cells from the column, eye above its top, level with
    it, below its bottom and above the volume; facing it within +-0.5 rad;
    looking up and down."""
    rng = np.random.default_rng(seed)
    cx, cz = COL_CELL
    out = []
    for i in range(n):
        eye_row = [20, 55, 80, 110, -30][i % 5]
        ang, dist = rng.uniform(0, 2 * np.pi), rng.uniform(4, 40)
        ex, ez = cx + 0.5 + dist * np.cos(ang), cz + 0.5 + dist * np.sin(ang)
        yaw = float(np.arctan2(cx + 0.5 - ex, -(cz + 0.5 - ez)) + rng.uniform(-0.5, 0.5))
        pitch = float(rng.choice([-1, 1]) * rng.uniform(0.02, 0.9))
        out.append((W.pose(float(ex), float(-eye_row), float(ez), pitch, yaw),
                    [(512, 384), (640, 480), (400, 300)][i % 3]))
    return out


def span_violations(orc, scene):
    """Lines whose drawn texels differ from the column's projected span, and
    texels whose colour index or depth differ from the model; plus counts of
    what was checked."""
    st = dict(lines=0, texels=0, bad_lines=0, bad_u=0, bad_depth=0, skipped=0)
    for pose, (Wd, Hd) in column_poses():
        rm = orc.raymap(pose, Wd, Hd, 0, 200)  # 200 DDA steps: no mip switch, no wrap
        raybuf, clip, _ = orc.raycast(scene.mips, rm)
        planes = orc.line_planes(rm)
        for L in range(rm.map_line_count):
            if clip[L, 1] < clip[L, 0]:
                continue
            row = raybuf[L]
            drawn = np.nonzero((row != CLEAR_TEXEL) & (row != UNTOUCHED))[0]
            e = F.column_span(rm, planes[L], clip[L], COL_CELL, COL_TOP, COL_H)
            if e is None:  # the line's ray misses the cell
                st["bad_lines"] += int(len(drawn) > 0)
                continue
            if isinstance(e, str) or e["uncertain_span"]:
                st["skipped"] += 1
                continue
            st["lines"] += 1
            if not np.array_equal(drawn, e["ys"]):
                st["bad_lines"] += 1
                continue
            st["texels"] += len(drawn)
            t = row[drawn]
            u = (t & 0xFFFF).astype(np.int64) - 0x1000
            d = (t >> 16).astype(np.int64)
            ok = ~e["ambiguous"]
            st["bad_u"] += int(((u != e["u"]) & ok).sum())
            st["bad_depth"] += int(((d != e["depth"]) & ok).sum())
    return st


def test_compose_samples_the_shader_texel(oracle):
    bad, near = compose_violations(oracle)
    assert bad == 0
    assert near < 0.02  # pixels left unchecked as too close to a texel edge


def test_clip_ranges_closed_form(oracle):
    bad, lines = clip_violations(oracle)
    assert lines > 5000 and bad == 0


def test_lone_column_spans(oracle, lone_column):
    st = span_violations(oracle, lone_column)
    assert st["lines"] >= 300 and st["texels"] >= 30000, st
    assert st["skipped"] <= 0.1 * st["lines"], st
    assert st["bad_lines"] == 0 and st["bad_u"] == 0 and st["bad_depth"] == 0, st


@pytest.mark.parametrize("plant", [1, 2, 3, 4, 6, 7, 8, 9, 10])
def test_planted_fine_mistakes_are_caught(plant, lone_column):
    orc = Oracle(plant)
    assert orc.plant == plant
    found = []
    if compose_violations(orc, n=4)[0] > 0:
        found.append("compose")
    if clip_violations(orc, n=4)[0] > 0:
        found.append("clip")
    st = span_violations(orc, lone_column)
    if st["bad_lines"] or st["bad_u"] or st["bad_depth"]:
        found.append("spans")
    assert found, "planted mistake %d passes every fine pin" % plant
