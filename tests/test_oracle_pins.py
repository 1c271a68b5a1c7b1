"""Pins of the oracle (CPU).

The reference ships no tests, fixtures or golden images and cannot be built
here, so the oracle's restatement of render_line / compose / gap-fill is
pinned by checks that do not reuse its arithmetic:

* an independent double-precision per-pixel 3D raycaster (oracle/brute.cpp)
  agrees with the line renderer's composed image (colour within a 3x3
  neighbourhood, depth within a few percent);
* closed forms: an empty scene leaves only the clear texel inside the line
  clip range; written texels have depth bit 0 clear; a floor of single-voxel
  columns only shows colours of those columns;
* invariants: texels outside the line clip range are never written, the
  result does not depend on the number of host threads;
* the gap-fill restatement against a floating-point emulation of
  bin/shader/soft.frag:8-74 (the shader's own loops, on a texture of 1/z);
* planted mistakes (oracle built with ORC_PLANT=1..5) fail these checks.
"""
import math

import numpy as np
import pytest

import rlevox
import workloads as W
from oracle import UNTOUCHED, Oracle

SKY_DEPTH = 0x7FFFFFFF
SKY_COLOUR = 0x8844
CLEAR_TEXEL = 0x00FF8844


def nbr_match(a, b, r=1):
    H, Wd = a.shape
    ok = np.zeros(a.shape, bool)
    bp = np.pad(b, r, mode="edge")
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            ok |= a == bp[r + dy:r + dy + H, r + dx:r + dx + Wd]
    return ok


@pytest.fixture(scope="module")
def c1_dense(c1_scene, oracle):
    return oracle.decode(c1_scene.mips[0])


@pytest.fixture(scope="module")
def c1_poses(c1_scene):
    return W.flythrough(c1_scene.mips[0], frames=5, seed=3)[:3]


BANDS = [(25, 50), (50, 100), (100, 200)]


def brute_agreement(orc, scene, dense, pose, size=512):
    surf, col = dense
    r = orc.render(scene.mips, pose, size, size, gapfill=False)
    bc, bd, bh = Oracle().brute(surf, col, pose, size, size, 2000.0)
    ok = nbr_match(r["colour"], bc, 1)
    bands = []
    for lo, hi in BANDS:
        sel = (bh >= lo) & (bh < hi)
        if sel.sum() >= 500:  # bands with too few pixels say nothing
            bands.append((ok & sel).sum() / sel.sum())
    assert bands
    return bands, r, (bc, bd, bh)


# colour agreement thresholds: every populated distance band, and their mean
MIN_BAND, MEAN_BAND = 0.92, 0.945


def test_line_renderer_agrees_with_per_pixel_raycaster(oracle, c1_scene, c1_dense, c1_poses):
    allb = []
    for pose in c1_poses:
        bands, r, (bc, bd, bh) = brute_agreement(oracle, c1_scene, c1_dense, pose)
        assert min(bands) >= MIN_BAND, bands
        allb += bands
        # depth: same camera-space units, truncated and with bit 0 cleared
        sel = (bh >= 10) & (bh < 200) & (r["colour"] == bc) & (r["depth"] < SKY_DEPTH)
        ratio = r["depth"][sel] / np.maximum(bd[sel], 1)
        assert 0.98 <= np.median(ratio) <= 1.005
        assert np.mean(np.abs(ratio - 1.0) < 0.05) >= 0.97
        # terrain the per-pixel raycaster hits within 200 voxels is not sky,
        # away from pixel column 0 (the reference's line clip leaves it sky)
        near = (bh[:, 1:] >= 0) & (bh[:, 1:] < 200)
        assert np.mean(r["depth"][:, 1:][near] < SKY_DEPTH) >= 0.96
    assert np.mean(allb) >= MEAN_BAND, allb


@pytest.mark.parametrize("plant", [1, 2, 3, 4])
def test_planted_render_mistakes_are_caught(plant, c1_scene, c1_dense, c1_poses):
    orc = Oracle(plant)
    assert orc.plant == plant
    allb = []
    for pose in c1_poses:
        allb += brute_agreement(orc, c1_scene, c1_dense, pose)[0]
    assert min(allb) < MIN_BAND and np.mean(allb) < MEAN_BAND, allb


def empty_scene(sx=64, sy=64, sz=64, nmips=2):
    solid = np.zeros((sz, sy, sx), np.uint8)
    return rlevox.encode_dense(solid, np.zeros_like(solid, np.uint16), nmips)


def test_empty_scene_renders_clear_texels_and_sky(oracle):
    sc = empty_scene()
    pose = W.pose(20.3, -30.0, 40.7, 0.3, 1.1)
    rm = oracle.raymap(pose, 256, 192)
    raybuf, clip, ctr = oracle.raycast(sc.mips, rm)
    rows = np.arange(rm.rs)
    for ln in range(rm.map_line_count):
        lo, hi = clip[ln]
        inside = (rows >= lo) & (rows <= hi)
        assert np.all(raybuf[ln][inside] == CLEAR_TEXEL)
        assert np.all(raybuf[ln][~inside] == UNTOUCHED)
    assert ctr["col_loads"] == 0 and ctr["spans"] == 0
    col, dep = oracle.compose(rm, raybuf, clip)
    assert np.all(col == SKY_COLOUR) and np.all(dep == SKY_DEPTH)


def floor_scene(sx=64, sy=64, sz=64, row=40, seed=3):
    """One voxel per column at `row`; colour = a random 16-bit value per column."""
    rng = np.random.default_rng(seed)
    solid = np.zeros((sz, sy, sx), np.uint8)
    colour = np.zeros((sz, sy, sx), np.uint16)
    solid[:, row, :] = 1
    cols = rng.integers(0, 0xFF00, size=(sz, sx), dtype=np.uint16) & 0xFEFF
    colour[:, row, :] = cols
    return rlevox.encode_dense(solid, colour, 1), cols


def test_floor_of_single_voxel_columns(oracle):
    sc, cols = floor_scene()
    pose = W.pose(30.5, 20.0 - 40, 30.5, 0.6, 0.4)  # 20 voxels above the floor
    r = oracle.render(sc.mips, pose, 256, 256, gapfill=False)
    raybuf = r["raybuf"]
    written = (raybuf != UNTOUCHED) & (raybuf != CLEAR_TEXEL)
    assert written.sum() > 1000
    assert np.all(((raybuf[written] >> 16) & 1) == 0)  # depth bit 0 cleared
    hit = r["depth"] < SKY_DEPTH
    assert hit.mean() > 0.3
    assert set(np.unique(r["colour"][hit])) <= set(np.unique(cols))
    # the same floor seen by the per-pixel raycaster
    surf, col = oracle.decode(sc.mips[0])
    bc, bd, bh = oracle.brute(surf, col, pose, 256, 256, 2000.0)
    near = (bh >= 3) & (bh < 60)
    assert (nbr_match(r["colour"], bc) & near).sum() / near.sum() >= 0.95


def test_threads_do_not_change_results(oracle, c1_scene, c1_poses):
    rm = oracle.raymap(c1_poses[0], 320, 240)
    a = oracle.raycast(c1_scene.mips, rm, threads=1)
    b = oracle.raycast(c1_scene.mips, rm, threads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_writes_stay_inside_line_clip(oracle, c1_scene, c1_poses):
    for pose in c1_poses:
        rm = oracle.raymap(pose, 400, 300)
        raybuf, clip, ctr = oracle.raycast(c1_scene.mips, rm)
        rows = np.arange(rm.rs)[None, :]
        lo, hi = clip[:, :1], clip[:, 1:2]
        inside = (rows >= lo) & (rows <= hi)
        assert np.all(raybuf[~inside] == UNTOUCHED)
        assert not np.any(raybuf[inside] == UNTOUCHED)
        # lines whose screen segment degenerates return before the clear with
        # an empty clip range (0, -1) and are not counted
        assert ctr["lines"] == int((clip[:, 1] >= clip[:, 0]).sum())
        assert np.all((clip[:, 1] >= clip[:, 0]) | ((clip[:, 0] == 0) & (clip[:, 1] == -1)))
        assert ctr["cleared"] == int(inside.sum())
        assert ctr["col_loads"] == int((raybuf[inside] != CLEAR_TEXEL).sum())


# ------------------------------------------------------------------ gap-fill

def soft_frag(col, dep):
    """bin/shader/soft.frag:8-74 in float64 on a 2048-wide target texture.

    w = 1/z as compose stores it (colorize_buddha_soft.frag:90-91: z is the
    16-bit depth over 65280, w = 0.001/z); sky has w = 0.  Texture lookups
    are GL_NEAREST with clamp-to-edge; +y of the texture is up.  The filtered
    quantity is the shade byte of colour16, rounded to the nearest byte.

    Returns the filtered colours and a mask of pixels whose result sits on a
    decision boundary the integer restatement may resolve the other way: a
    tap offset within 0.005 px of a texel edge (the integer form keeps the
    radius in 1/256 px), or a tap's w within 0.1% of the 0.7*radmax cut.
    """
    H, Wd = col.shape
    w = np.where(dep < SKY_DEPTH, 0.001 * 65280.0 / np.maximum(dep, 1), 0.0)
    shade = (col & 0xFF).astype(np.float64)
    ys, xs = np.mgrid[0:H, 0:Wd]
    scale = 2048.0

    def tex(a, ox, oy):  # sample at pixel centre + (ox, oy) px, oy up
        xx = np.clip(np.floor(xs + 0.5 + ox), 0, Wd - 1).astype(int)
        yy = np.clip(np.floor(ys + 0.5 - oy), 0, H - 1).astype(int)
        return a[yy, xx]

    radmax = w.copy()
    a = 0.0
    while a < 3.1415 * 2.0:
        radmax = np.maximum(radmax, tex(w, np.round(math.sin(a) * 0.005 * scale),
                                        np.round(math.cos(a) * 0.005 * scale)))
        a += 3.1415 * 1.9 / 6.0
    rad = 0.0023 * radmax
    total = shade.copy()
    n = np.ones_like(shade)
    edge = np.zeros(col.shape, bool)
    for ta in [-1.0, -0.6, -0.2, 0.2, 0.6]:
        ox = ta * rad * scale
        edge |= np.abs((ox + 0.5) - np.round(ox + 0.5)) < 0.005
        for tb in [-1.0, -0.6, -0.2, 0.2, 0.6]:
            oy = tb * rad * scale
            wt = tex(w, ox, oy)
            q = (wt > 0) & (wt >= radmax * 0.7)
            edge |= (wt > 0) & (np.abs(wt - radmax * 0.7) <= 1e-3 * radmax)
            total += np.where(q, tex(shade, ox, oy), 0.0)
            n += q
    filt = (rad > 0.00008) & (n > 6)
    edge &= rad > 0.00008
    avg = np.floor(total / n + 0.5)
    out = np.where(filt, (col & 0xFF00) | avg.astype(np.uint16), col)
    return out.astype(np.uint16), edge


def synthetic_frame(seed, H=120, Wd=160):
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:H, 0:Wd]
    dep = (30 + 1500 * (ys / H) + 40 * np.sin(xs / 7.0)).astype(np.int32) & ~1
    dep[rng.random((H, Wd)) < 0.03] = SKY_DEPTH
    dep[:10, :] = SKY_DEPTH
    col = (rng.integers(0, 4, (H, Wd)) << 8 | rng.integers(0, 256, (H, Wd))).astype(np.uint16)
    col[dep == SKY_DEPTH] = SKY_COLOUR
    return col, dep


def test_gapfill_matches_shader_emulation(oracle):
    for seed in range(3):
        col, dep = synthetic_frame(seed)
        out = oracle.gapfill(col, dep)
        ref, edge = soft_frag(col, dep)
        same = out == ref
        assert same.mean() >= 0.99, same.mean()
        assert np.all(same | edge)  # every difference sits on a decision boundary
        assert edge.mean() < 0.08
        assert (out != col).mean() > 0.2  # the filter actually ran


def test_gapfill_closed_forms(oracle):
    H, Wd = 64, 96
    col = np.full((H, Wd), 0x0233, np.uint16)
    for d in [2, 100, 1874, 1876, 60000]:
        dep = np.full((H, Wd), d, np.int32)
        assert np.array_equal(oracle.gapfill(col, dep), col)  # uniform image is a fixed point
    dep = np.full((H, Wd), SKY_DEPTH, np.int32)
    noise = np.random.default_rng(0).integers(0, 65535, (H, Wd)).astype(np.uint16)
    assert np.array_equal(oracle.gapfill(noise, dep), noise)  # all sky: untouched
    # beyond the radius threshold (0.0023 * 2048 * 65.28 / d <= 0.00008 * 2048 px)
    dep = np.full((H, Wd), 1900, np.int32)
    assert np.array_equal(oracle.gapfill(noise, dep), noise)
    # near: every output shade lies within the range of the input shades
    dep = np.full((H, Wd), 50, np.int32)
    out = oracle.gapfill(noise, dep)
    assert np.all(out >> 8 == noise >> 8)
    assert (out != noise).mean() > 0.9


def test_planted_gapfill_mistake_is_caught():
    orc = Oracle(5)
    col, dep = synthetic_frame(0)
    ref, edge = soft_frag(col, dep)
    assert not np.all((orc.gapfill(col, dep) == ref) | edge)
