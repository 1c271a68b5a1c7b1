"""RLE4 scene format: the product's encoder, mip reduction, pointer map,
validation and .rle4 file round trip (CPU).

Pins: the dense decode of an encoded volume equals its surface (closed-form
on boxes, brute force on random volumes), the product encoder and the
oracle's restatement of RLE4::compress produce identical slabs and pointer
maps, run splitting follows Rle4.cpp:168-179, and mips follow the 3-of-8 rule
of tree.h:23-85 (counted directly with numpy).
"""
import os

import numpy as np
import pytest

import rlevox
import workloads as W


def random_volume(seed, sx=32, sy=48, sz=16, fill=0.5):
    rng = np.random.default_rng(seed)
    solid = (rng.random((sz, sy, sx)) < fill).astype(np.uint8)
    colour = rng.integers(0, 65536, (sz, sy, sx)).astype(np.uint16)
    return solid, colour


def surface_numpy(solid):
    """A solid voxel is stored unless all in-range voxels of its 3x3x3 block
    are solid and the block lies wholly inside the volume (Rle4.cpp:100-132
    counts only in-range neighbours, so a volume-edge voxel has < 27)."""
    sz, sy, sx = solid.shape
    p = np.pad(solid, 1, constant_values=0).astype(np.int32)
    cnt = np.zeros(solid.shape, np.int32)
    for a in range(3):
        for b in range(3):
            for c in range(3):
                cnt += p[a:a + sz, b:b + sy, c:c + sx]
    return ((solid == 1) & (cnt < 27)).astype(np.uint8)


def test_surface_of_a_box_is_its_shell(oracle):
    solid = np.zeros((20, 30, 24), np.uint8)
    solid[3:15, 5:25, 4:20] = 1
    surf = oracle.surface_dense(solid)
    shell = solid.copy()
    shell[4:14, 6:24, 5:19] = 0
    assert np.array_equal(surf, shell)
    full = np.ones((8, 8, 8), np.uint8)  # the volume's own faces are surface
    s = oracle.surface_dense(full)
    assert s[1:7, 1:7, 1:7].sum() == 0 and s.sum() == 8 ** 3 - 6 ** 3


@pytest.mark.parametrize("seed,fill", [(1, 0.5), (2, 0.9), (3, 0.05)])
def test_surface_brute_force(oracle, seed, fill):
    solid, _ = random_volume(seed, fill=fill)
    assert np.array_equal(oracle.surface_dense(solid), surface_numpy(solid))


@pytest.mark.parametrize("seed,fill", [(1, 0.5), (4, 0.97), (5, 0.02)])
def test_product_encoder_matches_oracle_and_round_trips(oracle, seed, fill):
    solid, colour = random_volume(seed, fill=fill)
    hs = rlevox.encode_dense(solid, colour, 1)
    m = hs.mips[0]
    slabs, ptr = oracle.encode_dense(solid, colour)
    assert np.array_equal(m["slabs"], slabs)
    assert np.array_equal(m["ptrmap"], ptr)
    surf, col = oracle.decode(m)
    assert np.array_equal(surf, surface_numpy(solid))
    assert np.array_equal(col[surf == 1], colour[surf == 1])
    assert rlevox.validate(hs.mips)[0] == 0


def test_long_runs_are_split():
    sx = sz = 4
    sy = 2048
    solid = np.zeros((sz, sy, sx), np.uint8)
    colour = np.zeros_like(solid, np.uint16)
    solid[1, 1500, 2] = 1              # skip 1500 > 1023: a skip-only element first
    solid[2, 100:200, 1] = 1           # a pillar: 100 stored voxels > 63 per element
    colour[2, 100:200, 1] = np.arange(100)
    hs = rlevox.encode_dense(solid, colour, 1)
    m = hs.mips[0]

    def column(x, z):
        c = z * sx + x
        ofs = int(m["ptrmap"][2 * c])
        cnt, ncol = int(m["slabs"][ofs]), int(m["slabs"][ofs + 1])
        els = [int(v) for v in m["slabs"][ofs + 2:ofs + 2 + cnt]]
        cols = [int(v) for v in m["slabs"][ofs + 2 + cnt:ofs + 2 + cnt + ncol]]
        return els, cols, int(m["ptrmap"][2 * c + 1])

    els, cols, info = column(2, 1)
    assert els == [1023, (1 << 10) | 477]
    assert info == 2 | (1023 << 16)  # count | first element << 16
    els, cols, info = column(1, 2)
    assert els == [(63 << 10) | 100, (37 << 10) | 0]
    assert cols == list(range(100))
    els, cols, info = column(0, 0)
    assert els == [] and cols == [] and info == 0


def mip_numpy(solid, colour):
    """3-of-8 rule; colour of the third solid child in (x, y, z) nesting order."""
    sz, sy, sx = solid.shape
    s = solid.reshape(sz // 2, 2, sy // 2, 2, sx // 2, 2)
    c = colour.reshape(sz // 2, 2, sy // 2, 2, sx // 2, 2)
    # children ordered with x slowest, then y, then z (subx, suby, subz loops)
    sc = s.transpose(0, 2, 4, 5, 3, 1).reshape(sz // 2, sy // 2, sx // 2, 8)
    cc = c.transpose(0, 2, 4, 5, 3, 1).reshape(sz // 2, sy // 2, sx // 2, 8)
    run = np.cumsum(sc, axis=-1)
    on = run[..., -1] >= 3
    third = np.argmax((run == 3) & (sc == 1), axis=-1)
    col = np.take_along_axis(cc, third[..., None], -1)[..., 0]
    return on.astype(np.uint8), np.where(on, col, 0).astype(np.uint16)


def test_mip_rule(oracle):
    solid, colour = random_volume(7, fill=0.4)
    s2, c2 = oracle.mip_dense(solid, colour)
    ns, nc = mip_numpy(solid, colour)
    assert np.array_equal(s2, ns) and np.array_equal(c2, nc)
    # the product's mip chain: level 1 decodes to the surface of the reduced volume
    hs = rlevox.encode_dense(solid, colour, 3)
    assert len(hs.mips) == 3
    surf, col = oracle.decode(hs.mips[1])
    assert np.array_equal(surf, surface_numpy(ns))
    assert np.array_equal(col[surf == 1], nc[surf == 1])
    s4, c4 = mip_numpy(ns, nc)
    surf, col = oracle.decode(hs.mips[2])
    assert np.array_equal(surf, surface_numpy(s4))
    assert np.array_equal(col[surf == 1], c4[surf == 1])


def test_generated_scene_is_valid_and_deterministic():
    a = W.gen("tiny", seed=5)
    b = W.gen("tiny", seed=5, threads=1)
    assert rlevox.validate(a.mips)[0] == 0
    assert len(a.mips) == 3
    for ma, mb in zip(a.mips, b.mips):
        assert np.array_equal(ma["slabs"], mb["slabs"]) and np.array_equal(ma["ptrmap"], mb["ptrmap"])
        assert np.array_equal(rlevox.build_ptrmap(ma["slabs"], ma["sx"], ma["sz"]), ma["ptrmap"])
    c = W.gen("tiny", seed=6)
    assert not np.array_equal(a.mips[0]["slabs"], c.mips[0]["slabs"])


def test_generated_scene_mips_follow_the_rule(oracle):
    sc = W.gen("tiny", seed=5)
    # level 0 decodes; its volume is not recoverable (interior voxels are not
    # stored), so check level 1 against the stored surface's own occupancy
    # only through validity and size relations
    surf0, _ = oracle.decode(sc.mips[0])
    surf1, _ = oracle.decode(sc.mips[1])
    assert surf1.shape == tuple(s // 2 for s in surf0.shape)
    assert surf0.sum() > 0 and surf1.sum() > 0


def test_rle4_file_round_trip(tmp_path):
    sc = W.gen("tiny", seed=5)
    path = str(tmp_path / "tiny.rle4")
    rlevox.rle4_save(path, sc.mips)
    back = rlevox.rle4_load(path)
    assert len(back.mips) == len(sc.mips)
    for a, b in zip(sc.mips, back.mips):
        assert (a["sx"], a["sy"], a["sz"]) == (b["sx"], b["sy"], b["sz"])
        assert np.array_equal(a["slabs"], b["slabs"]) and np.array_equal(a["ptrmap"], b["ptrmap"])
    data = open(path, "rb").read()
    open(path, "wb").write(data[:len(data) // 2])  # truncated file
    with pytest.raises(rlevox.RlevoxError):
        rlevox.rle4_load(path)
    with pytest.raises(rlevox.RlevoxError):
        rlevox.rle4_load(str(tmp_path / "missing.rle4"))


def copy_mips(mips):
    return [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in m.items()} for m in mips]


def test_validate_rejects_malformed_scenes():
    sc = W.gen("tiny", seed=5)
    good = copy_mips(sc.mips)
    assert rlevox.validate(good)[0] == 0

    bad = copy_mips(sc.mips)  # element run past the column height
    m = bad[0]
    c = next(i for i in range(m["sx"] * m["sz"]) if m["slabs"][m["ptrmap"][2 * i]] > 0)
    ofs = int(m["ptrmap"][2 * c])
    m["slabs"][ofs + 2] = (m["slabs"][ofs + 2] & 0xFC00) | 1023
    m["ptrmap"] = rlevox.build_ptrmap(m["slabs"], m["sx"], m["sz"])
    assert rlevox.validate(bad)[0] != 0

    bad = copy_mips(sc.mips)  # colour count disagrees with the drawn voxels
    m = bad[0]
    ofs = int(m["ptrmap"][2 * c])
    m["slabs"][ofs + 1] += 1
    assert rlevox.validate(bad)[0] != 0

    bad = copy_mips(sc.mips)  # pointer map 'first' field disagrees with the slab
    bad[0]["ptrmap"][2 * c + 1] ^= 1 << 16
    assert rlevox.validate(bad)[0] != 0

    bad = copy_mips(sc.mips)  # offset beyond the slab array
    bad[0]["ptrmap"][2 * c] = bad[0]["slabs"].size + 10
    assert rlevox.validate(bad)[0] != 0

    bad = copy_mips(sc.mips)  # grid not a power of two
    bad[0]["sx"] = 48
    assert rlevox.validate(bad)[0] != 0

    bad = copy_mips(sc.mips)  # mip sizes do not halve
    bad[1]["sy"] = bad[0]["sy"]
    assert rlevox.validate(bad)[0] != 0


@pytest.mark.parametrize("shift", [1, 3, 4])
def test_repack_with_offset_shift_keeps_the_scene(oracle, shift):
    sc = W.gen("tiny", seed=5)
    rp = rlevox.repack(sc.mips, shift)
    assert rlevox.validate(rp.mips)[0] == 0
    for a, b in zip(sc.mips, rp.mips):
        assert b["shift"] == shift
        assert b["slabs"].size >= a["slabs"].size
        assert np.all((b["ptrmap"][0::2].astype(np.uint64) << shift) < b["slabs"].size)
        assert np.array_equal(a["ptrmap"][1::2], b["ptrmap"][1::2])
        assert np.array_equal(rlevox.build_ptrmap(b["slabs"], b["sx"], b["sz"], shift), b["ptrmap"])
        da, db = oracle.decode(a), oracle.decode(b)
        assert np.array_equal(da[0], db[0]) and np.array_equal(da[1], db[1])
    pose = W.random_views(sc.mips[0], 1, seed=4, alt=(4.0, 30.0))[0]
    ra = oracle.render(sc.mips, pose, 160, 120)
    rb = oracle.render(rp.mips, pose, 160, 120)
    for k in ("raybuf", "clip", "colour", "depth"):
        assert np.array_equal(ra[k], rb[k]), k
    assert ra["counters"] == rb["counters"]
    # camera placement reads the widened offsets too: same surface rows, same poses
    m0a, m0b = sc.mips[0], rp.mips[0]
    for x, z in ((3.2, 7.9), (40.5, 12.1), (63.9, 63.9), (0.0, 30.3)):
        assert W.surface_row(m0a, x, z) == W.surface_row(m0b, x, z)
    assert (W.random_views(m0b, 3, seed=9)[2].pos[:] == W.random_views(m0a, 3, seed=9)[2].pos[:])
    back = rlevox.repack(rp.mips, 0)  # and back to the reference layout
    for a, b in zip(sc.mips, back.mips):
        assert np.array_equal(a["slabs"], b["slabs"]) and np.array_equal(a["ptrmap"], b["ptrmap"])
    with pytest.raises(rlevox.RlevoxError):
        rlevox.rle4_save("/tmp/never_written.rle4", rp.mips)


def test_offset_shift_is_validated():
    sc = W.gen("tiny", seed=5)
    rp = rlevox.repack(sc.mips, 2)
    bad = copy_mips(rp.mips)
    bad[0]["shift"] = 0  # offsets read without the shift land inside other columns
    assert rlevox.validate(bad)[0] != 0
    bad = copy_mips(rp.mips)
    bad[0]["shift"] = 5
    assert rlevox.validate(bad)[0] != 0
