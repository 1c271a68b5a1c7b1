"""Writes tests/golden/golden.json: SHA-256 digests of seeded scenes, ray
maps and oracle renders (CPU).

The reference ships no fixtures, so these vectors are the oracle's own,
frozen once its pins pass (tests/test_oracle_pins.py).  They catch any later
change of the scene generator, the ray map, the oracle or the device path;
GPU tests compare device renders with the same digests.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "rle-based-voxel-raycasting_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import workloads as W  # noqa: E402

OUT = os.path.join(HERE, "golden.json")


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def raybuf_digest(raybuf, clip):
    """Texels inside each line's clip range (outside them nothing is written)."""
    rows = np.arange(raybuf.shape[1])[None, :]
    inside = (rows >= clip[:, :1]) & (rows <= clip[:, 1:2])
    return digest(np.where(inside, raybuf, 0).astype(np.uint32), clip)


def scene_digest(hs):
    return [digest(m["slabs"], m["ptrmap"]) for m in hs.mips]


def cases():
    """(name, scene name, seed, pose, W, H, mip_distance, far_distance)."""
    c1 = W.gen("c1", seed=1)
    tiny = W.gen("tiny", seed=5)
    out = []
    for i, p in enumerate(W.flythrough(c1.mips[0], frames=5, seed=3)[:3]):
        out.append(("c1_fly%d" % i, "c1", 1, p, 512, 512, 0, 0))
    out.append(("c1_graze", "c1", 1, W.grazing(c1.mips[0], n=1, seed=2)[0], 640, 480, 0, 0))
    out.append(("c1_mipfar", "c1", 1, W.random_views(c1.mips[0], 1, seed=9)[0], 800, 600, 48, 700))
    for i, p in enumerate(W.random_views(tiny.mips[0], 2, seed=4, alt=(4.0, 30.0))):
        out.append(("tiny_%d" % i, "tiny", 5, p, 160, 120, 0, 0))
    return out


def pose_list(p):
    return [float(v) for v in p.pos] + [float(v) for v in p.rot]


def main():
    from oracle import Oracle
    orc = Oracle()
    scenes = {"c1": W.gen("c1", seed=1), "tiny": W.gen("tiny", seed=5)}
    data = {"scenes": {k: {"seed": 1 if k == "c1" else 5, "mips": scene_digest(v)}
                       for k, v in scenes.items()},
            "renders": []}
    for name, sc, seed, p, w, h, mipd, far in cases():
        r = orc.render(scenes[sc].mips, p, w, h, mipd, far)
        data["renders"].append({
            "name": name, "scene": sc, "seed": seed, "pose": pose_list(p), "W": w, "H": h,
            "mip_distance": mipd, "far_distance": far,
            "raymap": hashlib.sha256(r["rm"].to_bytes()).hexdigest(),
            "raybuf": raybuf_digest(r["raybuf"], r["clip"]),
            "compose": digest(r["compose"], r["depth"]),
            "colour": digest(r["colour"]),
            "counters": r["counters"],
        })
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print("wrote", OUT, len(data["renders"]), "renders")


if __name__ == "__main__":
    main()
