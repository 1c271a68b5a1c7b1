"""Committed golden digests (tests/golden/golden.json) still hold for the
scene generator, the product ray map and the oracle (CPU)."""
import hashlib
import json
import os
import sys

import pytest

import rlevox

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as G  # noqa: E402

GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))


@pytest.fixture(scope="module")
def scenes():
    import workloads as W
    return {"c1": W.gen("c1", seed=1), "tiny": W.gen("tiny", seed=5)}


def test_scene_generator_digests(scenes):
    for name, rec in GOLDEN["scenes"].items():
        assert G.scene_digest(scenes[name]) == rec["mips"], name


@pytest.mark.parametrize("case", GOLDEN["renders"], ids=lambda c: c["name"])
def test_oracle_render_digests(case, scenes, oracle):
    pose = rlevox.Pose.make(case["pose"][:3], case["pose"][3:])
    rm = rlevox.raymap(pose, case["W"], case["H"], case["mip_distance"], case["far_distance"])
    assert hashlib.sha256(rm.to_bytes()).hexdigest() == case["raymap"]
    r = oracle.render(scenes[case["scene"]].mips, pose, case["W"], case["H"],
                      case["mip_distance"], case["far_distance"])
    assert G.raybuf_digest(r["raybuf"], r["clip"]) == case["raybuf"]
    assert G.digest(r["compose"], r["depth"]) == case["compose"]
    assert G.digest(r["colour"]) == case["colour"]
    assert r["counters"] == case["counters"]


def test_pose_round_trip():
    """Poses stored as float lists rebuild the identical fp32 pose."""
    for case in GOLDEN["renders"]:
        p = rlevox.Pose.make(case["pose"][:3], case["pose"][3:])
        assert G.pose_list(p) == case["pose"]
