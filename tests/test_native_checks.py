"""Native host checks (CPU): the wavefront kernel's chunked DDA against the
sequential walk, the early end of a line's walk against the oracle's full
walk, the batched post passes' tile queues, the gap-fill's divisions by a
reciprocal estimate, and the library's host code plus the oracle under
AddressSanitizer and UndefinedBehaviorSanitizer."""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "rle-based-voxel-raycasting_amd", "csrc")


def build_and_run(name, arg, tmp_path):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(NATIVE, name + ".cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, str(arg)], stdout=subprocess.PIPE, text=True)
    return r.returncode, r.stdout


def test_chunked_dda_matches_sequential_walk(tmp_path):
    rc, out = build_and_run("dda_chunk_check", 3000, tmp_path)
    assert rc == 0, out
    assert "mismatches 0" in out


def _tile_queue_run(tmp_path, defines):
    exe = str(tmp_path / ("tile_queue_check" + "".join(d.replace("=", "_") for d in defines)))
    subprocess.run(["g++", "-O2", "-std=c++17"] + ["-D" + d for d in defines] +
                   [os.path.join(NATIVE, "tile_queue_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    m = re.search(r"tile queue walks (\d+), stripe (\d+), chunk (\d+), failures (\d+)", r.stdout)
    assert m, r.stdout
    return r.returncode, [int(v) for v in m.groups()], r.stdout


@pytest.mark.parametrize("stripe,chunk", [(64, 1), (64, 3), (1024, 4), (1024, 16)])
def test_tile_queues_hand_out_every_tile_once(tmp_path, stripe, chunk):
    """The batched post passes' tile queues, walked by 1-2048 workgroups
    under random interleavings of their requests on batches of 1 to 710400
    tiles: every tile once, every walk ends."""
    rc, (runs, got_stripe, got_chunk, failures), out = _tile_queue_run(
        tmp_path, ["RLEVOX_TILE_STRIPE=%d" % stripe, "RLEVOX_TILE_CHUNK=%d" % chunk])
    assert rc == 0 and failures == 0 and runs > 100, out
    assert (got_stripe, got_chunk) == (stripe, chunk)


def test_tile_queue_check_catches_a_wrong_back_index(tmp_path):
    rc, (_, _, _, failures), _ = _tile_queue_run(tmp_path, ["TILE_QUEUE_PLANT=1"])
    assert rc != 0 and failures > 0


def _walk_end_run(tmp_path, plant):
    exe = str(tmp_path / ("walk_end_check_%d" % plant))
    srcs = [os.path.join(NATIVE, "walk_end_check.cpp")]
    srcs += [os.path.join(CSRC, f) for f in ("host.cpp", "raymap.cpp", "scene.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-pthread",
                    "-Wno-subobject-linkage", "-DWALK_END_PLANT=%d" % plant] + srcs + ["-o", exe],
                   check=True)
    r = subprocess.run([exe, "40"], stdout=subprocess.PIPE, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="4"))
    assert r.returncode == 0, r.stdout
    m = re.search(r"lines ended early (\d+), steps walked (\d+) of (\d+), violations (\d+)",
                  r.stdout)
    assert m, r.stdout
    return [int(v) for v in m.groups()]


def test_walk_end_never_cuts_a_draw(tmp_path):
    """WalkEnd (the early end of a line's walk) evaluated before every step
    of the oracle's full walk on 80 random views: no step after the claimed
    end draws or moves the window, and the walk needs far fewer steps."""
    ended, walked, full, violations = _walk_end_run(tmp_path, 0)
    assert violations == 0
    assert ended > 1000 and walked < 0.8 * full


@pytest.mark.parametrize("plant", [1, 4])
def test_walk_end_check_catches_a_wrong_top(tmp_path, plant):
    """The same check with every run top taken `plant` rows too low (a bound
    that is too optimistic) finds steps that still draw."""
    assert _walk_end_run(tmp_path, plant)[3] > 0


def _recip_div_run(tmp_path, plant):
    exe = str(tmp_path / ("recip_div_check_%d" % plant))
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-DRLEVOX_DIV_PLANT=%d" % plant, os.path.join(NATIVE, "recip_div_check.cpp"),
                    "-o", exe], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    m = re.search(r"recip_div cases (\d+), div_small mismatches (\d+), div_78720 mismatches (\d+)",
                  r.stdout)
    assert m, r.stdout
    return r.returncode, [int(v) for v in m.groups()]


def test_gapfill_reciprocal_divisions_are_exact(tmp_path):
    """The gap-fill's shade average and tap radius from a reciprocal that is
    up to 2 ulps off equal integer division on their whole input ranges
    (and 2 M sampled radii above 2^18)."""
    rc, (cases, small, big) = _recip_div_run(tmp_path, 0)
    assert rc == 0 and small == 0 and big == 0 and cases > 12000000


def test_recip_div_check_catches_planted_mistakes(tmp_path):
    """Without the half-step bias, or without the upward correction, the same
    check finds wrong quotients."""
    assert _recip_div_run(tmp_path, 1)[1][1] > 0
    assert _recip_div_run(tmp_path, 2)[1][2] > 0


def test_host_code_under_sanitizers(tmp_path):
    """Scene generation, validation (with 200 corrupted scenes), pointer-map
    rebuild, re-packing, .rle4 files (with a corrupted one), the dense
    encoder, 2000 random ray maps and oracle renders, built with
    -fsanitize=address,undefined: no report, every check passes."""
    exe = str(tmp_path / "host_sanitize")
    srcs = [os.path.join(NATIVE, "host_sanitize.cpp")]
    srcs += [os.path.join(CSRC, f) for f in ("host.cpp", "raymap.cpp", "scene.cpp")]
    srcs += [os.path.join(ROOT, "oracle", f) for f in ("oracle.cpp", "brute.cpp")]
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-pthread",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer"] + srcs + ["-o", exe], check=True)
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert r.returncode == 0, r.stdout[-4000:]
    assert r.stdout.strip().endswith("host sanitize ok"), r.stdout[-4000:]
