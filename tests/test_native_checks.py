"""Host checks of the closed-form DDA used by the wavefront kernel: the
product's ffadd()/SeqChunk code (csrc/ffadd.h) compiled for the host and
compared with plain sequential fp32 accumulation (CPU)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")


def build_and_run(name, arg, tmp_path):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(NATIVE, name + ".cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, str(arg)], stdout=subprocess.PIPE, text=True)
    return r.returncode, r.stdout


def test_ffadd_matches_sequential_additions(tmp_path):
    rc, out = build_and_run("ffadd_check", 400000, tmp_path)
    assert rc == 0, out
    assert "mismatches 0" in out


def test_chunked_dda_matches_sequential_walk(tmp_path):
    rc, out = build_and_run("dda_chunk_check", 3000, tmp_path)
    assert rc == 0, out
    assert "mismatches 0" in out
    # both the per-chunk closed form and its fallback were exercised
    closed = int(out.split("closed form ")[1].split(",")[0])
    assert closed > 0
