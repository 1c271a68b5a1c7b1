"""Host check of the wavefront kernel's chunked DDA: lane k holding term k
of each crossing sequence and the steps ordered by merge ranks reproduce
the sequential walk bit for bit (CPU)."""
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


def test_chunked_dda_matches_sequential_walk(tmp_path):
    rc, out = build_and_run("dda_chunk_check", 3000, tmp_path)
    assert rc == 0, out
    assert "mismatches 0" in out
