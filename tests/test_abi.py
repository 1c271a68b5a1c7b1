"""The C ABI: every entry point declared in include/rlevox.h is exported by
librlevox.so and bound by the Python mirror, struct layouts agree with the C
compiler's, and device entry points fail cleanly without a GPU (CPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import rlevox

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rlevox.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rlevox_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 40
    lib = C.CDLL(rlevox.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    unbound = [n for n in names if n not in rlevox.EXPORTED]
    assert not unbound, unbound


def test_struct_layouts_match_c(tmp_path):
    structs = {"rlevox_mip": rlevox.Mip, "rlevox_pose": rlevox.Pose,
               "rlevox_render_params": rlevox.RenderParams, "rlevox_raymap": rlevox.RayMap,
               "rlevox_counters": rlevox.Counters, "rlevox_gen_params": rlevox.GenParams,
               "rlevox_view": rlevox.View, "rlevox_tuning": rlevox.Tuning,
               "rlevox_server_config": rlevox.ServerConfig,
               "rlevox_server_stats": rlevox.ServerStats}
    src = tmp_path / "sizes.c"
    body = "".join('printf("%s %%zu\\n", sizeof(%s));\n' % (k, k) for k in structs)
    body += 'printf("raymap_cos_myaw %zu\\n", offsetof(rlevox_raymap, cos_myaw));\n'
    body += 'printf("raymap_ray_ratio %zu\\n", offsetof(rlevox_raymap, ray_ratio));\n'
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "%s"\nint main(void){\n%s return 0;}\n'
                   % (HEADER, body))
    exe = str(tmp_path / "sizes")
    subprocess.run(["gcc", "-std=c99", str(src), "-o", exe], check=True)
    out = dict(l.split() for l in subprocess.run([exe], stdout=subprocess.PIPE, text=True).stdout.split("\n") if l)
    for k, cls in structs.items():
        assert int(out[k]) == C.sizeof(cls), k
    assert int(out["raymap_cos_myaw"]) == rlevox.RayMap.cos_myaw.offset
    assert int(out["raymap_ray_ratio"]) == rlevox.RayMap.ray_ratio.offset


def test_abi_version():
    assert rlevox.lib.rlevox_abi_version() == 3


def test_no_device_is_reported_not_crashed():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(rlevox.RlevoxError) as e:
        rlevox.Context(0)
    assert e.value.name == "NO_DEVICE"


def test_null_arguments_are_rejected():
    st = rlevox.lib.rlevox_raymap_compute(None, 640, 480, None, None)
    assert st == 1 and "null" in rlevox.last_error()
    st = rlevox.lib.rlevox_render(None, None, None, 640, 480, None, None, None, None)
    assert st == 1


def test_counter_dicts():
    """as_dict() holds the reference's walk, walk_dict() the same keys with
    the kernels' walk values (the fields a walk that ends early changes)."""
    c = rlevox.Counters(*range(1, 13))
    ref = c.as_dict()
    assert list(ref) == list(rlevox.Counters.REFERENCE) and ref["cleared"] == 8
    walk = c.walk_dict()
    assert set(walk) == set(ref)
    assert (walk["dda_steps"], walk["ptr_loads"], walk["rle_loads"], walk["elems"]) == (9, 10, 11, 12)
    assert all(walk[k] == ref[k] for k in ("lines", "spans", "col_loads", "cleared"))
