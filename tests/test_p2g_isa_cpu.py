"""The 3D P2G kernel adds each cell's sums into its wavefront's LDS node tile with a plain
read-add-write per stencil node (no atomics).  That is exact only while every node's tile
read stays after the previous node's tile write in the compiled code: another lane's node can
be the same one, and a wavefront's LDS operations complete in issue order.  This compiles
the kernels for gfx950 with the library's flags and checks the order in the assembly; the
build (adaptive-gimp_amd/build.py) runs the same check and refuses to link a reordered kernel."""
import importlib.util
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "adaptive-gimp_amd")


def _build_module():
    spec = importlib.util.spec_from_file_location("agimp_build", os.path.join(PKG, "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_p2g_tile_updates_stay_in_order(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    mod = _build_module()
    out = str(tmp_path / "kernels.s")
    cmd = [hipcc, *mod.FLAGS, "-I", mod.INCLUDE, "-x", "hip", "--cuda-device-only", "-S", "-o", out,
           os.path.join(PKG, "csrc", "kernels.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    ok, seq = mod.p2g_tile_order_ok(open(out).read())
    assert ok, seq


def test_order_check_rejects_a_reordered_sequence():
    """The build's check itself: a P2G body whose next tile read moved above a tile write fails."""
    mod = _build_module()
    pairs = "".join(["  ds_read_b128 v[0:3], v1\n  ds_write_b128 v1, v[0:3]\n"] * 27)
    good = "_ZN5agimp5k_p2gILi3EEEvPKNS_9SceneDescEfi:\n  ds_write_b128 v1, v[0:3]\n" + pairs + \
        "  ds_read_b128 v[0:3], v1\n" * 4 + ".Lfunc_end0:\n"
    assert mod.p2g_tile_order_ok(good)[0]
    bad = good.replace("  ds_read_b128 v[0:3], v1\n  ds_write_b128 v1, v[0:3]\n  ds_read_b128 v[0:3], v1\n",
                       "  ds_read_b128 v[0:3], v1\n  ds_read_b128 v[0:3], v1\n  ds_write_b128 v1, v[0:3]\n", 1)
    assert bad != good and not mod.p2g_tile_order_ok(bad)[0]
