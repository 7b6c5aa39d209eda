"""The 3D P2G kernel adds each cell's sums into its wavefront's LDS node tile with a plain
read-add-write per stencil node (no atomics).  That is exact only while every node's tile
read stays after the previous node's tile write in the compiled code: another lane's node can
be the same one, and a wavefront's LDS operations complete in issue order.  This compiles
the kernels for gfx950 with the library's flags and checks the order in the assembly."""
import importlib.util
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "adaptive-gimp_amd")


def _build_flags():
    spec = importlib.util.spec_from_file_location("agimp_build", os.path.join(PKG, "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return list(mod.FLAGS), mod.INCLUDE


def test_p2g_tile_updates_stay_in_order(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    flags, include = _build_flags()
    out = str(tmp_path / "kernels.s")
    cmd = [hipcc, *flags, "-I", include, "-x", "hip", "--cuda-device-only", "-S", "-o", out,
           os.path.join(PKG, "csrc", "kernels.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    asm = open(out).read()
    names = [n for n in re.findall(r"^(_ZN5agimp\w+):", asm, re.M) if "k_p2gILi3" in n]
    assert len(names) == 1
    start = asm.index(names[0] + ":")
    body = asm[start:asm.index(".Lfunc_end", start)]
    seq = "".join("R" if op == "ds_read_b128" else "W"
                  for op in re.findall(r"^\s+(ds_read_b128|ds_write_b128)\b", body, re.M))
    # the tile clear's write, one read-add-write per stencil node, then the flush reads of
    # the four wavefront tiles
    m = re.fullmatch(r"W+((?:RW)+)(R+)", seq)
    assert m, seq
    assert len(m.group(1)) == 2 * 27 and len(m.group(2)) == 4, seq
