"""Build libagimp.so (HIP kernels for gfx950 + the C-ABI of include/agimp.h) in-tree.

    python adaptive-gimp_amd/build.py          # always recompiles

Each .hip translation unit is compiled with hipcc in parallel, then linked into one
shared library next to this file.  Packed-FP32 instructions are disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``); float atomics use the hardware
``global_atomic_add_f32`` (``-munsafe-fp-atomics``).  The build fails if the device code of
the 3D P2G kernel does not keep its LDS tile updates in program order (p2g_tile_order_ok).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
OUT = os.path.join(HERE, "libagimp.so")
SOURCES = ["kernels.hip", "capi.hip", "learned.hip", "mlp.hip"]
ARCH = os.environ.get("AGIMP_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libagimp.so")


FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
]


def p2g_tile_order_ok(asm: str) -> tuple[bool, str]:
    """The 3D P2G kernel adds each cell's sums into its wavefront's LDS node tile with a plain
    read-add-write per stencil node.  That is exact only while every node's tile read stays
    after the previous node's tile write in the compiled code (another lane's node can be the
    same one; a wavefront's LDS operations complete in issue order).  Checks that the kernel's
    ds_read_b128 / ds_write_b128 sequence is: the tile clear's writes, 27 read-write pairs, then
    only reads (the flush's, if it reads whole nodes).  Returns (ok, the sequence)."""
    import re

    names = [n for n in re.findall(r"^(_ZN5agimp\w+):", asm, re.M) if "k_p2gILi3" in n]
    if len(names) != 1:
        return False, f"expected one 3D P2G kernel, found {names}"
    start = asm.index(names[0] + ":")
    body = asm[start:asm.index(".Lfunc_end", start)]
    seq = "".join("R" if op == "ds_read_b128" else "W"
                  for op in re.findall(r"^\s+(ds_read_b128|ds_write_b128)\b", body, re.M))
    m = re.fullmatch(r"W+((?:RW)+)(R*)", seq)
    ok = bool(m) and len(m.group(1)) == 2 * 27
    return ok, seq


def build(verbose: bool = False) -> str:
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)

    def p2g_order_check():
        # the device assembly of kernels.hip with the same flags as the object that is linked
        asm_path = os.path.join(objdir, "kernels_gfx.s")
        cmd = [cc, *FLAGS, "-I", INCLUDE, "-x", "hip", "--cuda-device-only", "-S", "-o", asm_path,
               os.path.join(CSRC, "kernels.hip")]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc -S failed for kernels.hip:\n{r.stderr}")
        with open(asm_path) as f:
            ok, seq = p2g_tile_order_ok(f.read())
        if not ok:
            raise RuntimeError("the compiler reordered the 3D P2G LDS tile updates (read/write sequence "
                               f"{seq}); the plain read-add-write would race, refusing to build")
        return asm_path

    def compile_one(src):
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [cc, *FLAGS, "-I", INCLUDE, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            sys.stderr.write(r.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES) + 1) as ex:
        order = ex.submit(p2g_order_check)
        objs = list(ex.map(compile_one, SOURCES))
        order.result()
    tmp = OUT + ".tmp"
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
