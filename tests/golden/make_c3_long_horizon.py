"""Writes tests/golden/c3_long_horizon.npz: the 1M-particle 3-level sphere-on-slab workload
advanced 1000 substeps by the oracle (oracle/agimp_oracle.cpp), once in float64 (the error
oracle) and once in float32 (the reference's field type), both from the same float32 initial
state.  A seeded subset of 4096 particles (original indices) is stored at substeps 250, 500
and 1000: x, v, F, J_p and fl_p of both runs.  The float32-vs-float64 difference is the noise
floor that summation order alone produces; the GPU test compares the HIP path at the same
substeps with the float64 run against that floor.

    python tests/golden/make_c3_long_horizon.py [--threads 6] [--steps 1000]

Runs on the CPU in about 15 minutes per precision on 6 cores.  Test infrastructure only.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SUBSET_SEED = 1234
SUBSET_N = 4096
CHECKPOINTS = (250, 500, 1000)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=6)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--out", default=os.path.join(HERE, "c3_long_horizon.npz"))
    a = ap.parse_args()

    import __graft_entry__ as ge
    from oracle import oracle as O

    pkg = ge.package()
    wl = pkg.scenes.c3_sphere_on_slab()
    st = wl.state
    f32 = lambda v: np.asarray(v, np.float32)
    idx = np.sort(np.random.default_rng(SUBSET_SEED).choice(wl.n, SUBSET_N, replace=False))
    ckpts = [c for c in CHECKPOINTS if c <= a.steps]
    out = dict(idx=idx.astype(np.int64), checkpoints=np.asarray(ckpts, np.int64), dt=np.float64(wl.dt),
               x0=f32(st["x"])[idx])
    full = {}
    for name, fp64 in (("fp64", True), ("fp32", False)):
        o = O.OracleSim(3, wl.level, wl.coarsest_size, wl.n, fp64=fp64, threads=a.threads, **wl.params)
        o.set_particles(f32(st["x"]), f32(st["v"]), f32(st["F"]), f32(st["C"]), f32(st["m"]), st["g"])
        t0 = time.time()
        for s in range(1, a.steps + 1):
            o.substep(wl.dt)
            if s in ckpts:
                r = o.get_particles()
                for k in ("x", "v", "F", "Jp"):
                    out[f"{name}_{k}_{s}"] = r[k][idx]
                out[f"{name}_fl_{s}"] = r["fl"][idx].astype(np.int8)
                if fp64:
                    full[s] = (r["x"], r["v"], r["fl"])
                else:  # noise floor over every particle: float32 run against the float64 run
                    x64, v64, fl64 = full[s]
                    dx = np.abs(r["x"] - x64).max(axis=1)
                    out[f"floor_x_max_{s}"] = np.float64(dx.max())
                    out[f"floor_x_p99_{s}"] = np.float64(np.quantile(dx, 0.99))
                    out[f"floor_v_max_{s}"] = np.float64(np.abs(r["v"] - v64).max())
                    out[f"floor_fl_diff_{s}"] = np.int64((r["fl"] != fl64).sum())
                print(f"{name} substep {s}: {time.time() - t0:.0f} s", flush=True)
            elif s % 50 == 0:
                print(f"{name} substep {s}: {time.time() - t0:.0f} s", flush=True)
        del o
    np.savez_compressed(a.out, **out)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
