"""The headline workload (1M particles, 3 levels, sphere falling onto the slab) advanced 1000
substeps on the GPU, against the FP64 oracle's run stored in tests/golden/c3_long_horizon.npz
(written by tests/golden/make_c3_long_horizon.py; a seeded subset of 4096 particles at
substeps 250, 500 and 1000).

The bound is set by the noise floor that the fixture also holds: the float32 oracle (the same
algorithm; only the precision and the summation order differ) against the same FP64 run.  The
GPU path sums in another order again (float atomics, gathers), so its distance from the FP64
run must be of the same size as the float32 oracle's, not equal to it.  Over all 1M particles
that floor reaches 1.0e-4 at substep 500 and 2.2e-4 at 1000 (p99 1.7e-5), with 164 particles
on another transfer level: float32 positions near 0.7 round to ~6e-8 at every update, and a
particle whose base cell flips at a cell face follows another stencil from then on.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_long_horizon.npz")


def _errors(a, ref):
    d = np.abs(np.asarray(a, np.float64) - ref)
    return float(d.max()), float(np.quantile(d.reshape(d.shape[0], -1).max(axis=1), 0.99))


def test_headline_scene_1000_substeps_against_fp64_oracle(pkg):
    z = np.load(GOLDEN, allow_pickle=False)
    idx = z["idx"]
    wl = pkg.scenes.c3_sphere_on_slab()
    assert np.array_equal(np.asarray(wl.state["x"], np.float32)[idx], z["x0"])  # the same initial state
    sim = pkg.AdaptiveGIMP(3, wl.level, wl.coarsest_size, wl.n, wl.initializer(), **wl.params)
    done = 0
    report = {}
    for ck in z["checkpoints"]:
        sim.advance(wl.dt, int(ck) - done)
        done = int(ck)
        x = sim.x_p.to_numpy()[idx]
        v = sim.v_p.to_numpy()[idx]
        F = sim.F_p.to_numpy()[idx]
        assert np.isfinite(x).all() and np.isfinite(v).all() and np.isfinite(F).all()
        gx = _errors(x, z[f"fp64_x_{ck}"])
        fx = _errors(z[f"fp32_x_{ck}"], z[f"fp64_x_{ck}"])
        gF = _errors(F, z[f"fp64_F_{ck}"])
        fF = _errors(z[f"fp32_F_{ck}"], z[f"fp64_F_{ck}"])
        report[int(ck)] = dict(gpu_x=gx, floor_x=fx, gpu_F=gF, floor_F=fF)
    cnt = sim.counters()
    sim.close()
    print(report)
    assert cnt["nonfinite"] == 0 and cnt["oob"] == 0
    for ck, r in report.items():
        floor_all = float(z[f"floor_x_max_{ck}"])  # float32 oracle vs FP64 over all 1M particles
        # positions: the bulk tracks the FP64 run as closely as the float32 oracle does, and the
        # worst particle of the subset stays within 1e-4 or twice the float32 oracle's worst over
        # all particles (2.2e-4 at substep 1000: some particles change transfer level)
        assert r["gpu_x"][1] <= 3 * r["floor_x"][1] + 1e-6, (ck, r)
        assert r["gpu_x"][0] <= max(1e-4, 2 * floor_all), (ck, r, floor_all)
        # deformation gradients: within a few times the float32 oracle's (the clamp makes F
        # path-dependent)
        assert r["gpu_F"][0] <= max(1e-4, 5 * r["floor_F"][0]), (ck, r)
