"""GPU tests on the workload configurations that the headline scene does not cover, and the
HIP path's boundary conditions, plasticity and static activation against closed forms.

* closed forms (tests/closed_forms.py): one particle next to the sticky or slip wall band or
  the slab collider, the singular-value clamp and J_p, the hardened Kirchhoff stress, and
  hand-derived static octrees whose grid initializer returns coarse levels;
* the 2D static band (200k particles, 4 levels, finest 1/512) at full size: topology bit-exact
  against the oracle, one substep within the single-substep bars, six within the
  multi-substep bars;
* the 3D three-sphere scene (4M particles, 4 levels) at full size: three substeps, then the
  octree rebuilt from the GPU's snapshot bit-exact against the oracle, counters, finiteness
  and mass on canonical nodes; one substep of a 250k-particle draw of the same scene within
  the single-substep bars;
* the 64 batched 2D scenes (levels 2..4): one batched advance, then every scene's octree
  rebuilt from its snapshot bit-exact against the oracle, and scenes covering every level
  count against their own oracle runs.

The 2D workloads run with rho = 1000 where several substeps are compared: as written (rho = 1)
they break the P-wave CFL bound at their time step, so round-off grows exponentially
(tests/test_oracle_pins.py::test_c1_as_specified_is_unstable shows it for the two squares).

It also checks that PyTorch's GPU runtime, initialised first, shares a process with the
library.
"""
import os

import numpy as np
import pytest

from closed_forms import (BC_CASES, STATIC_CASES, bc_single_particle, check_static_case, clamp_closed_form,
                          kirchhoff_hardened, random_rotation, static_case_levels, stress_from_grid)
from test_gpu_parity import (assert_grid_close, assert_particles_close, assert_topology_equal, make_pair,
                             particle_errors)

pytestmark = pytest.mark.gpu


def _single(pkg, dim, L, c, x, v, F, Jp, **prm):
    def init(sim):
        sim.x_p.from_numpy(np.asarray(x, float)[None])
        sim.v_p.from_numpy(np.asarray(v, float)[None])
        sim.F_p.from_numpy(np.asarray(F, float)[None])
        sim.C_p.from_numpy(np.zeros((1, dim, dim)))
        sim.m_p.from_numpy(np.ones(1))
        sim.g_p.from_numpy(np.array([L - 1], np.int32))
        sim.Jp_p.from_numpy(np.array([Jp]))

    return pkg.AdaptiveGIMP(dim, L, c, 1, init, **prm)


# ------------------------------------------------------------------------- closed forms


@pytest.mark.parametrize("L,c", [(1, 16), (3, 4)])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("case", BC_CASES, ids=[cs[0] for cs in BC_CASES])
def test_walls_and_slab_closed_form(pkg, case, dim, L, c):
    _, pos, vel, mode, slab = case
    fs = c * 2 ** (L - 1)
    dx = 1.0 / fs
    x = np.array(pos[:dim]) * dx
    v = np.array(vel[:dim])
    grav = (0.0, -9.8, 0.0)[:dim]
    dt = 1e-3
    prm = dict(E=1e3, gravity=grav, wall_mode=mode, slab=slab)
    sim = _single(pkg, dim, L, c, x, v, np.eye(dim), 1.0, **prm)
    sim.substep(dt)
    xe, ve, Ce = bc_single_particle(x, v, dt, dx, fs, grav, wall_mode=mode, slab=slab)
    assert sim.fl_p.to_numpy()[0] == L - 1
    np.testing.assert_allclose(sim.v_p.to_numpy()[0], ve, rtol=0, atol=2e-6)
    np.testing.assert_allclose(sim.x_p.to_numpy()[0], xe, rtol=0, atol=1e-7)
    np.testing.assert_allclose(sim.C_p.to_numpy()[0], Ce, rtol=0, atol=1e-5)
    if not ve.any():  # the whole stencil in the band or the slab: at rest, position unchanged
        assert not sim.v_p.to_numpy().any()
        assert np.array_equal(sim.x_p.to_numpy()[0], np.float32(x))
    sim.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_clamp_closed_form(pkg, dim):
    rng = np.random.default_rng(21)
    tc, ts = 2.5e-2, 7.5e-3
    for k in range(8):
        s = rng.uniform(0.95, 1.03, dim)
        F = random_rotation(dim, rng) @ np.diag(s) @ random_rotation(dim, rng).T
        F = F.astype(np.float32).astype(np.float64)  # the stored F is float32
        Jp = float(np.float32(rng.uniform(0.9, 1.1)))
        sim = _single(pkg, dim, 1, 16, np.full(dim, 0.5 + 0.013), np.zeros(dim), F, Jp, E=0.0,
                      gravity=(0.0,) * dim, plastic=True, theta_c=tc, theta_s=ts, hardening=10.0)
        sim.substep(1e-4)
        Fe, Je = clamp_closed_form(F, Jp, tc, ts)
        np.testing.assert_allclose(sim.F_p.to_numpy()[0], Fe, rtol=0, atol=2e-7)
        assert sim.Jp_p.to_numpy()[0] == pytest.approx(Je, rel=2e-7)
        sim.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_hardened_stress_closed_form(pkg, dim):
    rng = np.random.default_rng(22)
    c, dt, pv, E, nu, xi = 16, 1e-4, 2e-3, 1e4, 0.3, 10.0
    dx = 1.0 / c
    for Jp in (1.0, 0.97, 1.02):
        F = random_rotation(dim, rng) @ np.diag(rng.uniform(0.99, 1.01, dim)) @ random_rotation(dim, rng).T
        F = F.astype(np.float32).astype(np.float64)
        x = (0.5 + rng.uniform(-0.4, 0.4, dim) * dx).astype(np.float32).astype(np.float64)
        sim = _single(pkg, dim, 1, c, x, np.zeros(dim), F, np.float32(Jp), E=E, nu=nu, gravity=(0.0,) * dim,
                      p_vol=pv, plastic=True, theta_c=0.5, theta_s=0.5, hardening=xi)
        sim.substep(dt)
        m = sim.ad_grid_m[0].to_numpy().astype(np.float64)
        v = sim.ad_grid_v[0].to_numpy().astype(np.float64)
        S = stress_from_grid(m, v, x, dx)
        Se = -float(np.float32(dt)) * pv / dx * kirchhoff_hardened(F, E, nu, float(np.float32(Jp)), xi)
        np.testing.assert_allclose(S, Se, rtol=0, atol=2e-4 * np.abs(Se).max())
        sim.close()


@pytest.mark.parametrize("case", STATIC_CASES, ids=[c["name"] for c in STATIC_CASES])
def test_static_coarse_levels_hand_derived(pkg, case):
    def gi(sim, I):
        lev = static_case_levels(case, I)
        sim.activate_cell(lev, I)
        return lev

    def init(sim):
        sim.x_p.from_numpy(np.array([[0.49, 0.5]]))
        sim.F_p.from_numpy(np.eye(2)[None])
        sim.m_p.from_numpy(np.ones(1))

    sim = pkg.AdaptiveGIMP(2, case["L"], case["c"], 1, init, gi)
    sim.refine()
    sim.substep(1e-5)  # fl_p is written by the substep's level mapping
    check_static_case(case, [sim.cell_mask[l].to_numpy() for l in range(case["L"])], sim.node_mask.to_numpy(),
                      sim.fl_p.to_numpy()[0])
    sim.close()


# ------------------------------------------------------------------------- 2D static band, full size


def _wl_pair(pkg, O, wl):
    return make_pair(pkg, O, wl.dim, wl.level, wl.coarsest_size, wl.state, wl.params, wl.grid_initializer)


def test_static_band_full_size_one_substep(pkg, O):
    wl = pkg.scenes.c2_static_band()
    sim, ora = _wl_pair(pkg, O, wl)
    sim.refine()
    ora.refine()
    assert_topology_equal(sim, ora, O, wl.level)
    sim.substep(wl.dt)
    ora.substep(wl.dt)
    assert_topology_equal(sim, ora, O, wl.level)
    assert_grid_close(sim, ora, wl.level)
    assert_particles_close(particle_errors(sim, ora))
    cnt = sim.counters()
    assert cnt["oob"] == ora.oob() == 0 and cnt["nonfinite"] == 0
    sim.close()


def test_static_band_full_size_six_substeps(pkg, O):
    wl = pkg.scenes.c2_static_band(rho=1000.0)
    sim, ora = _wl_pair(pkg, O, wl)
    for _ in range(6):
        sim.substep(wl.dt)
        ora.substep(wl.dt)
    assert_topology_equal(sim, ora, O, wl.level)
    e = particle_errors(sim, ora)
    assert e["fl"] == 0 and e["x"] <= 1e-5 and e["v"] <= 1e-3 and e["F"] <= 1e-4, e
    sim.close()


# ------------------------------------------------------------------------- 3D three spheres, 4 levels


def _canonical_mass(sim, L, dim):
    nm = sim.node_mask.to_numpy()
    total = 0.0
    for l in range(L):
        m = sim.ad_grid_m[l].to_numpy()
        nz = np.argwhere(m != 0)
        f = nz * 2 ** (L - 1 - l)
        M = nm[tuple(f.T)]
        act = (M & 5) == 1
        canon = act if l == 0 else act & np.any(nz % 2 == 1, axis=1)
        total += m[tuple(nz[canon].T)].astype(np.float64).sum() + m[tuple(nz[M == 0].T)].astype(np.float64).sum()
        del m
    return total


def test_three_spheres_full_size_properties(pkg, O):
    """4M particles, 4 levels: culling over four levels and T/ghost marks at L = 4 on the real
    scene.  After 3 substeps the octree rebuilt from the GPU's snapshot equals the oracle's bit
    for bit; every particle is finite and transferred; mass on canonical + unmarked nodes
    equals the particle mass to float32 accumulation accuracy."""
    wl = pkg.scenes.c5_three_spheres()
    sim = pkg.AdaptiveGIMP(3, wl.level, wl.coarsest_size, wl.n, wl.initializer(), **wl.params)
    for _ in range(3):
        sim.substep(wl.dt)
    cnt = sim.counters()
    assert cnt["oob"] == 0 and cnt["nonfinite"] == 0 and cnt["frozen"] == 0 and cnt["active"] == 3 * wl.n
    x = sim.x_p.to_numpy()
    assert np.isfinite(x).all()
    assert _canonical_mass(sim, wl.level, 3) == pytest.approx(wl.state["m"].sum(), rel=1e-4)
    o2 = O.OracleSim(3, wl.level, wl.coarsest_size, wl.n, fp64=False, **wl.params)
    o2.set_particles(x, sim.v_p.to_numpy(), sim.F_p.to_numpy(), sim.C_p.to_numpy(),
                     np.asarray(wl.state["m"], np.float32), sim.g_p.to_numpy())
    sim.refine()
    o2.refine()
    assert_topology_equal(sim, o2, O, wl.level)
    fl_gpu = sim.fl_p.to_numpy()
    sim.close()
    # the levels used by the last substep's transfers cover all four
    assert set(np.unique(fl_gpu)) == {0, 1, 2, 3}


def test_three_spheres_one_substep(pkg, O):
    """A 250k-particle draw of the 4-level three-sphere scene: one substep within the
    single-substep bars of the oracle (grid and particles), topology bit-exact."""
    wl = pkg.scenes.c5_three_spheres(n=250_000)
    sim, ora = _wl_pair(pkg, O, wl)
    sim.substep(wl.dt)
    ora.substep(wl.dt)
    assert_topology_equal(sim, ora, O, wl.level)
    assert_grid_close(sim, ora, wl.level)
    assert_particles_close(particle_errors(sim, ora))
    sim.close()


# ------------------------------------------------------------------------- 64 batched scenes


def test_batched_scenes_against_oracle(pkg, O):
    """All 64 scenes (levels 2..4, finest 1/256 or 1/512, 5k..200k particles) advanced 3
    substeps as one batch.  Every scene's octree rebuilt from its own snapshot equals the
    oracle's bit for bit; for one scene of each level count and the largest and smallest
    scenes, particles track a single-scene oracle run within the multi-substep bars."""
    wls = pkg.scenes.c4_batch(rho=1000.0)
    sims = [pkg.AdaptiveGIMP(2, w.level, w.coarsest_size, w.n, w.initializer(), **w.params) for w in wls]
    b = pkg.Batch(sims)
    steps = 3
    b.advance([w.dt for w in wls], steps)
    ns = [w.n for w in wls]
    sample = {int(np.argmax(ns)), int(np.argmin(ns))}
    for L in (2, 3, 4):
        sample.add(next(i for i, w in enumerate(wls) if w.level == L))
    for i, (w, s) in enumerate(zip(wls, sims)):
        snap = dict(x=s.x_p.to_numpy(), v=s.v_p.to_numpy(), F=s.F_p.to_numpy(), C=s.C_p.to_numpy(),
                    g=s.g_p.to_numpy())
        assert np.isfinite(snap["x"]).all(), i
        if i in sample:
            ora = O.OracleSim(2, w.level, w.coarsest_size, w.n, fp64=False, **w.params)
            f32 = lambda a: np.asarray(a, np.float32)
            st = w.state
            ora.set_particles(f32(st["x"]), f32(st["v"]), f32(st["F"]), f32(st["C"]), f32(st["m"]), st["g"])
            for _ in range(steps):
                ora.substep(w.dt)
            e = particle_errors(s, ora)
            assert e["fl"] == 0 and e["x"] <= 1e-5 and e["v"] <= 1e-3 and e["F"] <= 1e-4, (i, e)
        o2 = O.OracleSim(2, w.level, w.coarsest_size, w.n, fp64=False, **w.params)
        o2.set_particles(snap["x"], snap["v"], snap["F"], snap["C"], np.asarray(w.state["m"], np.float32), snap["g"])
        s.refine()
        o2.refine()
        assert_topology_equal(s, o2, O, w.level)
    b.close()
    for s in sims:
        s.close()


def test_autotune_batch_picks_fastest_recorded_chunk(pkg):
    """tune_batch on eight of the batched scenes: for every size class the chosen particles per
    chunk is one of 64..512 and the fastest of that class's recorded trials, every member runs
    with its class's choice, and tuning leaves the particles as they were (each trial restores
    the snapshot taken before tuning)."""
    from importlib import import_module

    at = import_module(pkg.__name__ + ".autotune")
    wls = pkg.scenes.c4_batch(n_scenes=8, rho=1000.0)
    sims = [pkg.AdaptiveGIMP(2, w.level, w.coarsest_size, w.n, w.initializer(), **w.params) for w in wls]
    b = pkg.Batch(sims)
    before = [s.x_p.to_numpy() for s in sims]
    ch, log = at.tune_batch(b, [w.dt for w in wls], steps=2)
    classes = at.size_classes([w.n for w in wls])
    assert len(ch) == len(wls)
    for c in set(int(k) for k in classes):
        trials = [t for t in log if t["class"] == c]
        assert sorted(t["chunk"] for t in trials) == [64, 128, 256, 512]
        assert all(t["s_per_substep"] > 0 for t in trials)
        best = min(trials, key=lambda t: t["s_per_substep"])["chunk"]
        assert all(int(ch[i]) == best for i in range(len(wls)) if classes[i] == c), (c, trials, ch)
        assert all(int(s._cfg.particles_per_chunk) > 0 for s in sims)
    for s, x0 in zip(sims, before):
        assert np.array_equal(s.x_p.to_numpy(), x0)
    b.advance([w.dt for w in wls], 2)  # the tuned batch runs
    assert all(np.isfinite(s.x_p.to_numpy()).all() for s in sims)
    b.close()
    for s in sims:
        s.close()


def _chain_scene(pkg, name):
    S = pkg.scenes
    if name == "small3d":
        return S.small_3d(seed=3, n=6000, L=3, c=16)
    if name == "c2_static_small":  # static octree, 4 levels, 2D
        return S.c2_static_band(seed=1, n=20_000, rho=1000.0)
    return S.c1_two_squares(rho=1000.0)  # dynamic, 2D


@pytest.mark.parametrize("name,n", [("small3d", 8), ("small3d", 11), ("small3d", 19), ("c2_static_small", 17),
                                    ("c1", 16)])
def test_chained_replays_equal_single_replays(pkg, O, name, n):
    """advance() replays chains of 8 substeps per graph launch, then single substeps; inside a
    chain the P2G and G2P launches clear the octree for the next rebuild and the T and ghost
    marks share the level map's launch.  The result equals n single eager substeps up to
    float32 summation order, and octrees rebuilt from both snapshots agree bit for bit."""
    wl = _chain_scene(pkg, name)
    d = wl.dim
    a = pkg.AdaptiveGIMP(d, wl.level, wl.coarsest_size, wl.n, wl.initializer(), wl.grid_initializer, **wl.params)
    e = pkg.AdaptiveGIMP(d, wl.level, wl.coarsest_size, wl.n, wl.initializer(), wl.grid_initializer, **wl.params)
    a.advance(wl.dt, n)
    for _ in range(n):
        e.substep(wl.dt)
    xa, xe = a.x_p.to_numpy(), e.x_p.to_numpy()
    assert np.abs(xa - xe).max() <= 1e-6
    assert np.abs(a.v_p.to_numpy() - e.v_p.to_numpy()).max() <= 1e-4 * max(np.abs(e.v_p.to_numpy()).max(), 1e-30)
    assert np.array_equal(a.fl_p.to_numpy(), e.fl_p.to_numpy())
    a.refine()
    e.refine()
    for l in range(wl.level):
        assert np.array_equal(a.cell_mask[l].to_numpy(), e.cell_mask[l].to_numpy()), l
    a.close()
    e.close()


_TORCH_FIRST = r"""
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import torch
a = torch.arange(1 << 16, device="cuda", dtype=torch.float32)
s0 = float((2 * a).sum().item())
import __graft_entry__ as g
pkg = g.package()
wl = pkg.scenes.small_3d(seed=1, n=6000, L=3, c=16)
sim = pkg.AdaptiveGIMP(3, wl.level, wl.coarsest_size, wl.n, wl.initializer(), None, **wl.params)
sim.advance(wl.dt, 3)
x = sim.x_p.to_numpy()
sim.close()
s1 = float((2 * a).sum().item())
assert np.isfinite(x).all() and s0 == s1 == float(2 * sum(range(1 << 16)))
print("torch-then-library ok")
"""


def test_torch_initialised_before_the_library_coexists():
    """PyTorch's GPU runtime and the library share one process if torch initialises first (the
    order bench.py --copy-mb uses); the other order fails in torch's initialisation.  Run in a
    fresh process, since this one has initialised the library already."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", _TORCH_FIRST, root], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "torch-then-library ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
