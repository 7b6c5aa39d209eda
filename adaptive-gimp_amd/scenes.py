"""Seeded scene generators and static-adaptivity level maps.

Particle initializers follow the reference's callback protocol (``init.py:3-32``,
``run_3d.py:14-32``): ``initializer(sim)`` fills ``sim.x_p``, ``sim.v_p``, ``sim.F_p``,
``sim.m_p``, ``sim.g_p``, ``sim.c_p`` before the first substep.  ``ti.random()`` is replaced by
``numpy.random.default_rng(seed)`` so every scene is reproducible.

Grid initializers (static adaptivity) are vectorised versions of ``init.py:34-99``: called
as ``gi(sim, I)`` with ``I`` an int array ``[M, d]`` of finest-cell indices, they call
``sim.activate_cell(level, I_subset)`` and return the level array, exactly like the scalar
Taichi functions do one cell at a time.

The workload configurations (``workload(name)``) add what the reference lacks (seeded
jittered sampling, spheres, plasticity, sticky walls, a slab collider); each choice that the
workload text leaves open is fixed here and stated in its docstring.
"""
from __future__ import annotations

import math

import numpy as np

# ----------------------------------------------------------------------------- reference demo scenes


def reference_case0_2d(case: int = 0, seed: int = 0):
    """``init.py:3-32`` ``initialize_particle(simulator, case)`` as a Python callback."""

    def init(sim):
        n = sim.n_particles
        rng = np.random.default_rng(seed)
        xy = rng.random((n, 2))
        half = n // 2
        x = np.empty((n, 2))
        v = np.empty((n, 2))
        x[:half] = xy[:half] * 0.4 + np.array([0.2, 0.1])
        v[:half] = [0.0, 20.0]
        x[half:] = xy[half:] * 0.4 + np.array([0.4, 0.55])
        v[half:] = [0.0, -20.0]
        g = np.full(n, -1, np.int32)
        c = np.tile(np.array([0.93, 0.33, 0.23]), (n, 1))
        if case == 0:
            inner = np.all((0.2 <= xy) & (xy <= 0.8), axis=1)
            mid = np.all((0.1 <= xy) & (xy <= 0.9), axis=1)
            outer = np.all((0.03 <= xy) & (xy <= 0.97), axis=1)
            g = np.where(inner, 0, np.where(mid, 1, np.where(outer, 2, 3))).astype(np.int32)
            cols = np.array([[0.93, 0.33, 0.23], [0.95, 0.80, 0.43], [0.80, 0.49, 0.80], [0.34, 0.51, 0.44]])
            c = cols[g]
        sim.x_p.from_numpy(x)
        sim.v_p.from_numpy(v)
        sim.F_p.from_numpy(np.tile(np.eye(2), (n, 1, 1)))
        sim.m_p.from_numpy(np.full(n, sim.p_mass))
        sim.g_p.from_numpy(g)
        sim.c_p.from_numpy(c)

    return init


def reference_3d(seed: int = 0):
    """``run_3d.py:14-32`` ``init_p``: two boxes moving towards each other, 3-level tags."""

    def init(sim):
        n = sim.n_particles
        rng = np.random.default_rng(seed)
        xyz = rng.random((n, 3))
        x = np.empty((n, 3))
        v = np.zeros((n, 3))
        first = (np.arange(n) // 25000) == 0
        x[first] = xyz[first] * 0.3 + np.array([0.2, 0.1, 0.2])
        v[first] = [0.0, 20.0, 0.0]
        x[~first] = xyz[~first] * 0.3 + np.array([0.2, 0.5, 0.2])
        v[~first] = [0.0, -20.0, 0.0]
        inner = np.all((0.05 <= xyz) & (xyz <= 0.95), axis=1)
        mid = np.all((0.01 <= xyz) & (xyz <= 0.99), axis=1)
        g = np.where(inner, 0, np.where(mid, 1, 2)).astype(np.int32)
        sim.x_p.from_numpy(x)
        sim.v_p.from_numpy(v)
        sim.F_p.from_numpy(np.tile(np.eye(3), (n, 1, 1)))
        sim.m_p.from_numpy(np.full(n, sim.p_mass))
        sim.g_p.from_numpy(g)

    return init


# ----------------------------------------------------------------------------- static grid initializers


def _all(cond):
    return np.all(cond, axis=-1)


def initialize_mask1(sim, I):
    """``init.py:34-37``: everything at the finest level."""
    lev = np.full(I.shape[0], sim.level - 1, np.int64)
    sim.activate_cell(lev, I)
    return lev


def initialize_mask2(sim, I):
    """``init.py:39-48``."""
    sz = sim.finest_size
    a = (I[:, 0] < sz / 2) & _all((8 <= I) & (I < sz - 8))
    lev = np.where(a, 0, 3)
    sim.activate_cell(lev, I)
    return lev


def initialize_mask3(sim, I):
    """``init.py:50-63``."""
    sz = sim.finest_size
    inside = _all((8 <= I) & (I < sz - 8))
    lev = np.where((I[:, 0] < sz / 4) & inside, 0, np.where((I[:, 0] < sz / 2) & inside, 1, 2))
    sim.activate_cell(lev, I)
    return lev


def initialize_mask4(sim, I):
    """``init.py:65-79`` (cells outside [0, sz) return level 3 without activation)."""
    sz = sim.finest_size
    left = (I[:, 0] < sz / 2) & _all((8 <= I) & (I < sz - 8))
    dom = _all((0 <= I) & (I < sz))
    lev = np.where(left, np.where(I[:, 1] < sz / 2, 1, 2), 3)
    act = left | dom
    sim.activate_cell(lev[act], I[act])
    return lev


def initialize_mask5(sim, I):
    """``init.py:81-98``."""
    sz = sim.finest_size
    left = (I[:, 0] < sz / 2) & _all((8 <= I) & (I < sz - 8))
    dom = _all((0 <= I) & (I < sz))
    lev = np.where(left, np.where(I[:, 1] < sz / 4, 0, np.where(I[:, 1] < sz / 2, 1, 2)), 3)
    act = left | dom
    sim.activate_cell(lev[act], I[act])
    return lev


REFERENCE_GRID_INITIALIZERS = [None, initialize_mask1, initialize_mask2, initialize_mask3, initialize_mask4,
                               initialize_mask5]  # run.py:16


def band_initializer(center=0.3, half_widths=(0.2, 0.1, 0.05), axis=1):
    """Static map refined around the plane ``x[axis] = center``.

    A level-l cell (l < L-1) is split when it is itself split-eligible and its extent along
    ``axis`` intersects ``|y - center| <= half_widths[l]``; every finest cell returns (and
    activates) the level of the unsplit cell containing it.  The rule is hierarchical, so
    the activated cells never overlap (static mode runs no culling).
    """

    def gi(sim, I):
        L, c = sim.level, sim.coarsest_size
        lev = np.zeros(I.shape[0], np.int64)
        going = np.ones(I.shape[0], bool)
        for l in range(L - 1):
            P = I[:, axis] >> (L - 1 - l)
            dx = 1.0 / (c * 2 ** l)
            lo, hi = P * dx, (P + 1) * dx
            w = half_widths[min(l, len(half_widths) - 1)]
            split = going & (hi >= center - w) & (lo <= center + w)
            lev = np.where(split, l + 1, lev)
            going = split
        sim.activate_cell(lev, I)
        return lev

    return gi


# ----------------------------------------------------------------------------- seeded samplers


def jittered_lattice(lo, hi, spacing, rng):
    """One particle per lattice sub-cell of side ``spacing`` inside the box [lo, hi), each at a
    uniform random offset inside its sub-cell."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    counts = np.maximum(np.floor((hi - lo) / spacing + 1e-9).astype(int), 1)
    axes = [lo[a] + spacing * np.arange(counts[a]) for a in range(len(lo))]
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, len(lo))
    return grid + rng.random(grid.shape) * spacing


def sphere_lattice(center, radius, spacing, rng):
    c = np.asarray(center, float)
    pts = jittered_lattice(c - radius, c + radius, spacing, rng)
    return pts[np.sum((pts - c) ** 2, axis=1) <= radius * radius]


def _state(x, v, g, m, dim, color=None):
    n = x.shape[0]
    return dict(x=x.astype(np.float64), v=v.astype(np.float64), F=np.tile(np.eye(dim), (n, 1, 1)),
                C=np.zeros((n, dim, dim)), m=np.full(n, m), g=g.astype(np.int32),
                c=np.tile(np.array([0.93, 0.33, 0.23]), (n, 1)) if color is None else color)


# ----------------------------------------------------------------------------- workload configurations


class Workload:
    """A simulator configuration plus a seeded particle state (original particle order)."""

    def __init__(self, name, dim, level, coarsest_size, state, dt, steps, params, grid_initializer=None,
                 notes=""):
        self.name, self.dim, self.level, self.coarsest_size = name, dim, level, coarsest_size
        self.state, self.dt, self.steps, self.params = state, dt, steps, params
        self.grid_initializer, self.notes = grid_initializer, notes

    @property
    def n(self):
        return self.state["x"].shape[0]

    def initializer(self):
        st = self.state

        def init(sim):
            sim.x_p.from_numpy(st["x"])
            sim.v_p.from_numpy(st["v"])
            sim.F_p.from_numpy(st["F"])
            sim.C_p.from_numpy(st["C"])
            sim.m_p.from_numpy(st["m"])
            sim.g_p.from_numpy(st["g"])
            sim.c_p.from_numpy(st["c"])

        return init


def _box_tags(rel, cuts):
    """Level tag from the relative position in a box, ``init.py:17-29`` style: cuts is a list
    of margins; inside all(cuts[k] <= rel <= 1-cuts[k]) gives level k, outside all gives len."""
    g = np.full(rel.shape[0], len(cuts), np.int32)
    for k in reversed(range(len(cuts))):
        inside = np.all((cuts[k] <= rel) & (rel <= 1 - cuts[k]), axis=1)
        g = np.where(inside, k, g)
    return g


def c1_two_squares(seed=0, scale=1.0, rho=1.0):
    """2D dynamic adaptivity: L=3, coarsest 64 (finest dx 1/256), two squares of side 0.2 at
    (0.3,0.6) and (0.7,0.4), 4 particles per finest cell on a jittered 2x2 sub-lattice
    (~20,972 particles), E=1e4, nu=0.2, rho=1, gravity (0,-9.8), singular-value clamp
    [1-2.5e-2, 1+7.5e-3] without hardening, sticky walls, dt=1e-4, 2000 substeps.
    Level tags: the reference's 4-level rule (init.py:17-29) with its innermost split
    dropped: all(0.1<=rel<=0.9) -> 0, all(0.03<=rel<=0.97) -> 1, else 2."""
    rng = np.random.default_rng(seed)
    L, c = 3, 64
    dxf = 1.0 / (c * 2 ** (L - 1))
    sp = dxf / 2
    xs, gs = [], []
    for cen in [(0.3, 0.6), (0.7, 0.4)]:
        lo = np.array(cen) - 0.1
        p = jittered_lattice(lo, lo + 0.2, sp / math.sqrt(scale), rng)
        xs.append(p)
        gs.append(_box_tags((p - lo) / 0.2, [0.1, 0.03]))
    x = np.concatenate(xs)
    g = np.concatenate(gs)
    ppc = 4.0
    p_vol = dxf ** 2 / ppc
    st = _state(x, np.zeros_like(x), g, rho * p_vol, 2)
    prm = dict(E=1e4, nu=0.2, gravity=(0.0, -9.8), p_vol=p_vol, wall_mode=1, plastic=1, theta_c=2.5e-2,
               theta_s=7.5e-3, hardening=0.0)
    return Workload("c1_2d_dynamic_two_squares", 2, L, c, st, 1e-4, 2000, prm)


def c2_static_band(seed=0, n=200_000, rho=1.0):
    """2D static adaptivity: L=4, coarsest 64 (finest dx 1/512), refinement fixed around
    y=0.3 (finest within |y-0.3|<=0.05, level 2 within 0.1, level 1 within 0.2, coarsest
    elsewhere; hierarchical so cells never overlap); 200k particles uniformly in 5 seeded
    axis-aligned boxes (sides U[0.05,0.25], count proportional to area), one velocity
    U[-1,1]^2 per box; material as c1; dt=5e-5; 5000 substeps."""
    rng = np.random.default_rng(seed)
    L, c = 4, 64
    sides = rng.uniform(0.05, 0.25, (5, 2))
    los = rng.uniform(0.1, 0.9 - sides)
    vel = rng.uniform(-1, 1, (5, 2))
    area = np.prod(sides, axis=1)
    counts = np.floor(n * area / area.sum()).astype(int)
    counts[-1] += n - counts.sum()
    x = np.concatenate([los[b] + rng.random((counts[b], 2)) * sides[b] for b in range(5)])
    v = np.concatenate([np.tile(vel[b], (counts[b], 1)) for b in range(5)])
    p_vol = float(area.sum()) / n
    st = _state(x, v, np.zeros(n, np.int32), rho * p_vol, 2)
    prm = dict(E=1e4, nu=0.2, gravity=(0.0, -9.8), p_vol=p_vol, wall_mode=1, plastic=1, theta_c=2.5e-2,
               theta_s=7.5e-3, hardening=0.0)
    return Workload("c2_2d_static_band", 2, L, c, st, 5e-5, 5000, prm, grid_initializer=band_initializer())


def c3_sphere_on_slab(seed=0, n=1_000_000):
    """3D dynamic adaptivity (the headline workload): L=3, coarsest 64 (finest dx 1/256);
    n particles (default 1,000,000) drawn without replacement (seeded) from a jittered
    lattice of spacing dx/2 inside the sphere r=0.15 at (0.5,0.7,0.5); v=0, F=I.
    Snow: E=1.4e5, nu=0.2, theta_c=2.5e-2, theta_s=7.5e-3, hardening 10, rho=400 (Stomakhin
    et al. 2013's snow density; with rho=1 the P-wave CFL limit dx/c=9.9e-6 is below dt).
    Gravity (0,-9.8,0), sticky walls, and a sticky slab collider occupying 0.45<=y<=0.55
    over the whole x/z extent, whose top face touches the sphere's lowest point at t=0.
    Level tags by radial shell (run_3d.py:27-32 adapted to a sphere): r/R<=0.9 -> 0,
    r/R<=0.98 -> 1, else 2.  dt=2e-5, 1000 substeps."""
    rng = np.random.default_rng(seed)
    L, c = 3, 64
    dxf = 1.0 / (c * 2 ** (L - 1))
    center, R = np.array([0.5, 0.7, 0.5]), 0.15
    pts = sphere_lattice(center, R, dxf / 2, rng)
    if n < pts.shape[0]:
        pts = pts[np.sort(rng.choice(pts.shape[0], n, replace=False))]
    rr = np.sqrt(np.sum((pts - center) ** 2, axis=1)) / R
    g = np.where(rr <= 0.9, 0, np.where(rr <= 0.98, 1, 2)).astype(np.int32)
    rho = 400.0
    p_vol = (4.0 / 3.0) * math.pi * R ** 3 / pts.shape[0]
    st = _state(pts, np.zeros_like(pts), g, rho * p_vol, 3)
    prm = dict(E=1.4e5, nu=0.2, gravity=(0.0, -9.8, 0.0), p_vol=p_vol, wall_mode=1, plastic=1, theta_c=2.5e-2,
               theta_s=7.5e-3, hardening=10.0, slab=(1, 0.45, 0.55))
    return Workload("c3_3d_sphere_1M_3lv", 3, L, c, st, 2e-5, 1000, prm)


def small_3d(seed=0, n=20_000, L=3, c=16):
    """A small 3D scene of the c3 kind for parity tests (sphere r=0.15 at (0.5,0.6,0.5))."""
    rng = np.random.default_rng(seed)
    dxf = 1.0 / (c * 2 ** (L - 1))
    center, R = np.array([0.5, 0.6, 0.5]), 0.15
    pts = sphere_lattice(center, R, dxf / 2, rng)
    if n < pts.shape[0]:
        pts = pts[np.sort(rng.choice(pts.shape[0], n, replace=False))]
    rr = np.sqrt(np.sum((pts - center) ** 2, axis=1)) / R
    g = np.where(rr <= 0.6, 0, np.where(rr <= 0.85, 1, 2)).astype(np.int32)
    g = np.minimum(g, L - 1)
    v = np.tile(np.array([0.0, -1.0, 0.0]), (pts.shape[0], 1))
    p_vol = (4.0 / 3.0) * math.pi * R ** 3 / pts.shape[0]
    st = _state(pts, v, g, 400.0 * p_vol, 3)
    prm = dict(E=1.4e5, nu=0.2, gravity=(0.0, -9.8, 0.0), p_vol=p_vol, wall_mode=1, plastic=1, theta_c=2.5e-2,
               theta_s=7.5e-3, hardening=10.0, slab=(1, 0.3, 0.4))
    return Workload("small_3d_sphere", 3, L, c, st, 2e-5, 100, prm)


def _c4_scene(i, rho):
    """Scene i of the batched 2D workload (seed = scene index)."""
    r = np.random.default_rng(i)
    n = int(round(math.exp(r.uniform(math.log(5e3), math.log(2e5)))))
    L = int(r.integers(2, 5))
    finest = int(r.choice([256, 512]))
    c = finest >> (L - 1)
    nb = int(r.integers(1, 4))
    sides = r.uniform(0.05, 0.25, (nb, 2))
    los = r.uniform(0.1, 0.9 - sides)
    vel = r.uniform(-1, 1, (nb, 2))
    area = np.prod(sides, axis=1)
    counts = np.floor(n * area / area.sum()).astype(int)
    counts[-1] += n - counts.sum()
    cuts = {2: [0.1], 3: [0.1, 0.03], 4: [0.2, 0.1, 0.03]}[L]
    xs, vs, gs = [], [], []
    for b in range(nb):
        rel = r.random((counts[b], 2))
        xs.append(los[b] + rel * sides[b])
        vs.append(np.tile(vel[b], (counts[b], 1)))
        gs.append(_box_tags(rel, cuts))
    x, v, g = np.concatenate(xs), np.concatenate(vs), np.concatenate(gs)
    p_vol = float(area.sum()) / n
    st = _state(x, v, g, rho * p_vol, 2)
    prm = dict(E=1e4, nu=0.2, gravity=(0.0, -9.8), p_vol=p_vol, wall_mode=1, plastic=1, theta_c=2.5e-2,
               theta_s=7.5e-3, hardening=0.0)
    return Workload(f"c4_scene{i}", 2, L, c, st, 5e-5, 1000, prm)


def c4_batch(n_scenes=64, rho=1.0):
    """Batched 2D workload: n_scenes independent scenes, scene i seeded with i: particle count
    log-uniform in [5k, 200k], levels U{2..4}, finest dx in {1/256, 1/512} (coarsest size =
    finest / 2^(L-1)), 1-3 boxes with sides U[0.05,0.25] and one velocity U[-1,1]^2 each,
    level tags by relative position in the box (init.py:17-29 rule cut to L levels), dynamic
    adaptivity, material of the static-band workload (E=1e4, nu=0.2, clamp, sticky walls),
    dt=5e-5, 1000 substeps.  rho=1 as written violates the P-wave CFL bound at that dt."""
    return [_c4_scene(i, rho) for i in range(n_scenes)]


def c5_three_spheres(seed=0, n=4_000_000):
    """3D learned-adaptivity workload (plain physics part): L=4, coarsest 64 (finest 1/512);
    three spheres r=0.1 on a seeded triangle around (0.5,0.5,0.5), each moving towards the
    centroid with a speed U[1,3]; n particles (default 4,000,000) drawn without replacement
    from jittered lattices of spacing dx/2; snow material as the 1M sphere; level tags by
    radial shell r/R <= 0.7 -> 0, <= 0.85 -> 1, <= 0.95 -> 2, else 3; dt=2e-5, 2000 substeps."""
    rng = np.random.default_rng(seed)
    L, c = 4, 64
    dxf = 1.0 / (c * 2 ** (L - 1))
    R = 0.1
    phi0 = rng.uniform(0, 2 * math.pi)
    centre = np.array([0.5, 0.5, 0.5])
    pts, vels, gs = [], [], []
    for k in range(3):
        phi = phi0 + 2 * math.pi * k / 3
        ck = centre + 0.16 * np.array([math.cos(phi), 0.3 * rng.uniform(-1, 1), math.sin(phi)])
        p = sphere_lattice(ck, R, dxf / 2, rng)
        speed = rng.uniform(1, 3)
        d = centre - ck
        v = speed * d / np.linalg.norm(d)
        rr = np.sqrt(np.sum((p - ck) ** 2, axis=1)) / R
        gs.append(np.where(rr <= 0.7, 0, np.where(rr <= 0.85, 1, np.where(rr <= 0.95, 2, 3))).astype(np.int32))
        pts.append(p)
        vels.append(np.tile(v, (p.shape[0], 1)))
    x, v, g = np.concatenate(pts), np.concatenate(vels), np.concatenate(gs)
    if n < x.shape[0]:
        keep = np.sort(rng.choice(x.shape[0], n, replace=False))
        x, v, g = x[keep], v[keep], g[keep]
    p_vol = 3 * (4.0 / 3.0) * math.pi * R ** 3 / x.shape[0]
    st = _state(x, v, g, 400.0 * p_vol, 3)
    prm = dict(E=1.4e5, nu=0.2, gravity=(0.0, -9.8, 0.0), p_vol=p_vol, wall_mode=1, plastic=1, theta_c=2.5e-2,
               theta_s=7.5e-3, hardening=10.0)
    return Workload("c5_3d_three_spheres_4M_4lv", 3, L, c, st, 2e-5, 2000, prm)


def uniform_ball(center, radius, n, rng):
    """n points uniform in a ball (used by the training scenes, which need many cheap draws)."""
    d = len(center)
    u = rng.normal(size=(n, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = radius * rng.random(n) ** (1.0 / d)
    return np.asarray(center) + u * r[:, None]


def training_scene(seed, n=200_000):
    """Scenes of the kinds the learned refinement prediction is used on, for training data
    (seeds 1 and up; the benchmark scenes use their own generators and seed 0):

    * odd seeds: three spheres (L=4, coarsest 64) of radius U[0.07, 0.11] on a triangle around
      (0.5, 0.5, 0.5), each moving to the centroid at U[1, 3] m/s, surfaces U[0.002, 0.05]
      apart, shells r/R <= a0 < a1 < a2 with a0 ~ U[0.6, 0.8] (levels 0..3, finest outside);
    * even seeds: one sphere (L=3, coarsest 64) of radius U[0.08, 0.16], U[0, 0.03] above a
      sticky slab at y in [0.45, 0.55], falling at U[0, 3] m/s, shells r/R <= a0 < a1 (0..2).
    Snow material of the benchmark scenes, rho = 400, n uniform particles."""
    rng = np.random.default_rng(seed)
    prm = dict(E=1.4e5, nu=0.2, gravity=(0.0, -9.8, 0.0), wall_mode=1, plastic=1, theta_c=2.5e-2,
               theta_s=7.5e-3, hardening=10.0)
    if seed % 2 == 1:
        L, c = 4, 64
        R = rng.uniform(0.07, 0.11)
        gap = rng.uniform(0.002, 0.05)
        rc = (2 * R + gap) / math.sqrt(3.0)  # centres on a circle: pairwise distance 2R + gap
        phi0 = rng.uniform(0, 2 * math.pi)
        a0 = rng.uniform(0.6, 0.8)
        cuts = (a0, a0 + (1 - a0) * 0.5, a0 + (1 - a0) * 0.8)
        xs, vs, gs = [], [], []
        for k, nk in enumerate(np.diff(np.linspace(0, n, 4).astype(int))):
            phi = phi0 + 2 * math.pi * k / 3
            ck = np.array([0.5 + rc * math.cos(phi), 0.5 + 0.3 * rc * rng.uniform(-1, 1), 0.5 + rc * math.sin(phi)])
            p = uniform_ball(ck, R, nk, rng)
            d = np.array([0.5, 0.5, 0.5]) - ck
            vs.append(np.tile(rng.uniform(1, 3) * d / np.linalg.norm(d), (nk, 1)))
            rr = np.linalg.norm(p - ck, axis=1) / R
            gs.append(np.searchsorted(np.asarray(cuts), rr).astype(np.int32))
            xs.append(p)
        x, v, g = np.concatenate(xs), np.concatenate(vs), np.concatenate(gs)
        p_vol = 3 * (4.0 / 3.0) * math.pi * R ** 3 / n
    else:
        L, c = 3, 64
        R = rng.uniform(0.08, 0.16)
        ck = np.array([rng.uniform(0.3, 0.7), 0.55 + R + rng.uniform(0.0, 0.03), rng.uniform(0.3, 0.7)])
        x = uniform_ball(ck, R, n, rng)
        v = np.tile(np.array([0.0, -rng.uniform(0, 3), 0.0]), (n, 1))
        a0 = rng.uniform(0.8, 0.92)
        cuts = (a0, a0 + (1 - a0) * 0.8)
        g = np.searchsorted(np.asarray(cuts), np.linalg.norm(x - ck, axis=1) / R).astype(np.int32)
        p_vol = (4.0 / 3.0) * math.pi * R ** 3 / n
        prm["slab"] = (1, 0.45, 0.55)
    st = _state(x, v, g, 400.0 * p_vol, 3)
    prm["p_vol"] = p_vol
    return Workload(f"train_{seed}", 3, L, c, st, 2e-5, 200, prm)


WORKLOADS = {
    "c1": c1_two_squares,
    "c2": c2_static_band,
    "c3": c3_sphere_on_slab,
    "c4": c4_batch,
    "c5": c5_three_spheres,
}


# ----------------------------------------------------------------------------- static map evaluation

PAD = 4  # cells of padding around every level (one 4-cell leaf block on each side)


def extents(level, coarsest_size):
    """Per-level index extents e_l = 4 * align_size(n_l // 4 + 1, 4) (adaptive_gimp.py:36-37)."""
    return [4 * ((((coarsest_size << l) // 4 + 1) + 3) & ~3) for l in range(level)]


class _LevelMapRecorder:
    """Stand-in for the simulator inside a vectorised grid_initializer call: records the
    activate_cell(level, I) calls (adaptive_gimp.py:68-70) per finest cell."""

    def __init__(self, dim, level, coarsest_size):
        self.dim, self.level, self.coarsest_size = dim, level, coarsest_size
        self.finest_size = coarsest_size * 2 ** (level - 1)
        self.coarsest_dx = 1.0 / coarsest_size
        self._act = {}

    def activate_cell(self, l, I):
        I = np.asarray(I).reshape(-1, self.dim)
        l = np.broadcast_to(np.asarray(l), (I.shape[0],))
        for row, lv in zip(map(tuple, I.tolist()), l.tolist()):
            self._act[row] = lv


def static_level_maps(gi, dim, level, coarsest_size):
    """Evaluate a grid initializer over the padded finest-cell box [-PAD, e_{L-1}+PAD)^d.
    Returns (ret_level, act_level) int8 arrays (act_level -1 where no activation)."""
    e = extents(level, coarsest_size)[-1]
    E = e + 2 * PAD
    axes = [np.arange(-PAD, e + PAD)] * dim
    I = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, dim).astype(np.int64)
    rec = _FastRecorder(dim, level, coarsest_size, I)
    ret = np.asarray(gi(rec, I)).astype(np.int64).reshape(-1)
    if ret.shape[0] != I.shape[0]:
        raise ValueError("grid initializer must return one level per cell")
    if np.any((ret < 0) | (ret >= level)):
        raise ValueError("grid initializer returned a level outside [0, level)")
    return ret.astype(np.int8).reshape((E,) * dim), rec.act.reshape((E,) * dim)


class _FastRecorder(_LevelMapRecorder):
    """Vectorised recorder: activate_cell receives a subset of the evaluated index array."""

    def __init__(self, dim, level, coarsest_size, I):
        super().__init__(dim, level, coarsest_size)
        self._I = I
        self._e = extents(level, coarsest_size)[-1]
        self.act = np.full(I.shape[0], -1, np.int8)

    def activate_cell(self, l, I):
        I = np.asarray(I).reshape(-1, self.dim).astype(np.int64)
        l = np.broadcast_to(np.asarray(l), (I.shape[0],)).astype(np.int64)
        if np.any((l < 0) | (l >= self.level)):
            raise ValueError("activate_cell level outside [0, level)")
        E = self._e + 2 * PAD
        J = I + PAD
        ok = np.all((J >= 0) & (J < E), axis=1)
        k = np.zeros(I.shape[0], np.int64)
        for a in range(self.dim):
            k = k * E + J[:, a]
        self.act[k[ok]] = l[ok].astype(np.int8)
