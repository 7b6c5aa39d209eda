"""Pins of the oracle (oracle/agimp_oracle.cpp) against closed forms, library routines,
invariants and hand-derived cases -- the reference itself (Taichi) cannot run here and
ships no tests or fixtures.  CPU only."""
import numpy as np
import pytest
from scipy.linalg import polar

# ------------------------------------------------------------------------- GIMP weights (utils.py:18-42)


def bspline(t):
    return (np.array([0.5 * (0.5 - t) ** 2, 0.75 - t * t, 0.5 * (0.5 + t) ** 2]),
            np.array([t - 0.5, -2.0 * t, t + 0.5]))


def closed_form_stencil(t):
    d = len(t)
    W = np.zeros((3 ** d, d + 1))
    N = [bspline(ta) for ta in t]
    for idx in range(3 ** d):
        i = [(idx // 3 ** a) % 3 for a in range(d)]
        W[idx, 0] = np.prod([N[a][0][i[a]] for a in range(d)])
        for a in range(d):
            W[idx, 1 + a] = N[a][1][i[a]] * np.prod([N[b][0][i[b]] for b in range(d) if b != a])
    return W


@pytest.mark.parametrize("dim", [2, 3])
def test_weights_equal_quadratic_bspline(O, dim):
    rng = np.random.default_rng(0)
    for _ in range(200):
        t = rng.uniform(-0.5, 0.5, dim)
        np.testing.assert_allclose(O.weight_stencil(dim, t), closed_form_stencil(t), atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_weights_known_answer_at_cell_centre(O, dim):
    W = O.weight_stencil(dim, np.zeros(dim))
    w1 = np.array([1 / 8, 3 / 4, 1 / 8])
    g1 = np.array([-0.5, 0.0, 0.5])
    for idx in range(3 ** dim):
        i = [(idx // 3 ** a) % 3 for a in range(dim)]
        assert W[idx, 0] == pytest.approx(np.prod(w1[i]), abs=1e-15)
        assert W[idx, 1] == pytest.approx(g1[i[0]] * np.prod(w1[i[1:]]), abs=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_weight_invariants(O, dim):
    """sum w = 1, sum grad w = 0, sum w D = 0, sum grad w (x) D = I with D = i - (t+1)."""
    rng = np.random.default_rng(1)
    for _ in range(100):
        t = rng.uniform(-0.5, 0.5, dim)
        W = O.weight_stencil(dim, t)
        D = np.array([[(idx // 3 ** a) % 3 - (t[a] + 1) for a in range(dim)] for idx in range(3 ** dim)])
        assert W[:, 0].sum() == pytest.approx(1.0, abs=1e-14)
        np.testing.assert_allclose(W[:, 1:].sum(0), 0, atol=1e-14)
        np.testing.assert_allclose((W[:, :1] * D).sum(0), 0, atol=1e-14)
        np.testing.assert_allclose(D.T @ W[:, 1:], np.eye(dim), atol=1e-13)


# ------------------------------------------------------------------------- SVD and stress (utils.py:11-16)


@pytest.mark.parametrize("dim", [2, 3])
def test_svd_against_numpy(O, dim):
    rng = np.random.default_rng(2)
    for k in range(1000):
        F = np.eye(dim) + (0.01 if k % 2 else 0.6) * rng.standard_normal((dim, dim))
        if k % 5 == 0:
            F[0] = -F[0]  # inverted
        U, s, V = O.svd(F)
        np.testing.assert_allclose(U @ np.diag(s) @ V.T, F, atol=1e-13)
        assert np.linalg.det(U) == pytest.approx(1.0, abs=1e-12)
        assert np.linalg.det(V) == pytest.approx(1.0, abs=1e-12)
        np.testing.assert_allclose(np.sort(np.abs(s))[::-1], np.linalg.svd(F)[1], atol=1e-13)
        assert np.sign(np.prod(s)) == np.sign(np.linalg.det(F))


@pytest.mark.parametrize("dim", [2, 3])
def test_stress_against_scipy_polar(O, dim):
    mu, la = 1234.5, 678.9
    rng = np.random.default_rng(3)
    for _ in range(300):
        F = np.eye(dim) + 0.2 * rng.standard_normal((dim, dim))
        if np.linalg.det(F) <= 0:
            continue
        R, _ = polar(F)
        J = np.linalg.det(F)
        ref = 2 * mu * (F - R) @ F.T + la * J * (J - 1) * np.eye(dim)
        np.testing.assert_allclose(O.stress(F, mu, la), ref, atol=1e-10, rtol=1e-10)


@pytest.mark.parametrize("dim", [2, 3])
def test_stress_vanishes_for_rotations(O, dim):
    rng = np.random.default_rng(4)
    for _ in range(50):
        Q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
        if np.linalg.det(Q) < 0:
            Q[:, 0] = -Q[:, 0]
        np.testing.assert_allclose(O.stress(Q, 100.0, 50.0), 0, atol=1e-11)
    np.testing.assert_allclose(O.stress(np.eye(dim), 100.0, 50.0), 0, atol=1e-14)


def test_stress_uniaxial_closed_form(O):
    mu, la, s = 10.0, 5.0, 1.3
    F = np.diag([s, 1.0, 1.0])
    tau = O.stress(F, mu, la)
    J = s
    np.testing.assert_allclose(np.diag(tau), [2 * mu * (s - 1) * s + la * J * (J - 1), la * J * (J - 1),
                                              la * J * (J - 1)], rtol=1e-14)


# ------------------------------------------------------------------------- L=1: uniform MPM


def uniform_mpm_step(x, v, F, C, m, c, dt, mu, la, gravity, gap, p_vol):
    """Textbook quadratic-B-spline MPM step with the affine term C (cell units), written from
    the closed-form weights (independent of the oracle's GIMP overlap formula)."""
    n, d = x.shape
    dx = 1.0 / c
    q = x / dx
    b = np.floor(q + 0.5).astype(int) - 1
    t = q - (b + 1)
    N = 0.5 * (0.5 - t) ** 2, 0.75 - t * t, 0.5 * (0.5 + t) ** 2
    dN = t - 0.5, -2 * t, t + 0.5
    N, dN = np.stack(N, -1), np.stack(dN, -1)  # [n, d, 3]
    tau = np.zeros((n, d, d))
    for p in range(n):
        R, _ = polar(F[p])
        J = np.linalg.det(F[p])
        tau[p] = 2 * mu * (F[p] - R) @ F[p].T + la * J * (J - 1) * np.eye(d)
    S = -dt / dx * p_vol * tau
    size = c + 8
    gm = np.zeros((size,) * d)
    gp = np.zeros((size,) * d + (d,))
    offs = np.stack(np.meshgrid(*[np.arange(3)] * d, indexing="ij"), -1).reshape(-1, d)
    rows = np.arange(n)
    for o in offs:
        w = np.prod(N[rows[:, None], np.arange(d)[None, :], o[None, :]], axis=1)
        gw = np.stack([dN[rows, a, o[a]] * np.prod([N[rows, bb, o[bb]] for bb in range(d) if bb != a], axis=0)
                       for a in range(d)], -1)
        Dl = o[None, :] - (t + 1)
        mom = (w * m)[:, None] * (v + np.einsum("pij,pj->pi", C, Dl)) + np.einsum("pij,pj->pi", S, gw)
        idx = tuple((b + o[None, :] + 4).T)  # 4 cells of padding: base cells can be -1
        np.add.at(gm, idx, w * m)
        np.add.at(gp, idx, mom)
    # the grid update touches nodes inside [0, c] only (adaptive_gimp.py:281); a node at -1
    # keeps its momentum, which G2P then reads as a velocity
    coords = np.stack(np.meshgrid(*[np.arange(size)] * d, indexing="ij"), -1) - 4
    inside = np.all((coords >= 0) & (coords <= c), axis=-1)
    gv = gp.copy()
    has = (gm > 0) & inside
    gv[has] = gp[has] / gm[has][:, None] + np.asarray(gravity) * dt
    wall = (coords < gap) | (coords > c - gap)  # per axis: zero that component
    gv[wall & inside[..., None]] = 0.0
    nv = np.zeros((n, d))
    G = np.zeros((n, d, d))
    for o in offs:
        w = np.prod(N[rows[:, None], np.arange(d)[None, :], o[None, :]], axis=1)
        gw = np.stack([dN[rows, a, o[a]] * np.prod([N[rows, bb, o[bb]] for bb in range(d) if bb != a], axis=0)
                       for a in range(d)], -1)
        vi = gv[tuple((b + o[None, :] + 4).T)]
        nv += w[:, None] * vi
        G += vi[:, :, None] * gw[:, None, :]
    Fn = np.einsum("pij,pjk->pik", np.eye(d)[None] + dt / dx * G, F)
    return x + dt * nv, nv, Fn, G


@pytest.mark.parametrize("dim", [2, 3])
def test_single_level_equals_uniform_mpm(O, dim):
    """level=1 reduces the adaptive method to uniform quadratic MPM (every stencil node is a
    corner of an activated cell, no T/ghost nodes)."""
    rng = np.random.default_rng(5)
    c, n = 16, 400 if dim == 2 else 300
    x = 0.5 + 0.2 * (rng.random((n, dim)) - 0.5)
    x[: n // 10, 0] = rng.uniform(0.02, 0.2, n // 10)  # some particles inside the wall band
    v = rng.standard_normal((n, dim))
    F = np.eye(dim) + 0.05 * rng.standard_normal((n, dim, dim))
    C = 0.3 * rng.standard_normal((n, dim, dim))
    m = rng.uniform(0.5, 1.5, n)
    E, nu = 5e3, 0.2
    mu, la = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    dt = 2.0 ** -14
    g = [0.0, -100.0, 0.0][:dim]
    ora = O.OracleSim(dim, 1, c, n, fp64=True, E=E, nu=nu, gravity=g)
    ora.set_particles(x, v, F, C, m, np.zeros(n, np.int32))
    ora.substep(dt)
    got = ora.get_particles()
    ex, ev, eF, eC = uniform_mpm_step(x, v, F, C, m, c, dt, mu, la, g, 4, 1.0)
    assert (got["fl"] == 0).all()
    np.testing.assert_allclose(got["v"], ev, atol=1e-10, rtol=1e-10)
    np.testing.assert_allclose(got["C"], eC, atol=1e-9, rtol=1e-10)
    np.testing.assert_allclose(got["x"], ex, atol=1e-14)
    np.testing.assert_allclose(got["F"], eF, atol=1e-12)


# ------------------------------------------------------------------------- topology


def _run_refine(O, dim, L, c, x, g, fp64=True):
    n = x.shape[0]
    ora = O.OracleSim(dim, L, c, n, fp64=fp64)
    ora.set_particles(x, np.zeros_like(x), np.tile(np.eye(dim), (n, 1, 1)), np.zeros((n, dim, dim)), np.ones(n), g)
    ora.refine()
    cells = [ora.get_level(l)[0] for l in range(L)]
    return ora, cells, ora.get_node_mask()


def test_hand_derived_ghost_and_T_2d(O):
    """2D, L=2, c=4 (n_0=4, n_1=8).  Particle p1 (g=1) at x=(0.375,0.375): base cell (2,2) at
    level 1, so cells (2..3, 2..3) of level 1 are activated.  Particle p2 (g=0) at
    x=(0.75,0.5): base (2,1) at level 0, cells (2..3, 1..2) of level 0.  No cell has an
    activated ancestor, so culling changes nothing.  Level-1 cell (3,2) has its +x
    neighbour (4,2) covered by level-0 cell (2,1): the shared face holds level-1 nodes
    (4,2) (on the level-0 lattice: not T) and (4,3) (odd: T); the ghost pass marks cell
    (4,2) of level 1 GHOST and its corner nodes (5,2), (5,3) (unmarked so far) GHOST."""
    x = np.array([[0.375, 0.375], [0.75, 0.5]])
    g = np.array([1, 0], np.int32)
    ora, cells, nm = _run_refine(O, 2, 2, 4, x, g)
    c0, c1 = cells
    assert sorted(map(tuple, np.argwhere(c0 == 1))) == [(2, 1), (2, 2), (3, 1), (3, 2)]
    assert sorted(map(tuple, np.argwhere(c1 == 1))) == [(2, 2), (2, 3), (3, 2), (3, 3)]
    A, G, T = 1, 2, 4
    assert nm[4, 2] & A and not nm[4, 2] & T
    assert nm[4, 3] & T and nm[4, 3] & A
    assert nm[5, 2] == G and nm[5, 3] == G
    assert c1[4, 2] == G and c1[4, 3] == G
    # the -y face of (2,2)/(3,2) at level 1 borders level-1 cells (2,1),(3,1) inside level-0 cell
    # (1,0), which is not activated: l_p = -1, nothing marked there
    assert c1[2, 1] == 0 and c1[3, 1] == 0
    assert nm[2, 1] == 0 and nm[3, 1] == 0
    # level-0 corners map to finest nodes 2*I
    for I in [(2, 1), (2, 2), (3, 1), (3, 2)]:
        for dI in [(0, 0), (0, 1), (1, 0), (1, 1)]:
            f = (2 * (I[0] + dI[0]), 2 * (I[1] + dI[1]))
            assert nm[f] & A


def _cover(cells, L):
    """Paint each finest cell with the level of the activated cell covering it."""
    e = cells[-1].shape[0]
    dim = cells[-1].ndim
    cover = np.full((e,) * dim, -1)
    overlaps = 0
    for l in range(L):
        s = 2 ** (L - 1 - l)
        for I in np.argwhere(cells[l] == 1):
            sl = tuple(slice(I[a] * s, min((I[a] + 1) * s, e)) for a in range(dim))
            overlaps += int(np.count_nonzero(cover[sl] >= 0))
            cover[sl] = l
    return cover, overlaps


@pytest.mark.parametrize("dim,L,c,seed", [(2, 3, 4, 0), (2, 4, 4, 1), (2, 3, 8, 2), (3, 3, 4, 3), (3, 2, 8, 4)])
def test_topology_properties(O, dim, L, c, seed):
    rng = np.random.default_rng(seed)
    n = 60 if dim == 2 else 40
    x = rng.uniform(0.1, 0.9, (n, dim))
    g = rng.integers(0, L, n).astype(np.int32)
    ora, cells, nm = _run_refine(O, dim, L, c, x, g)
    # after culling every stored value is 0, ACTIVATED or GHOST
    for l in range(L):
        assert set(np.unique(cells[l])) <= {0, 1, 2}
    cover, overlaps = _cover(cells, L)
    assert overlaps == 0, "activated cells overlap"
    # every requested cell (g_p, base + dI) is covered by activated cells of level >= g_p
    for p in range(n):
        gp = g[p]
        dx = 1.0 / (c * 2 ** gp)
        b = np.floor(x[p] / dx + 0.5).astype(int) - 1
        s = 2 ** (L - 1 - gp)
        for dI in np.ndindex(*(2,) * dim):
            I = b + np.array(dI)
            sl = tuple(slice(I[a] * s, (I[a] + 1) * s) for a in range(dim))
            assert (cover[sl] >= gp).all()
    # node predicates (adaptive_gimp.py:130-140) are mutually exclusive
    act = (nm & 5) == 1
    tj = (nm & 4) == 4
    gh = (nm & 7) == 2
    assert not (act & tj).any() and not (act & gh).any() and not (tj & gh).any()
    # T-junction nodes are corners of activated cells; ghost nodes never are
    assert ((nm[tj] & 1) == 1).all()
    cn = O.canonical_node_mask(nm)
    assert ((cn[(cn & 2) == 2] & 1) == 0).all()
    # fl_p: every particle maps to the finest level at which one of its cells is activated
    got = ora.get_particles()["fl"]
    for p in range(n):
        want = -1
        for l in range(L):
            dx = 1.0 / (c * 2 ** l)
            b = np.floor(x[p] / dx + 0.5).astype(int) - 1
            if any(cells[l][tuple(b + np.array(dI))] == 1 for dI in np.ndindex(*(2,) * dim)):
                want = l
        assert got[p] == want


def test_static_mask1_is_single_finest_level(O, pkg):
    """init.py:34-37 activates every cell at the finest level: no coarse cells, no T/ghost."""
    dim, L, c, n = 2, 3, 8, 200
    rng = np.random.default_rng(6)
    x = rng.uniform(0.2, 0.8, (n, dim))
    ora = O.OracleSim(dim, L, c, n, fp64=True)
    ora.set_particles(x, np.zeros_like(x), np.tile(np.eye(dim), (n, 1, 1)), np.zeros((n, dim, dim)), np.ones(n),
                      np.zeros(n, np.int32))
    ret, act = pkg.scenes.static_level_maps(pkg.scenes.initialize_mask1, dim, L, c)
    ora.set_static_map(ret, act)
    ora.refine()
    for l in range(L - 1):
        assert not ora.get_level(l)[0].any()
    nm = ora.get_node_mask()
    assert set(np.unique(nm)) <= {0, 1}
    assert (ora.get_particles()["fl"] == L - 1).all()


# ------------------------------------------------------------------------- conservation / uniform fields


def test_mass_conservation_on_canonical_nodes(O, pkg):
    """After restriction, summing m over nodes that are activated and (level 0 or not
    all-even) counts every particle's mass exactly once (even activated, T and ghost nodes
    have pushed their mass to the coarser level); the rest sits on stencil nodes whose
    node_mask is 0, which no grid pass touches (the reference's unmarked-node quirk)."""
    wl = pkg.scenes.small_3d(seed=2, n=4000, L=3, c=16)
    st = wl.state
    ora = O.OracleSim(3, wl.level, wl.coarsest_size, wl.n, fp64=True, **wl.params)
    ora.set_particles(st["x"], st["v"], st["F"], st["C"], st["m"], st["g"])
    ora.substep(wl.dt)
    nm = ora.get_node_mask()
    total = 0.0
    unmarked = 0.0
    for l in range(wl.level):
        _, m, _ = ora.get_level(l)
        s = 2 ** (wl.level - 1 - l)
        e = m.shape[0]
        idx = np.arange(e)
        I = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), -1)
        f = np.minimum(I * s, nm.shape[0] - 1)
        M = nm[f[..., 0], f[..., 1], f[..., 2]]
        inside = np.all(I * s < nm.shape[0], axis=-1)
        act = ((M & 5) == 1) & inside
        canon = act if l == 0 else act & np.any(I % 2 == 1, axis=-1)
        total += m[canon].sum()
        unmarked += m[(M == 0) & inside].sum()
    # mass stranded on unmarked stencil nodes (never restricted) is the only leak
    assert unmarked > 0
    assert total + unmarked == pytest.approx(st["m"].sum(), rel=1e-12)


def test_uniform_velocity_free_fall_multilevel(O, pkg):
    """No stress (E=0), C=0, uniform v0: restriction/prolongation weights keep the grid
    velocity uniform, so every transferred particle gets exactly v0 + g dt."""
    rng = np.random.default_rng(7)
    dim, L, c, n = 2, 3, 8, 3000
    x = rng.uniform(0.3, 0.7, (n, dim))
    g = rng.integers(0, L, n).astype(np.int32)
    v0 = np.array([0.3, -0.2])
    dt = 2.0 ** -12
    ora = O.OracleSim(dim, L, c, n, fp64=True, E=0.0, nu=0.0, gravity=[0.0, -10.0])
    ora.set_particles(x, np.tile(v0, (n, 1)), np.tile(np.eye(dim), (n, 1, 1)), np.zeros((n, dim, dim)),
                      np.ones(n), g)
    ora.substep(dt)
    got = ora.get_particles()
    ok = got["fl"] >= 0
    assert ok.all()
    np.testing.assert_allclose(got["v"][ok], np.tile(v0 + np.array([0.0, -10.0]) * dt, (ok.sum(), 1)), atol=1e-12)
    np.testing.assert_allclose(got["C"][ok], 0, atol=1e-11)


def test_c1_as_specified_is_unstable(O, pkg):
    """The 2D two-squares workload with rho=1, E=1e4, dt=1e-4 has a P-wave speed of ~105, so
    dt exceeds dx/c = 3.7e-5: the oracle's own FP32 and FP64 runs agree to 1e-9 after one
    substep and then separate exponentially.  Multi-substep parity is meaningless there."""
    wl = pkg.scenes.c1_two_squares()
    st = wl.state

    def run(fp64, steps):
        o = O.OracleSim(2, wl.level, wl.coarsest_size, wl.n, fp64=fp64, **wl.params)
        o.set_particles(st["x"].astype(np.float32), st["v"], st["F"], st["C"], st["m"].astype(np.float32), st["g"])
        for _ in range(steps):
            o.substep(wl.dt)
        return o.get_particles()["v"]

    one = np.abs(run(False, 1) - run(True, 1)).max()
    ten32, ten64 = run(False, 10), run(True, 10)
    assert one < 1e-8
    assert np.abs(ten32 - ten64).max() > 0.1 * np.abs(ten64).max()
    assert np.abs(ten64).max() > 100 * np.abs(st["v"]).max() + 1.0  # blow-up from rest


# ------------------------------------------------------------------------- boundary conditions, plasticity
# (extensions the workloads need: sticky walls, the slab collider, the singular-value clamp and
# hardening; the reference has slip walls only, adaptive_gimp.py:284-287, and no plasticity)

from closed_forms import (BC_CASES, STATIC_CASES, bc_single_particle, check_static_case,  # noqa: E402
                          clamp_closed_form, kirchhoff_hardened, random_rotation, static_case_levels,
                          stress_from_grid)


def _single_particle_oracle(O, dim, L, c, x, v, F, Jp, dt, **prm):
    ora = O.OracleSim(dim, L, c, 1, fp64=True, **prm)
    ora.set_particles(np.asarray(x, float)[None], np.asarray(v, float)[None], np.asarray(F, float)[None],
                      np.zeros((1, dim, dim)), np.ones(1), np.array([L - 1], np.int32), Jp=np.array([Jp]))
    ora.substep(dt)
    return ora


@pytest.mark.parametrize("L,c", [(1, 16), (3, 4)])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("case", BC_CASES, ids=[cs[0] for cs in BC_CASES])
def test_walls_and_slab_closed_form(O, case, dim, L, c):
    """One finest-level particle next to a wall band or the slab: the oracle's v', x', C' equal
    the closed form (nodes in the band or slab hold zero velocity, slip zeroes only the normal
    component).  A slab applied on another axis, open at its ends, or sticky walls that keep the
    tangential velocity all fail it; a particle whose stencil lies wholly inside ends at rest."""
    _, pos, vel, mode, slab = case
    fs = c * 2 ** (L - 1)
    dx = 1.0 / fs
    x = np.array(pos[:dim]) * dx
    v = np.array(vel[:dim])
    grav = (0.0, -9.8, 0.0)[:dim]
    dt = 1e-3
    prm = dict(E=1e3, gravity=grav, wall_mode=1 if mode == "sticky" else 0)
    if slab is not None:
        prm["slab"] = slab
    ora = _single_particle_oracle(O, dim, L, c, x, v, np.eye(dim), 1.0, dt, **prm)
    r = ora.get_particles()
    xe, ve, Ce = bc_single_particle(x, v, dt, dx, fs, grav, wall_mode=mode, slab=slab)
    assert r["fl"][0] == L - 1
    np.testing.assert_allclose(r["v"][0], ve, rtol=0, atol=1e-13)
    np.testing.assert_allclose(r["x"][0], xe, rtol=0, atol=1e-15)
    np.testing.assert_allclose(r["C"][0], Ce, rtol=0, atol=1e-12)
    if not ve.any():
        assert np.array_equal(r["x"][0], x)


@pytest.mark.parametrize("dim", [2, 3])
def test_clamp_closed_form(O, dim):
    """The singular-value clamp with J_p bookkeeping (Stomakhin et al. 2013) in the oracle's
    G2P: with E = 0, v = 0 and no gravity the trial F is the stored F exactly."""
    rng = np.random.default_rng(21)
    tc, ts = 2.5e-2, 7.5e-3
    for k in range(12):
        s = rng.uniform(0.95, 1.03, dim)
        if k == 0:
            s[:] = 1.0 + 0.5 * ts  # inside the band: unchanged
        F = random_rotation(dim, rng) @ np.diag(s) @ random_rotation(dim, rng).T
        Jp = rng.uniform(0.9, 1.1)
        ora = _single_particle_oracle(O, dim, 1, 16, np.full(dim, 0.5 + 0.013), np.zeros(dim), F, Jp, 1e-4,
                                      E=0.0, gravity=(0.0,) * dim, plastic=1, theta_c=tc, theta_s=ts, hardening=10.0)
        r = ora.get_particles()
        Fe, Je = clamp_closed_form(F, Jp, tc, ts)
        np.testing.assert_allclose(r["F"][0], Fe, rtol=0, atol=1e-13)
        assert r["Jp"][0] == pytest.approx(Je, rel=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_hardened_stress_closed_form(O, dim):
    """P2G's stress term with hardening: S recovered from the node momenta by the B-spline
    moment identity equals -dt V_p/dx * tau(F) with mu, lambda scaled by exp(xi (1 - J_p))."""
    rng = np.random.default_rng(22)
    c, dt, pv, E, nu, xi = 16, 1e-4, 2e-3, 1e4, 0.3, 10.0
    dx = 1.0 / c
    for Jp in (1.0, 0.97, 1.02):
        F = random_rotation(dim, rng) @ np.diag(rng.uniform(0.99, 1.01, dim)) @ random_rotation(dim, rng).T
        x = 0.5 + rng.uniform(-0.4, 0.4, dim) * dx
        ora = _single_particle_oracle(O, dim, 1, c, x, np.zeros(dim), F, Jp, dt, E=E, nu=nu, gravity=(0.0,) * dim,
                                      p_vol=pv, plastic=1, theta_c=0.5, theta_s=0.5, hardening=xi)
        _, m, v = ora.get_level(0)
        S = stress_from_grid(m, v, x, dx)
        Se = -float(np.float32(dt)) * pv / dx * kirchhoff_hardened(F, E, nu, Jp, xi)
        np.testing.assert_allclose(S, Se, rtol=1e-9, atol=1e-12 * np.abs(Se).max())


@pytest.mark.parametrize("case", STATIC_CASES, ids=[c["name"] for c in STATIC_CASES])
def test_static_coarse_levels_against_hand_derived_masks(O, case):
    """The same hand-derived cases through the oracle (expectations in tests/closed_forms.py)."""
    L, c = case["L"], case["c"]
    x = np.array([[0.49, 0.5]])
    ora = O.OracleSim(2, L, c, 1, fp64=True)
    ora.set_particles(x, np.zeros((1, 2)), np.eye(2)[None], np.zeros((1, 2, 2)), np.ones(1), np.zeros(1, np.int32))
    ret, act = _static_maps(2, L, c, lambda I: static_case_levels(case, I))
    ora.set_static_map(ret, act)
    ora.refine()
    check_static_case(case, [ora.get_level(l)[0] for l in range(L)], ora.get_node_mask(), ora.get_particles()["fl"][0])


def _static_maps(dim, L, c, level_of):
    """ret/act level arrays over the padded finest box for a map that activates, in every
    finest cell I, the cell of level level_of(I) containing it (the init.py mask2/mask3
    protocol: activate_cell(l, I) then return l, adaptive_gimp.py:68-70)."""
    e = O_extents(L, c)[-1]
    E = e + 8
    I = np.stack(np.meshgrid(*[np.arange(-4, e + 4)] * dim, indexing="ij"), -1)
    lev = level_of(I).astype(np.int8)
    return lev, lev.copy()


def O_extents(L, c):
    return [4 * ((((c << l) // 4 + 1) + 3) & ~3) for l in range(L)]


def test_static_coarse_levels_hand_derived(O):
    """2D, L=2, c=4 (finest 8): the map returns level 0 left of finest x-index 4, level 1
    elsewhere.  One particle at (0.49, 0.5): finest base cell f = (3, 3).
    1. The 16 calls on f + {-1..2}^2 activate level-0 cells (1,1), (1,2) (I // 2 for
       I_x in {2,3}) and level-1 cells (4..5, 2..5) (I_x in {4,5}).
    2. GI(f) = 0, so the base cell at level 0 is (1,1); the calls on ((1..2, 1..2) * 2)
       activate (1,1), (1,2) at level 0 and (4,2), (4,4) at level 1: nothing new.
    The level-1 cells (4, 2..5) face level-0 cells (1,1), (1,2) across x = 4: their -x
    neighbours (3, 2..5) become GHOST, and the face nodes (4,3), (4,5) (odd, off the level-0
    lattice) T_JUNCTION.  fl_p = 1: the particle's level-1 cells (3..4, 3..4) hold the activated
    (4,3) and (4,4)."""
    x = np.array([[0.49, 0.5]])
    ora = O.OracleSim(2, 2, 4, 1, fp64=True)
    ora.set_particles(x, np.zeros((1, 2)), np.eye(2)[None], np.zeros((1, 2, 2)), np.ones(1), np.zeros(1, np.int32))
    ret, act = _static_maps(2, 2, 4, lambda I: np.where(I[..., 0] < 4, 0, 1))
    ora.set_static_map(ret, act)
    ora.refine()
    c0, c1 = ora.get_level(0)[0], ora.get_level(1)[0]
    A, G, T = 1, 2, 4
    assert sorted(map(tuple, np.argwhere(c0 == A))) == [(1, 1), (1, 2)]
    assert not (c0 & ~A).any()
    assert sorted(map(tuple, np.argwhere(c1 == A))) == [(i, j) for i in (4, 5) for j in (2, 3, 4, 5)]
    assert sorted(map(tuple, np.argwhere(c1 == G))) == [(3, j) for j in (2, 3, 4, 5)]
    assert set(np.unique(c1)) == {0, A, G}
    nm = ora.get_node_mask()
    assert sorted(map(tuple, np.argwhere(nm & T))) == [(4, 3), (4, 5)]
    assert ora.get_particles()["fl"][0] == 1


def test_static_three_levels_hand_derived(O):
    """2D, L=3, c=2 (finest 8): the map returns level 0 left of finest x-index 4, level 2
    elsewhere (a shift of 2 between them).  One particle at (0.49, 0.5): f = (3, 3).
    1. Calls on f + {-1..2}^2: I_x in {2,3} -> level-0 cells I // 4 = (0,0), (0,1);
       I_x in {4,5} -> level-2 cells (4..5, 2..5).
    2. GI(f) = 0; the level-0 base cell is (0,0); calls on (0..1, 0..1) * 4 add level-0
       (0,0), (0,1) and level-2 (4,0), (4,4).
    Ghosts: each level-2 cell (4,j), j in {0,2,3,4,5}, sees level-0 cell (0, j // 4) across
    x = 4 (l_p = 0), so level-2 cells (3,j) and their level-1 parents (1, j // 2) =
    (1,0), (1,1), (1,2) become GHOST.  T: face nodes (4, y) with y not a multiple of 4 on those
    faces: y in {1, 2, 3, 5, 6}.  fl_p = 2 (level-2 cells (4,3), (4,4) are among its cells)."""
    x = np.array([[0.49, 0.5]])
    ora = O.OracleSim(2, 3, 2, 1, fp64=True)
    ora.set_particles(x, np.zeros((1, 2)), np.eye(2)[None], np.zeros((1, 2, 2)), np.ones(1), np.zeros(1, np.int32))
    ret, act = _static_maps(2, 3, 2, lambda I: np.where(I[..., 0] < 4, 0, 2))
    ora.set_static_map(ret, act)
    ora.refine()
    c = [ora.get_level(l)[0] for l in range(3)]
    A, G, T = 1, 2, 4
    assert sorted(map(tuple, np.argwhere(c[0] == A))) == [(0, 0), (0, 1)]
    assert not c[0][c[0] != A].any()
    assert not (c[1] == A).any()
    assert sorted(map(tuple, np.argwhere(c[1] == G))) == [(1, 0), (1, 1), (1, 2)]
    want2 = [(4, 0)] + [(i, j) for i in (4, 5) for j in (2, 3, 4, 5)]
    assert sorted(map(tuple, np.argwhere(c[2] == A))) == sorted(want2)
    assert sorted(map(tuple, np.argwhere(c[2] == G))) == [(3, j) for j in (0, 2, 3, 4, 5)]
    nm = ora.get_node_mask()
    assert sorted(map(tuple, np.argwhere(nm & T))) == [(4, y) for y in (1, 2, 3, 5, 6)]
    assert ora.get_particles()["fl"][0] == 2
