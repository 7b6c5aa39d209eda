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
