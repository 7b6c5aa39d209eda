"""Closed-form expectations for single-particle substeps, shared by the oracle pins (CPU) and
the GPU tests.  Each is derived from the physics, not from the oracle's or the kernels' code:

* boundary conditions: a lone particle with F = I and C = 0 gives every stencil node the
  velocity u = v_p + g dt (its own momentum over its own mass).  The grid update then zeroes
  nodes in the wall band (slip: only the normal component; sticky: every component) or in the
  slab, and G2P returns v' = sum_i w_i u_i with the quadratic B-spline weights, x' = x + dt v',
  C' = sum_i u_i (x) grad w_i.  With a single finest-level particle the restriction and
  prolongation pass these values through unchanged (each coarse node copies one fine node).
* plasticity (Stomakhin et al. 2013): with E = 0, v = 0 and no gravity, G2P's trial F is the
  stored F exactly; its singular values are clamped to [1 - theta_c, 1 + theta_s] and
  J_p <- J_p * prod(sigma / clamp(sigma)).
* stress with hardening: with v = 0 and C = 0, P2G's node momenta are p_i = S grad w_i, and
  the B-spline moment identity sum_i grad w_i (x) (x_i - x_p)/dx = I gives back
  S = sum_i p_i (x) (x_i - x_p)/dx = -dt V_p / dx * tau, where
  tau = 2 mu_h (F - R) F^T + lambda_h J (J - 1) I (fixed-corotated Kirchhoff stress, R the
  polar rotation of F) and mu_h, lambda_h carry the factor exp(xi (1 - J_p)).
"""
from __future__ import annotations

import itertools

import numpy as np
from scipy.linalg import polar


def bspline(t):
    """Quadratic B-spline weights and their derivatives at the 3 nodes b, b+1, b+2 of a point
    at x/dx = b + 1 + t, t in [-1/2, 1/2)."""
    return (np.array([0.5 * (0.5 - t) ** 2, 0.75 - t * t, 0.5 * (0.5 + t) ** 2]),
            np.array([t - 0.5, -2.0 * t, t + 0.5]))


def f32(v):
    return float(np.float32(v))


def bc_single_particle(x, v, dt, dx, finest_size, gravity, gap=4, wall_mode="sticky", slab=None):
    """(x', v', C') of one finest-level particle with F = I, C = 0 after one substep."""
    d = len(x)
    dt = f32(dt)  # the step is taken in float32 (adaptive_gimp.py:231)
    q = np.asarray(x, np.float64) / dx
    b = np.floor(q + 0.5).astype(int) - 1
    t = q - (b + 1)
    W = [bspline(t[a]) for a in range(d)]
    u = np.asarray(v, np.float64) + np.asarray(gravity, np.float64)[:d] * dt
    vn = np.zeros(d)
    C = np.zeros((d, d))
    for i in itertools.product(range(3), repeat=d):
        f = b + np.array(i)  # finest node index
        band = (f < gap) | (f > finest_size - gap)
        un = u.copy()
        if wall_mode == "sticky":
            if band.any():
                un[:] = 0.0
        else:
            un[band] = 0.0
        if slab is not None:
            ax, lo, hi = slab
            y = f[ax] * dx
            if lo <= y <= hi:
                un[:] = 0.0
        w = np.prod([W[a][0][i[a]] for a in range(d)])
        gw = np.array([W[a][1][i[a]] * np.prod([W[c][0][i[c]] for c in range(d) if c != a]) for a in range(d)])
        vn += w * un
        C += np.outer(un, gw)
    return np.asarray(x, np.float64) + dt * vn, vn, C


def clamp_closed_form(F, Jp, theta_c, theta_s):
    """(F', J_p') of the singular-value clamp on a trial F with det F > 0."""
    U, s, Vt = np.linalg.svd(F)
    if np.linalg.det(U) < 0:  # make both factors rotations (keeps U diag(s) V^T)
        U[:, -1] *= -1
        s[-1] *= -1
    if np.linalg.det(Vt) < 0:
        Vt[-1] *= -1
        s[-1] *= -1
    ns = np.clip(s, 1.0 - theta_c, 1.0 + theta_s)
    return U @ np.diag(ns) @ Vt, Jp * np.prod(s / ns)


def kirchhoff_hardened(F, E, nu, Jp, xi):
    mu = E / (2 * (1 + nu)) * np.exp(xi * (1 - Jp))
    la = E * nu / ((1 + nu) * (1 - 2 * nu)) * np.exp(xi * (1 - Jp))
    R, _ = polar(F)
    J = np.linalg.det(F)
    return 2 * mu * (F - R) @ F.T + la * J * (J - 1) * np.eye(F.shape[0])


def stress_from_grid(m, v, x, dx):
    """S = sum_i p_i (x) (x_i - x_p)/dx over the nodes of a dense level with mass (p = m v)."""
    d = len(x)
    idx = np.argwhere(m > 0)
    q = np.asarray(x, np.float64) / dx
    S = np.zeros((d, d))
    for I in idx:
        p = m[tuple(I)] * v[tuple(I)]
        S += np.outer(p, I - q)
    return S


def random_rotation(d, rng):
    Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


# single-particle boundary cases: (name, position in finest cells, velocity, wall mode, slab);
# 2D cases use the first two coordinates
# finest_size 16 (coarsest 4 at 3 levels, or 16 at one level), boundary gap 4: the wall band is
# finest node index < 4 or > 12.
BC_CASES = [
    ("x_wall_straddle_sticky", (4.3, 8.1, 7.6), (0.3, -0.2, 0.1), "sticky", None),
    ("x_wall_straddle_slip", (4.3, 8.1, 7.6), (0.3, -0.2, 0.1), "slip", None),
    ("high_wall_straddle_sticky", (8.4, 11.8, 7.7), (-0.1, 0.4, 0.2), "sticky", None),
    ("inside_band_sticky", (2.2, 8.3, 7.9), (0.3, -0.2, 0.1), "sticky", None),
    ("inside_band_slip", (2.2, 8.3, 7.9), (0.3, -0.2, 0.1), "slip", None),
    # slab on axis 1 over [0.5, 0.6]: nodes 8 (y = 0.5 exactly, the closed lower end) and 9
    ("slab_lower_end", (5.3, 7.4, 8.8), (0.2, -0.5, 0.1), "sticky", (1, 0.5, 0.6)),
    ("slab_straddle", (5.3, 9.8, 8.8), (0.2, -0.5, 0.1), "sticky", (1, 0.5, 0.6)),
    ("slab_inside", (5.3, 8.6, 8.8), (0.2, -0.5, 0.1), "sticky", (1, 0.45, 0.7)),
]


# hand-derived static-adaptivity cases (2D, one particle at (0.49, 0.5)): the grid initializer
# returns (and activates) level `left` for finest x-index < 4 and `right` elsewhere.  Expected
# activated (A) and ghost (G) cells per level, T-junction nodes (finest index) and fl_p; the
# derivations are in the docstrings of tests/test_oracle_pins.py's test_static_*_hand_derived.
STATIC_CASES = [
    dict(name="two_levels_shift1", L=2, c=4, left=0, right=1,
         A={0: [(1, 1), (1, 2)], 1: [(i, j) for i in (4, 5) for j in (2, 3, 4, 5)]},
         G={0: [], 1: [(3, j) for j in (2, 3, 4, 5)]},
         T=[(4, 3), (4, 5)], fl=1),
    dict(name="three_levels_shift2", L=3, c=2, left=0, right=2,
         A={0: [(0, 0), (0, 1)], 1: [], 2: sorted([(4, 0)] + [(i, j) for i in (4, 5) for j in (2, 3, 4, 5)])},
         G={0: [], 1: [(1, 0), (1, 1), (1, 2)], 2: [(3, j) for j in (0, 2, 3, 4, 5)]},
         T=[(4, y) for y in (1, 2, 3, 5, 6)], fl=2),
]


def static_case_levels(case, I):
    """Level the case's grid initializer returns for finest cells I[..., 2]."""
    return np.where(np.asarray(I)[..., 0] < 4, case["left"], case["right"])


def check_static_case(case, cells, node_mask, fl):
    A, G, T = 1, 2, 4
    for l in range(case["L"]):
        assert sorted(map(tuple, np.argwhere(cells[l] == A))) == sorted(case["A"][l]), l
        assert sorted(map(tuple, np.argwhere(cells[l] == G))) == sorted(case["G"][l]), l
        assert set(np.unique(cells[l])) <= {0, A, G}, l
    assert sorted(map(tuple, np.argwhere(node_mask & T))) == sorted(case["T"])
    assert fl == case["fl"]
