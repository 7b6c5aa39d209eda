"""CPU pins of the learned-prediction oracles (oracle/mlp_oracle.py, oracle/units_oracle.py).

The network and the unit features have no counterpart in the reference; what pins them is
mathematics: central finite differences and an independent autograd for the gradient,
PyTorch's own bfloat16 conversion for the rounding, brute force for the candidate set, and
hand-made particle sets for the features and labels.
"""
import itertools
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import mlp_oracle as M  # noqa: E402
from oracle import units_oracle as U  # noqa: E402


def test_bf16_rounding_known_values():
    one = 1.0
    # spacing of bfloat16 at 1 is 2^-7: halfway cases round to even
    assert M.bf16(np.float32(one + 2.0 ** -8)) == 1.0
    assert M.bf16(np.float32(one + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    assert M.bf16(np.float32(one + 2.0 ** -8 + 2.0 ** -20)) == 1.0 + 2.0 ** -7
    assert np.isinf(M.bf16(np.float32(np.inf)))
    assert np.isnan(M.bf16(np.float32(np.nan)))
    assert M.bf16(np.float32(-0.0)) == 0.0


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = (rng.standard_normal(100000) * 10.0 ** rng.uniform(-20, 20, 100000)).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(M.bf16(a), ref)


def _data(rows=37, seed=1):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.1, 64)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.1, 64)
    p[M.OFF_B3:] = rng.normal(0, 0.1, 4)
    X = rng.normal(0, 1, (rows, 16)).astype(np.float32)
    y = rng.integers(0, 4, rows)
    return p.astype(np.float64), X, y


def test_exact_gradient_central_differences():
    p, X, y = _data()
    cw = np.array([1.0, 2.0, 0.5, 1.5])
    _, g = M.loss_grad(p, X, y, cw, bf16_rounding=False)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False) for a, b in
                          ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                           (M.OFF_W3, M.OFF_B3))] + [np.arange(M.OFF_B3, M.NPARAM)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, X, y, cw, bf16_rounding=False)[0] - M.loss_grad(pm, X, y, cw, bf16_rounding=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), (i, fd, g[i])


def test_exact_gradient_matches_autograd():
    torch = pytest.importorskip("torch")
    p, X, y = _data(rows=50, seed=4)
    loss, g = M.loss_grad(p, X, y, bf16_rounding=False)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    Xt = torch.tensor(X, dtype=torch.float64)
    W1 = pt[M.OFF_W1:M.OFF_B1].view(64, 16)
    W2 = pt[M.OFF_W2:M.OFF_B2].view(64, 64)
    W3 = pt[M.OFF_W3:M.OFF_B3].view(4, 64)
    h1 = torch.relu(Xt @ W1.T + pt[M.OFF_B1:M.OFF_W2])
    h2 = torch.relu(h1 @ W2.T + pt[M.OFF_B2:M.OFF_W3])
    z = h2 @ W3.T + pt[M.OFF_B3:]
    lt = torch.nn.functional.cross_entropy(z, torch.tensor(y), reduction="mean")
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_rounded_gradient_equals_exact_when_rounding_is_exact():
    # weights, inputs and every intermediate exactly representable in bfloat16 -> the
    # rounded network is the exact one: small-integer weights and inputs, one hidden unit path
    p = np.zeros(M.NPARAM)
    p[M.OFF_W1 + 0 * 16 + 0] = 1.0
    p[M.OFF_W2 + 0 * 64 + 0] = 2.0
    p[M.OFF_W3 + 1 * 64 + 0] = 0.5
    X = np.zeros((3, 16), np.float32)
    X[:, 0] = [1.0, 2.0, 0.0]
    y = np.array([1, 0, 3])
    l1, g1 = M.loss_grad(p, X, y, bf16_rounding=False)
    l2, g2 = M.loss_grad(p, X, y, bf16_rounding=True)
    assert abs(l1 - l2) < 1e-15
    # dz is not representable in bfloat16 in general: only the bias gradients (unrounded sums)
    # and the paths through exact values must agree exactly
    np.testing.assert_allclose(g2[M.OFF_B3:], g1[M.OFF_B3:], rtol=0, atol=0)


def test_decide_rule():
    z = np.log(np.array([[0.5, 0.2, 0.2, 0.1],     # p0 >= tau
                         [0.1, 0.15, 0.1, 0.65],   # cum 0.1, 0.25, 0.35
                         [0.05, 0.05, 0.05, 0.85]]))
    lev = M.decide(z, 0.3, level=4, unit_level=1)
    assert lev.tolist() == [3, 1, -1]
    lev = M.decide(z, 0.2, level=3, unit_level=0)
    assert lev.tolist() == [2, 1, -1]


# ----------------------------------------------------------------------------- units oracle
def test_bidx_coords_roundtrip():
    u = U.Units(3, 4, 64)
    rng = np.random.default_rng(0)
    J = rng.integers(0, u.E, (1000, 3))
    np.testing.assert_array_equal(U.coords(U.bidx(J, u.NB, 3), u.NB, 3), J)
    assert U.bidx(np.array([u.E - 1] * 3), u.NB, 3) == u.nu - 1


def test_candidates_brute_force():
    u = U.Units(2, 3, 16)
    rng = np.random.default_rng(1)
    occ = np.zeros(u.nu, bool)
    occ[rng.choice(u.nu, 20, replace=False)] = True
    got = set(u.candidates(occ).tolist())
    J = U.coords(np.arange(u.nu), u.NB, 2)
    Jo = U.coords(np.flatnonzero(occ), u.NB, 2)
    want = {i for i in range(u.nu) if np.any(np.max(np.abs(Jo - J[i]), axis=1) <= 1)}
    assert got == want


def test_requests_hand_case():
    # L=3, c=16 (finest 64): ulev 0, unit = coarsest cell (1/16).  A level-2 particle at
    # x = (0.25 + 0.5/64, ...) has base cell int(16.5 + 0.5) - 1 = 16 -> cells 16, 17 ->
    # units 4 (both), so only unit (4, 4) gets level 2; a level-1 particle at x = 0.2 has
    # base int(6.4 + 0.5) - 1 = 5 -> cells 5, 6 -> units 2, 3.
    u = U.Units(2, 3, 16)
    x = np.array([[0.25 + 0.5 / 64, 0.25 + 0.5 / 64], [0.2, 0.2]], np.float32)
    g = np.array([2, 1])
    req = u.requests(x, g)
    nz = {tuple(U.coords([q], u.NB, 2)[0] - U.PAD): int(req[q]) for q in np.flatnonzero(req)}
    assert nz == {(4, 4): 3, (2, 2): 2, (2, 3): 2, (3, 2): 2, (3, 3): 2}


def test_features_hand_case():
    # two particles in one unit (L=3, c=16, ulev 0, dx_u = 1/16), v along +x
    u = U.Units(3, 3, 16)
    x = np.array([[0.51, 0.51, 0.51], [0.52, 0.53, 0.51]], np.float32)
    v = np.array([[1.0, 0, 0], [3.0, 0, 0]], np.float32)
    F = np.tile(np.eye(3), (2, 1, 1))
    F[1] *= 1.1
    C = np.zeros((2, 3, 3))
    C[0, 0, 1] = 2.0
    m = np.array([1.0, 3.0])
    g = np.array([2, 0])
    feats = u.features(x, v, F, C, m, g, p_vol=0.5, window=20, dt=1e-3)
    assert len(feats) == 27  # the occupied unit and its 26 neighbours
    q = int(u.unit_of(x[:1])[0])
    f = feats[q]
    dxu = 1 / 16
    win = np.float32(20) * np.float32(1e-3)
    assert np.isclose(f[0], 4.0 * 0.5 / (2.0 * dxu ** 3))          # mass / (rho_ref vol), rho_ref = 2 / 0.5
    assert np.isclose(f[1], 2.0 * win * 16)                          # mean |v| in units per window
    assert np.isclose(f[2], 3.0 * win * 16)
    assert np.isclose(f[3], 1.0 * win)                               # mean |C|
    assert np.isclose(f[4], (1.331 - 1) / 2 * 10)                    # mean det F - 1, x10
    assert np.isclose(f[5], 0.0)                                     # min det F - 1 = 0
    assert f[6] == 1.0                                               # max g = 2 (finest)
    assert f[7] == 1.0                                               # moved by v * window, still requests level 2 here
    assert f[8] == 1.0                                               # the level-2 request is in this unit
    assert np.isclose(f[12], 2.0 * win * 16) and f[13] == 0 and f[14] == 0
    assert f[9] == 0 and f[10] == 0 and f[15] == 0                   # no occupied neighbours
    # the +x neighbour sees this unit moving towards it: upwind level = finest
    J = U.coords([q], u.NB, 3)[0]
    qx = int(U.bidx(J + np.array([1, 0, 0]), u.NB, 3))
    qmx = int(U.bidx(J - np.array([1, 0, 0]), u.NB, 3))
    assert feats[qx][15] == 1.0 and feats[qmx][15] == 0.0
    assert np.isclose(feats[qx][9], f[0] / 26)
