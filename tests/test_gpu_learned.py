"""GPU tests of the learned refinement prediction: the network kernels (BF16 matrix cores)
against the CPU oracle (oracle/mlp_oracle.py), the unit features and window labels against
oracle/units_oracle.py, and the predicted-octree substeps (criterion check, rollback).

Bars (stated per test): the network's logits and gradients follow the oracle's BF16 rounding
points, so they agree to FP32 accumulation order (plus rare one-ulp BF16 rounding flips
where FP32 and FP64 intermediates straddle a rounding boundary); features are FP32 sums.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    from oracle import mlp_oracle

    return mlp_oracle


@pytest.fixture(scope="module")
def UO():
    from oracle import units_oracle

    return units_oracle


@pytest.fixture(scope="module")
def learned(pkg):
    return __import__(pkg.__name__ + ".learned", fromlist=["Trainer"])


def _params(M, seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.1, 64)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.1, 64)
    p[M.OFF_B3:] = rng.normal(0, 0.3, 4)
    return p.astype(np.float32)


def test_mlp_forward_matches_oracle(learned, M):
    """Logits of the BF16 matrix-core forward vs the oracle with the same rounding points:
    bound 5e-3 of the largest logit (FP32 vs FP64 accumulation, rare BF16 rounding flips)."""
    p = _params(M, 1)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (1007, 16)).astype(np.float32)  # not a multiple of the 16-row tile
    tr = learned.Trainer(2048, p)
    Z = tr.logits(X)
    tr.close()
    ref, _ = M.forward(p, X, bf16_rounding=True)
    err = np.abs(Z - ref).max() / np.abs(ref).max()
    print(f"forward max rel err {err:.2e}")
    assert err < 5e-3


def test_mlp_gradient_matches_oracle(learned, M):
    """Loss within 1e-5 relative, gradient within 5e-3 relative L2 (1e-2 per parameter block)
    of the CPU oracle computed in FP64 with the same BF16 rounding points."""
    p = _params(M, 3)
    rng = np.random.default_rng(4)
    rows = 4099
    X = rng.normal(0, 1, (rows, 16)).astype(np.float32)
    y = rng.integers(0, 4, rows).astype(np.int32)
    cw = np.array([1.0, 2.0, 0.5, 1.5], np.float32)
    scale = 1.0 / float(cw[y].sum())
    tr = learned.Trainer(rows, p)
    loss, gg = tr.grad(X, y, cw, scale)
    tr.close()
    lref, gref = M.loss_grad(p, X, y, cw, scale, bf16_rounding=True)
    gg = gg.astype(np.float64)
    rel = np.linalg.norm(gg - gref) / np.linalg.norm(gref)
    print(f"loss {loss:.6f} vs {lref:.6f}; gradient rel L2 err {rel:.2e}")
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    assert rel < 5e-3
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPARAM)):
        assert np.linalg.norm(gg[a:b] - gref[a:b]) <= 1e-2 * np.linalg.norm(gref[a:b]) + 1e-9, (a, b)


def test_mlp_adam_matches_numpy(learned):
    rng = np.random.default_rng(5)
    p = rng.normal(0, 1, 5508).astype(np.float32)
    tr = learned.Trainer(16, p)
    pn, mn, vn = p.astype(np.float64), np.zeros(5508), np.zeros(5508)
    isw = np.zeros(5508, bool)
    isw[0:1024] = isw[1088:5184] = isw[5248:5504] = True
    lr, b1, b2, eps, wd = 1e-2, 0.9, 0.999, 1e-8, 1e-3
    for step in range(1, 4):
        gr = rng.normal(0, 1, 5508).astype(np.float32)
        tr.apply(gr, lr, wd)
        mn = b1 * mn + (1 - b1) * gr
        vn = b2 * vn + (1 - b2) * gr.astype(np.float64) ** 2
        pn = np.where(isw, pn - lr * wd * pn, pn)
        pn = pn - lr * (mn / (1 - b1 ** step)) / (np.sqrt(vn / (1 - b2 ** step)) + eps)
    got = tr.params()
    tr.close()
    np.testing.assert_allclose(got, pn, rtol=1e-5, atol=1e-6)


def test_mlp_training_fits_a_teacher(learned, M):
    """Epochs of the device-resident trainer (gradient + Adam kernels) fit a random teacher's
    labels far better than chance (4 classes)."""
    rng = np.random.default_rng(6)
    teacher = M.init_params(7)
    X = rng.normal(0, 1, (32768, 16)).astype(np.float32)
    y = np.argmax(M.forward(teacher, X, bf16_rounding=False)[0], axis=1).astype(np.int32)
    tr = learned.Trainer(len(y), M.init_params(8))
    losses = []
    for ep in range(12):
        tr.set_data(X, y)
        losses.append(tr.epoch(rng.permutation(len(y)), 1024, np.ones(4), 3e-3))
    acc = float(np.mean(np.argmax(tr.logits(X), 1) == y))
    tr.close()
    print(f"teacher fit: losses {losses[0]:.3f} -> {losses[-1]:.3f}, accuracy {acc:.3f}")
    assert losses[-1] < 0.5 * losses[0] and acc > 0.8


# ----------------------------------------------------------------------------- unit features
def _small(pkg, n=20_000, steps=3):
    wl = pkg.scenes.small_3d(n=n)
    sim = pkg.AdaptiveGIMP(wl.dim, wl.level, wl.coarsest_size, wl.n, wl.initializer(), None, **wl.params)
    for _ in range(steps):
        sim.substep(wl.dt)
    return wl, sim


def test_unit_features_match_oracle(pkg, UO):
    wl, sim = _small(pkg)
    sim.learned_enable(20, 0.3)
    X, units = sim.learned_features(wl.dt)
    x, v, F, C, g = (getattr(sim, k).to_numpy() for k in ("x_p", "v_p", "F_p", "C_p", "g_p"))
    ref = UO.Units(3, wl.level, wl.coarsest_size).features(x, v, F, C, wl.state["m"], g, wl.params["p_vol"], 20,
                                                           wl.dt)
    sim.close()
    assert sorted(units.tolist()) == sorted(ref.keys())
    R = np.stack([ref[int(u)] for u in units])
    scale = np.maximum(np.abs(R).max(axis=0), 1e-6)
    err = (np.abs(X - R) / scale).max(axis=0)
    print("feature rel errs", np.array2string(err, precision=1))
    assert np.all(err < 2e-4)


def test_window_labels_match_oracle(pkg, UO):
    wl, sim = _small(pkg, steps=1)
    sim.learned_enable(8, 0.3)
    _, units = sim.learned_features(wl.dt)
    sim.learned_labels_reset()
    xs = []
    for _ in range(8):
        xs.append(sim.x_p.to_numpy().copy())
        sim.learned_labels_accumulate()
        sim.substep(wl.dt)
    got = sim.learned_labels()
    g = sim.g_p.to_numpy()
    sim.close()
    want = UO.Units(3, wl.level, wl.coarsest_size).window_classes(xs, g, units)
    assert np.array_equal(got, want), f"{(got != want).sum()} of {len(got)} labels differ"
    assert set(np.unique(got).tolist()) >= {0, 3}


def test_fused_predict_matches_oracle(pkg, M):
    wl, sim = _small(pkg)
    sim.learned_enable(20, 0.3)
    p = _params(M, 9)
    sim.learned_set_model(p)
    X, units = sim.learned_features(wl.dt)
    cls, lev = sim.learned_predict(wl.dt)
    sim.close()
    z, _ = M.forward(p, X, bf16_rounding=True)
    e = np.exp(z - z.max(axis=1, keepdims=True))
    pr = e / e.sum(axis=1, keepdims=True)
    cum = np.cumsum(pr[:, :3], axis=1)
    clear = np.all(np.abs(cum - 0.3) > 1e-3, axis=1)  # rows not on the decision boundary
    want = M.decide(z, 0.3, wl.level, 0)
    assert clear.mean() > 0.9
    assert len(cls) == len(units)
    # candidate order is not fixed between calls: compare through the dense per-unit levels
    assert np.array_equal(lev[units][clear], want[clear])
    assert np.all(lev[np.setdiff1d(np.arange(len(lev)), units)] == -1)


# ----------------------------------------------------------------------------- predicted octrees
def _const_model(cls):
    p = np.zeros(5508, np.float32)
    p[5504 + cls] = 20.0
    return p


def test_check_is_silent_on_plain_octrees(pkg):
    wl, sim = _small(pkg, steps=0)
    sim.learned_enable(20, 0.3)
    sim.advance(wl.dt, 30)
    sim.substep(wl.dt)
    c = sim.counters()
    sim.close()
    assert c["criterion_failures"] == 0


def _single(pkg, v=5.0):
    """One particle (level 2 of 3, coarsest 16) moving along x at constant speed, no gravity."""
    def init(sim):
        sim.x_p.from_numpy(np.array([[0.3, 0.5, 0.5]]))
        sim.v_p.from_numpy(np.array([[v, 0.0, 0.0]]))
        sim.F_p.from_numpy(np.eye(3)[None])
        sim.m_p.from_numpy(np.array([1.0]))
        sim.g_p.from_numpy(np.array([2], np.int32))

    return pkg.AdaptiveGIMP(3, 3, 16, 1, init, None, gravity=(0.0, 0.0, 0.0))


def test_rollback_keeps_the_trajectory(pkg):
    """A model that never refines ahead: every crossing of a finest-cell boundary fails the
    check, the window rolls back and is rebuilt there.  For a lone particle a reused octree
    gives the same transfers as a fresh one, so the trajectory must equal the plain one;
    a lost or repeated substep would shift x by v dt = 5e-4."""
    dt, n = 1e-4, 100
    plain = _single(pkg)
    plain.advance(dt, n)
    xp = plain.x_p.to_numpy()
    plain.close()
    sim = _single(pkg)
    sim.learned_enable(20, 0.3)
    sim.learned_set_model(_const_model(3))
    sim.learned_set_parts(network=True, sweep=False)
    sim.advance_learned(dt, n)
    xl = sim.x_p.to_numpy()
    st = sim.learned_stats()
    sim.close()
    print(st, xp, xl)
    assert st["forced"] >= 2 and st["substeps"] == n and st["rebuilds"] >= 5 + st["forced"] - 1
    assert np.abs(xl - xp).max() < 1e-6


def test_finest_prediction_never_forces(pkg):
    wl, sim = _small(pkg, steps=0)
    sim.learned_enable(20, 0.3)
    sim.learned_set_model(_const_model(0))
    sim.advance_learned(wl.dt, 60)
    st = sim.learned_stats()
    c = sim.counters()
    sim.close()
    print(st)
    assert st["forced"] == 0 and st["rebuilds"] == 3 and st["refined"] > 0
    assert c["nonfinite"] == 0


def test_swept_requests_cover_straight_motion(pkg):
    """With the swept requests the lone particle's window octree already holds every cell it
    requests on its straight path: no rollback, same trajectory as the plain method."""
    dt, n = 1e-4, 100
    plain = _single(pkg)
    plain.advance(dt, n)
    xp = plain.x_p.to_numpy()
    plain.close()
    sim = _single(pkg)
    sim.learned_enable(20, 0.3)
    sim.learned_set_model(_const_model(3))
    sim.learned_set_parts(network=False, sweep=True)
    sim.advance_learned(dt, n)
    st = sim.learned_stats()
    xl = sim.x_p.to_numpy()
    sim.close()
    assert st["forced"] == 0 and st["rebuilds"] == 5
    assert np.abs(xl - xp).max() < 1e-6


def test_trained_network_prevents_rollbacks(pkg, learned):
    """What the trained network gains over the method without it: on the small falling
    sphere, 100 substeps in windows of 20 with the predicted octree made of the current
    requests plus the network's units (no swept requests).  A model that predicts nothing
    (class 3 everywhere) leaves every boundary crossing to the criterion check, which cuts
    most windows short and rolls them back; the trained network predicts the refinement the
    window needs, so no window is cut short.  Both runs end at the same substep count."""
    wl = pkg.scenes.small_3d(seed=0)
    params, meta = learned.load_model()
    stats = {}
    for name, p in (("none", _const_model(3)), ("trained", params)):
        sim = pkg.AdaptiveGIMP(wl.dim, wl.level, wl.coarsest_size, wl.n, wl.initializer(), None, **wl.params)
        sim.learned_enable(meta["window"], 0.3)
        sim.learned_set_model(p)
        sim.learned_set_parts(network=True, sweep=False)
        sim.advance_learned(wl.dt, 100)
        stats[name] = sim.learned_stats()
        c = sim.counters()
        sim.close()
        assert stats[name]["substeps"] == 100 and c["nonfinite"] == 0
    print(stats)
    assert stats["none"]["forced"] >= 5 and stats["none"]["discarded"] >= 50
    assert stats["trained"]["forced"] <= 1 and stats["trained"]["refined"] > 0
    assert stats["trained"]["discarded"] * 4 < stats["none"]["discarded"]
