"""Learned refinement prediction: training data from the plain method's own runs, training on
the GPU with the library's BF16 matrix-core kernels (``agimp_mlp_*``), and the model file.

Not in the reference, which rebuilds its octree from the current particles every substep
(adaptive_gimp.py:323-337).  What is learned (see include/agimp.h, agimp_learned_*):

* a sample is one candidate unit (a cell of level max(0, L-3), i.e. 4^d finest cells) at a
  window start t, with the 16 unit features the library computes from the particles;
* its label is the class of the finest level the plain criterion requests inside the unit
  during substeps t .. t+window-1 of a plain run (0 finest, 1 one coarser, 2 two coarser,
  3 none or coarser);
* the network (16-64-64-4) is trained by class-weighted softmax cross-entropy with Adam;
  the dataset, parameters and optimiser state live in device memory owned by the library
  (agimp_trainer_*), and forward, backward and update run in its kernels.

    python adaptive-gimp_amd/learned.py train --runs 200 --out refine_mlp.npz

The trained parameters ship as ``refine_mlp.npz`` next to this file.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "refine_mlp.npz")
NPARAM = 5508
WINDOW = 20


def _pkg():
    root = os.path.dirname(HERE)
    if root not in sys.path:
        sys.path.insert(0, root)
    import __graft_entry__ as ge

    return ge.package()


def init_params(seed=0):
    """He-uniform weights, zero biases (float32), parameter layout of agimp_mlp_forward."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for off, fan_in, n in ((0, 16, 1024), (1088, 64, 4096), (5248, 64, 256)):
        lim = np.sqrt(6.0 / fan_in)
        p[off:off + n] = rng.uniform(-lim, lim, n).astype(np.float32)
    return p


def save_model(path, params, meta):
    np.savez(path, params=np.asarray(params, np.float32), meta=np.array(json.dumps(meta)))


def load_model(path=MODEL_PATH):
    """(parameters float32 [5508], metadata dict)."""
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), json.loads(str(z["meta"]))


# ----------------------------------------------------------------------------- data
def run_samples(pkg, wl, windows, window=WINDOW, cap_per_window=8192, seed=0, tau=0.3):
    """Features and window labels from one plain run of workload `wl`: `windows` windows of
    `window` substeps each, at most `cap_per_window` random candidates kept per window."""
    rng = np.random.default_rng(seed)
    sim = pkg.AdaptiveGIMP(wl.dim, wl.level, wl.coarsest_size, wl.n, wl.initializer(), wl.grid_initializer,
                           **wl.params)
    try:
        sim.learned_enable(window, tau)
        Xs, ys = [], []
        for _ in range(windows):
            X, _units = sim.learned_features(wl.dt)
            sim.learned_labels_reset()
            for _k in range(window):
                sim.learned_labels_accumulate()
                sim.substep(wl.dt)
            y = sim.learned_labels()
            if len(y) > cap_per_window:
                keep = rng.choice(len(y), cap_per_window, replace=False)
                X, y = X[keep], y[keep]
            Xs.append(X)
            ys.append(y)
        return np.concatenate(Xs), np.concatenate(ys).astype(np.int32)
    finally:
        sim.close()


def generate(seeds, n=200_000, windows=10, window=WINDOW, log=None):
    pkg = _pkg()
    Xs, ys, owner = [], [], []
    for s in seeds:
        t0 = time.time()
        wl = pkg.scenes.training_scene(s, n)
        X, y = run_samples(pkg, wl, windows, window, seed=s)
        Xs.append(X)
        ys.append(y)
        owner.append(np.full(len(y), s, np.int32))
        if log:
            log(f"run seed={s} L={wl.level} rows={len(y)} classes={np.bincount(y, minlength=4).tolist()} "
                f"{time.time() - t0:.1f}s")
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(owner)


# ----------------------------------------------------------------------------- training
class Trainer:
    """The library's device-resident trainer (agimp_trainer_*): forward, backward and Adam
    run in the library's kernels on data it holds in device memory."""

    def __init__(self, capacity, params=None, device=0):
        self._lib = _pkg()._lib
        self.L = self._lib.load()
        h = ctypes.c_void_p()
        self._lib.check(self.L.agimp_trainer_create(int(capacity), int(device), ctypes.byref(h)),
                        what="agimp_trainer_create")
        self.h = h
        self.cap = int(capacity)
        self.set_params(init_params(0) if params is None else params)

    def close(self):
        if getattr(self, "h", None):
            self.L.agimp_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _f(self, a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def _i(self, a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._lib.check(self.L.agimp_trainer_set_params(self.h, self._f(p)), what="agimp_trainer_set_params")

    def params(self):
        p = np.empty(NPARAM, np.float32)
        self._lib.check(self.L.agimp_trainer_get_params(self.h, self._f(p)), what="agimp_trainer_get_params")
        return p

    def set_data(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.int32)
        self._lib.check(self.L.agimp_trainer_set_data(self.h, self._f(X), self._i(y), len(y)),
                        what="agimp_trainer_set_data")
        self.rows = len(y)

    def epoch(self, perm, batch, class_weight, lr, weight_decay=1e-4):
        perm = np.ascontiguousarray(perm, np.int32)
        cw = np.ascontiguousarray(class_weight, np.float32)
        out = ctypes.c_double(0.0)
        self._lib.check(self.L.agimp_trainer_epoch(self.h, self._i(perm), int(batch), self._f(cw),
                                                   1.0 / min(batch, self.rows), float(lr), float(weight_decay),
                                                   ctypes.byref(out)), what="agimp_trainer_epoch")
        return out.value

    def logits(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Z = np.empty((len(X), 4), np.float32)
        for a in range(0, len(X), self.cap):
            b = min(len(X), a + self.cap)
            self._lib.check(self.L.agimp_trainer_logits(self.h, self._f(X[a:b]), b - a, self._f(Z[a:b])),
                            what="agimp_trainer_logits")
        self.rows = 0
        return Z

    def grad(self, X, y, class_weight=None, scale=None):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.int32)
        cw = None if class_weight is None else np.ascontiguousarray(class_weight, np.float32)
        if scale is None:
            scale = 1.0 / (float(cw[y].sum()) if cw is not None else len(y))
        g = np.empty(NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        self._lib.check(self.L.agimp_trainer_grad(self.h, self._f(X), self._i(y), len(y),
                                                  self._f(cw) if cw is not None else None, float(scale),
                                                  self._f(g), self._f(loss)), what="agimp_trainer_grad")
        self.rows = 0
        return float(loss[0]), g

    def apply(self, g, lr, weight_decay=0.0):
        g = np.ascontiguousarray(g, np.float32)
        self._lib.check(self.L.agimp_trainer_apply(self.h, self._f(g), float(lr), float(weight_decay)),
                        what="agimp_trainer_apply")


def decide(z, tau):
    """Class per row: first c whose cumulative softmax probability reaches tau (3 if none)."""
    z = np.asarray(z, np.float64)
    e = np.exp(z - z.max(axis=1, keepdims=True))
    p = e / e.sum(axis=1, keepdims=True)
    cum = np.cumsum(p[:, :3], axis=1)
    c = np.full(len(z), 3)
    for k in (2, 1, 0):
        c = np.where(cum[:, k] >= tau, k, c)
    return c


def class_weights(y):
    """Inverse-square-root class frequencies, normalised to a mean of 1 over the rows."""
    freq = np.bincount(y, minlength=4).astype(np.float64) + 1.0
    cw = (freq.sum() / freq) ** 0.5
    return cw / cw[y].mean()


def train(X, y, X_val=None, y_val=None, epochs=8, batch=16384, lr=3e-3, seed=0, log=None):
    rng = np.random.default_rng(seed)
    tr = Trainer(len(y), init_params(seed))
    cw = class_weights(y)
    hist = []
    for ep in range(epochs):
        tr.set_data(X, y)
        cur = lr * (0.02 + 0.98 * 0.5 * (1 + np.cos(np.pi * ep / max(1, epochs - 1))))  # cosine decay per epoch
        loss = tr.epoch(rng.permutation(len(y)), batch, cw, cur)
        rec = {"epoch": ep, "lr": float(cur), "loss": float(loss)}
        if X_val is not None:
            rec.update(evaluate(tr, X_val, y_val))
        hist.append(rec)
        if log:
            log(json.dumps(rec))
    p = tr.params()
    tr.close()
    return p, hist, cw


def evaluate(tr, X, y, taus=(0.1, 0.2, 0.3, 0.5)):
    z = tr.logits(X)
    out = {"val_acc_argmax": float(np.mean(np.argmax(z, 1) == y))}
    for t in taus:
        c = decide(z, t)
        # a prediction coarser than the label (larger class) risks a failed criterion check
        out[f"under_{t}"] = float(np.mean(c > y))
        out[f"over_{t}"] = float(np.mean(c < y))
    return out


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--runs", type=int, default=200)
    t.add_argument("--n", type=int, default=200_000)
    t.add_argument("--windows", type=int, default=10)
    t.add_argument("--epochs", type=int, default=8)
    t.add_argument("--val-runs", type=int, default=20)
    t.add_argument("--out", default=MODEL_PATH)
    a = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    seeds = list(range(1, a.runs + 1))
    t0 = time.time()
    X, y, owner = generate(seeds, a.n, a.windows, log=log)
    val = owner > a.runs - a.val_runs
    log(f"dataset rows={len(y)} classes={np.bincount(y, minlength=4).tolist()} gen={time.time() - t0:.0f}s")
    params, hist, cw = train(X[~val], y[~val], X[val], y[val], epochs=a.epochs, log=log)
    meta = {"window": WINDOW, "runs": a.runs, "val_runs": a.val_runs, "n": a.n, "windows": a.windows,
            "rows": int(len(y)), "class_counts": np.bincount(y, minlength=4).tolist(), "class_weight": cw.tolist(),
            "history": hist, "seeds": [1, a.runs]}
    save_model(a.out, params, meta)
    log(f"saved {a.out} ({time.time() - t0:.0f}s)")


if __name__ == "__main__":
    main()
