"""Learned refinement prediction: training data from the plain method's own runs, training on
the GPU with the library's BF16 matrix-core kernels (``agimp_mlp_*``), and the model file.

Not in the reference, which rebuilds its octree from the current particles every substep
(adaptive_gimp.py:323-337).  What is learned (see include/agimp.h, agimp_learned_*):

* a sample is one candidate unit (a cell of level max(0, L-3), i.e. 4^d finest cells) at a
  window start t, with the 16 unit features the library computes from the particles;
* its label is the class of the finest level the plain criterion requests inside the unit
  during substeps t .. t+window-1 of a plain run (0 finest, 1 one coarser, 2 two coarser,
  3 none or coarser);
* the network (16-64-64-4) is trained by softmax cross-entropy with Adam; forward, backward
  and the update all run in the library's kernels, PyTorch only holds the device tensors.

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
    """Adam on the library's kernels; tensors live on the GPU (PyTorch as device memory)."""

    def __init__(self, params, lr=3e-3, weight_decay=1e-4):
        import torch

        self.torch = torch
        self.L = _pkg()._lib.load()
        dev = torch.device("cuda")
        self.p = torch.tensor(np.asarray(params, np.float32), device=dev)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.g = torch.zeros_like(self.p)
        self.loss = torch.zeros(1, device=dev)
        self.step_no = 0
        self.lr, self.wd = lr, weight_decay

    def _s(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def grad(self, X, y, cw=None, scale=None):
        """loss and gradient (device tensors) of scale * sum cw[y] CE on the rows (X, y)."""
        rows = X.shape[0]
        if scale is None:
            scale = 1.0 / float(cw[y].sum().item()) if cw is not None else 1.0 / rows
        rc = self.L.agimp_mlp_loss_grad(ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(y.data_ptr()), rows,
                                        ctypes.c_void_p(self.p.data_ptr()),
                                        ctypes.c_void_p(cw.data_ptr()) if cw is not None else None,
                                        float(scale), ctypes.c_void_p(self.g.data_ptr()),
                                        ctypes.c_void_p(self.loss.data_ptr()), self._s())
        if rc != 0:
            raise RuntimeError(f"agimp_mlp_loss_grad failed ({rc})")
        return self.loss, self.g

    def step(self, lr=None):
        self.step_no += 1
        rc = self.L.agimp_mlp_adam(ctypes.c_void_p(self.p.data_ptr()), ctypes.c_void_p(self.g.data_ptr()),
                                   ctypes.c_void_p(self.m.data_ptr()), ctypes.c_void_p(self.v.data_ptr()),
                                   self.step_no, float(lr if lr is not None else self.lr), 0.9, 0.999, 1e-8,
                                   float(self.wd), self._s())
        if rc != 0:
            raise RuntimeError(f"agimp_mlp_adam failed ({rc})")

    def logits(self, X):
        out = self.torch.empty((X.shape[0], 4), device=X.device, dtype=self.torch.float32)
        rc = self.L.agimp_mlp_forward(ctypes.c_void_p(X.data_ptr()), X.shape[0], ctypes.c_void_p(self.p.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr()), self._s())
        if rc != 0:
            raise RuntimeError(f"agimp_mlp_forward failed ({rc})")
        return out


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


def train(X, y, X_val=None, y_val=None, epochs=8, batch=16384, lr=3e-3, seed=0, log=None):
    import torch

    tr = Trainer(init_params(seed), lr=lr)
    dev = torch.device("cuda")
    Xd = torch.tensor(np.ascontiguousarray(X, np.float32), device=dev)
    yd = torch.tensor(np.ascontiguousarray(y, np.int32), device=dev)
    freq = np.bincount(y, minlength=4).astype(np.float64) + 1.0
    cw_np = (freq.sum() / freq) ** 0.5
    cw_np /= cw_np[y].mean() if len(y) else 1.0
    cw = torch.tensor(cw_np.astype(np.float32), device=dev)
    n = len(y)
    steps = epochs * max(1, n // batch)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    hist = []
    k = 0
    for ep in range(epochs):
        perm = torch.randperm(n, device=dev, generator=gen)
        tot, cnt = 0.0, 0
        for b0 in range(0, n - batch + 1 if n >= batch else 1, batch):
            idx = perm[b0:b0 + batch]
            xb = Xd.index_select(0, idx).contiguous()
            yb = yd.index_select(0, idx).contiguous()
            loss, _ = tr.grad(xb, yb, cw)
            cur = lr * 0.5 * (1 + np.cos(np.pi * k / steps))  # cosine decay
            tr.step(max(cur, lr * 0.02))
            k += 1
            tot += float(loss.item())
            cnt += 1
        rec = {"epoch": ep, "loss": tot / max(cnt, 1)}
        if X_val is not None:
            rec.update(evaluate(tr, X_val, y_val))
        hist.append(rec)
        if log:
            log(json.dumps(rec))
    return tr.p.cpu().numpy(), hist, cw_np


def evaluate(tr, X, y, taus=(0.2, 0.3, 0.5)):
    import torch

    z = tr.logits(torch.tensor(np.ascontiguousarray(X, np.float32), device="cuda")).cpu().numpy()
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
