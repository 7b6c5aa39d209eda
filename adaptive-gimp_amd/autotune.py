"""Autotuner for the P2G work-chunk size (particles of one leaf block handled by one
workgroup; the "particles per block" of the batched workload).

Small chunks spread a dense leaf block over more workgroups (more parallelism, more tile
flushes); large chunks amortise the LDS tile and its global flush.  The best size depends
on particle density per level and on scene size, so it is measured, not guessed:

* ``tune(sim, dt)`` times whole substeps of one simulator for every candidate;
* ``tune_batch(batch, dt)`` groups the scenes of a batch into size classes and runs a
  coordinate descent over the classes, timing the captured batched substep.

Each trial restores the particle snapshot taken before tuning, so tuning has no effect on
the simulation.  Results can be cached in a JSON file (``AGIMP_TUNE_CACHE``) keyed by the
scene signature.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

CANDIDATES = (64, 128, 256, 512, 1024, 2048)
_FIELDS = ("x_p", "v_p", "F_p", "C_p", "Jp_p", "g_p")


def snapshot(sim):
    return {k: getattr(sim, k).to_numpy() for k in _FIELDS}


def restore(sim, snap):
    for k, a in snap.items():
        getattr(sim, k).from_numpy(a)
    sim._push()


def signature(sim):
    n = sim.n_particles
    bucket = int(round(n, -max(0, len(str(n)) - 2)))  # two significant digits
    return f"d{sim.dim}_L{sim.level}_c{sim.coarsest_size}_n{bucket}"


def _cache_load():
    path = os.environ.get("AGIMP_TUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _cache_store(key, value):
    path = os.environ.get("AGIMP_TUNE_CACHE")
    if not path:
        return
    c = _cache_load()
    c[key] = value
    with open(path, "w") as f:
        json.dump(c, f, indent=1)


def _time_advance(fn, steps):
    fn(1)  # warm the chosen configuration (graph capture on first use)
    t0 = time.perf_counter()
    fn(steps)
    return (time.perf_counter() - t0) / steps


def tune(sim, dt, candidates=CANDIDATES, steps=4, use_cache=True):
    """Pick the chunk size with the fastest substep for this simulator; returns
    (best, {candidate: seconds per substep})."""
    key = signature(sim)
    if use_cache:
        c = _cache_load().get(key)
        if c:
            sim.set_particles_per_chunk(c["best"])
            return c["best"], c["times"]
    snap = snapshot(sim)
    times = {}
    for ch in candidates:
        sim.set_particles_per_chunk(ch)
        restore(sim, snap)

        def run(k):
            sim.advance(dt, k)
            sim.sync()

        times[ch] = _time_advance(run, steps)
    best = min(times, key=times.get)
    sim.set_particles_per_chunk(best)
    restore(sim, snap)
    _cache_store(key, {"best": best, "times": times})
    return best, times


def size_classes(ns, edges=(20_000, 80_000)):
    """Class index of every scene by particle count."""
    return np.searchsorted(np.asarray(edges), np.asarray(ns), side="right")


def tune_batch(batch, dt, candidates=(64, 128, 256, 512), steps=3, sweeps=1):
    """Coordinate descent over size classes: for each class try every candidate with the
    other classes fixed, keep the fastest.  Returns (per-scene chunk sizes, trial log)."""
    sims = batch.sims
    cls = size_classes([s.n_particles for s in sims])
    snaps = [snapshot(s) for s in sims]
    choice = {int(c): candidates[len(candidates) // 2] for c in np.unique(cls)}
    log = []

    def apply():
        ch = np.array([choice[int(c)] for c in cls], np.int32)
        batch.set_particles_per_chunk(ch)
        for s, sn in zip(sims, snaps):
            restore(s, sn)
        return ch

    def run(k):
        batch.advance(dt, k)

    for _ in range(sweeps):
        for c in sorted(choice):
            best_t, best_ch = None, choice[c]
            for cand in candidates:
                choice[c] = cand
                apply()
                t = _time_advance(run, steps)
                log.append({"class": c, "chunk": cand, "s_per_substep": t})
                if best_t is None or t < best_t:
                    best_t, best_ch = t, cand
            choice[c] = best_ch
    ch = apply()
    return ch, log
