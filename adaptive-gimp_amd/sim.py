"""``AdaptiveGIMP``: drop-in for the simulator class of g1n0st/adaptive-gimp.

Same constructor (``adaptive_gimp.py:6-11``), ``substep(dt0)`` (``:323``), the same particle
fields (``x_p v_p F_p C_p m_p c_p g_p fl_p``, ``:55-62``) and per-level grid fields
(``cell_mask ad_grid_v ad_grid_m``, ``:39-42``; ``node_mask``, ``:23-26``), plus an explicit
``refine()``.  Fields behave like Taichi fields from Python scope: ``to_numpy()``,
``from_numpy()``, indexing.  All computation runs in libagimp.so on the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .scenes import extents, static_level_maps

UNACTIVATED, ACTIVATED, GHOST, T_JUNCTION, CULLING = 0, 1, 2, 4, 8  # utils.py:5-9


class ParticleField:
    """Host view of one particle field, synchronised with the device on access."""

    def __init__(self, sim, name):
        self._sim, self._name = sim, name

    @property
    def shape(self):
        return self._sim._host[self._name].shape

    def to_numpy(self):
        self._sim._pull()
        return self._sim._host[self._name].copy()

    def from_numpy(self, arr):
        self._sim._pull()
        h = self._sim._host[self._name]
        h[...] = np.asarray(arr, h.dtype).reshape(h.shape)
        self._sim._host_dirty = True

    def fill(self, value):
        self._sim._pull()
        self._sim._host[self._name][...] = value
        self._sim._host_dirty = True

    def __getitem__(self, i):
        self._sim._pull()
        return self._sim._host[self._name][i]

    def __setitem__(self, i, value):
        self._sim._pull()
        self._sim._host[self._name][i] = value
        self._sim._host_dirty = True

    def __array__(self, dtype=None, copy=None):
        a = self.to_numpy()
        return a if dtype is None else a.astype(dtype)


class GridField:
    """Read-only dense view of one level's cell_mask / ad_grid_m / ad_grid_v."""

    def __init__(self, sim, level, kind):
        self._sim, self._l, self._kind = sim, level, kind

    @property
    def shape(self):
        e = self._sim.extents[self._l]
        shp = (e,) * self._sim.dim
        return shp + ((self._sim.dim,) if self._kind == "v" else ())

    def to_numpy(self):
        return self._sim._read_level(self._l)[self._kind]

    def __getitem__(self, i):
        return self.to_numpy()[i]


class NodeMaskField:
    def __init__(self, sim):
        self._sim = sim

    @property
    def shape(self):
        return (self._sim.extents[-1],) * self._sim.dim

    def to_numpy(self):
        return self._sim._read_node_mask()

    def __getitem__(self, i):
        return self.to_numpy()[i]


class AdaptiveGIMP:
    """GPU adaptive GIMP simulator (reference: ``adaptive_gimp.py:4-352``).

    Keyword parameters extend the reference's hard-coded constants; every default is the
    reference's value: ``E=5e3, nu=0.2`` (``:52``), ``gravity`` -100 on axis 1 (``:50``),
    ``p_mass=1`` (``:49``), particle volume 1 (implicit in the stress scale, ``:240``),
    ``boundary_gap=4`` (``:18``), slip walls (``:284-287``), no plasticity.
    """

    def __init__(self, dim, level, coarsest_size, n_particles, particle_initializer, grid_initializer=None, *,
                 E=5e3, nu=0.2, gravity=None, p_mass=1.0, p_vol=1.0, boundary_gap=4, wall_mode="slip",
                 plastic=False, theta_c=2.5e-2, theta_s=7.5e-3, hardening=0.0, slab=None,
                 particles_per_chunk=None, device=0):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self._kwargs = dict(E=E, nu=nu, gravity=None if gravity is None else [float(v) for v in gravity],
                            p_mass=p_mass, p_vol=p_vol, boundary_gap=boundary_gap, wall_mode=wall_mode,
                            plastic=bool(plastic), theta_c=theta_c, theta_s=theta_s, hardening=hardening,
                            slab=None if slab is None else [int(slab[0]), float(slab[1]), float(slab[2])],
                            particles_per_chunk=None if particles_per_chunk is None else int(particles_per_chunk),
                            device=int(device))
        self.dim = dim
        self.level = level
        self.coarsest_size = coarsest_size
        self.coarsest_dx = 1 / coarsest_size                      # adaptive_gimp.py:16
        self.finest_size = coarsest_size * (2 ** (level - 1))     # :17
        self.boundary_gap = boundary_gap                          # :18
        self.leaf_size = 4                                        # :28
        self.radius = 0.5                                         # :48
        self.p_mass = p_mass                                      # :49
        g = [-100.0 if a == 1 else 0.0 for a in range(dim)] if gravity is None else list(gravity)
        self.gravity = np.asarray(g, np.float64)                  # :50
        self.n_particles = int(n_particles)                       # :51
        self.E, self.nu = E, nu
        self.mu = E / (2 * (1 + nu))                              # :53
        self.la = E * nu / ((1 + nu) * (1 - 2 * nu))
        self.extents = extents(level, coarsest_size)

        self._L = _lib.load()
        cfg = _lib.AgimpConfig()
        _lib.check(self._L.agimp_config_init(ctypes.byref(cfg), dim, level, coarsest_size, self.n_particles),
                   what="agimp_config_init")
        cfg.boundary_gap = boundary_gap
        cfg.E, cfg.nu = E, nu
        for a in range(3):
            cfg.gravity[a] = float(g[a]) if a < len(g) else 0.0
        cfg.p_vol = p_vol
        cfg.wall_mode = {"slip": 0, "sticky": 1}.get(wall_mode, wall_mode) if isinstance(wall_mode, str) \
            else int(wall_mode)
        cfg.plastic = int(bool(plastic))
        cfg.theta_c, cfg.theta_s, cfg.hardening = theta_c, theta_s, hardening
        if slab is not None:
            cfg.slab_enable, cfg.slab_axis, cfg.slab_lo, cfg.slab_hi = 1, int(slab[0]), float(slab[1]), float(slab[2])
        if particles_per_chunk is not None:  # else the library default (1024)
            cfg.particles_per_chunk = int(particles_per_chunk)
        cfg.device = int(device)
        self._cfg = cfg
        h = ctypes.c_void_p()
        _lib.check(self._L.agimp_create(ctypes.byref(cfg), ctypes.byref(h)), what="agimp_create")
        self._h = h

        n, d = self.n_particles, dim
        self._host = {
            "x_p": np.zeros((n, d), np.float32), "v_p": np.zeros((n, d), np.float32),
            "F_p": np.zeros((n, d, d), np.float32), "C_p": np.zeros((n, d, d), np.float32),
            "m_p": np.zeros(n, np.float32), "c_p": np.zeros((n, 3), np.float32),
            "g_p": np.zeros(n, np.int32), "fl_p": np.zeros(n, np.int32), "Jp_p": np.ones(n, np.float32),
        }
        self._host_dirty = True
        self._device_newer = False
        self._last_candidates = 0
        self._level_cache = None
        self._node_cache = None
        for name in self._host:
            setattr(self, name, ParticleField(self, name))
        self.cell_mask = [GridField(self, l, "cell") for l in range(level)]
        self.ad_grid_m = [GridField(self, l, "m") for l in range(level)]
        self.ad_grid_v = [GridField(self, l, "v") for l in range(level)]
        self.node_mask = NodeMaskField(self)

        particle_initializer(self)                                # :64
        self.set_grid_initializer(grid_initializer)               # :65-66
        self._push()

    def set_grid_initializer(self, grid_initializer):
        """Static adaptivity from ``grid_initializer`` (adaptive_gimp.py:65-66, :72-85), or
        dynamic adaptivity with None.  Takes effect at the next rebuild."""
        self.grid_initializer = grid_initializer
        self.static_adaptivity = grid_initializer is not None
        if self.static_adaptivity:
            ret, act = static_level_maps(grid_initializer, self.dim, self.level, self.coarsest_size)
            self._static_maps = (np.ascontiguousarray(ret), np.ascontiguousarray(act))
            _lib.check(self._L.agimp_set_static_levels(self._h, _lib.ptr(self._static_maps[0], ctypes.c_int8),
                                                       _lib.ptr(self._static_maps[1], ctypes.c_int8)),
                       self._h, "agimp_set_static_levels")
        else:
            self._static_maps = None
            _lib.check(self._L.agimp_set_static_levels(self._h, None, None), self._h, "agimp_set_static_levels")

    # ------------------------------------------------------------------ lifecycle
    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.agimp_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ host <-> device
    def _push(self):
        if not self._host_dirty:
            return
        H = self._host
        arrs = [np.ascontiguousarray(H[k], np.float32) for k in ("x_p", "v_p", "F_p", "C_p", "m_p")]
        g = np.ascontiguousarray(H["g_p"], np.int32)
        jp = np.ascontiguousarray(H["Jp_p"], np.float32)
        f = ctypes.c_float
        _lib.check(self._L.agimp_set_particles(self._h, *[_lib.ptr(a, f) for a in arrs], _lib.ptr(g, ctypes.c_int32),
                                               _lib.ptr(jp, f)), self._h, "agimp_set_particles")
        self._host_dirty = False

    def _pull(self):
        if not self._device_newer:
            return
        H = self._host
        f, i = ctypes.c_float, ctypes.c_int32
        _lib.check(self._L.agimp_get_particles(self._h, _lib.ptr(H["x_p"], f), _lib.ptr(H["v_p"], f),
                                               _lib.ptr(H["F_p"], f), _lib.ptr(H["C_p"], f),
                                               _lib.ptr(H["Jp_p"], f), _lib.ptr(H["fl_p"], i),
                                               _lib.ptr(H["g_p"], i)), self._h, "agimp_get_particles")
        self._device_newer = False

    def _invalidate(self):
        self._device_newer = True
        self._level_cache = None
        self._node_cache = None

    def _read_level(self, l):
        if self._level_cache is None:
            self._level_cache = {}
        if l not in self._level_cache:
            e, d = self.extents[l], self.dim
            cell = np.empty((e,) * d, np.int32)
            m = np.empty((e,) * d, np.float32)
            v = np.empty((e,) * d + (d,), np.float32)
            _lib.check(self._L.agimp_get_level(self._h, l, _lib.ptr(cell, ctypes.c_int32), _lib.ptr(m, ctypes.c_float),
                                               _lib.ptr(v, ctypes.c_float)), self._h, "agimp_get_level")
            self._level_cache[l] = {"cell": cell, "m": m, "v": v}
        return self._level_cache[l]

    def _read_node_mask(self):
        if self._node_cache is None:
            out = np.empty((self.extents[-1],) * self.dim, np.int32)
            _lib.check(self._L.agimp_get_node_mask(self._h, _lib.ptr(out, ctypes.c_int32)), self._h,
                       "agimp_get_node_mask")
            self._node_cache = out
        return self._node_cache

    # ------------------------------------------------------------------ simulation
    def refine(self):
        """Rebuild the octree from the current particles (adaptive_gimp.py:324-335)."""
        self._push()
        _lib.check(self._L.agimp_refine(self._h), self._h, "agimp_refine")
        self._level_cache = None
        self._node_cache = None

    def substep(self, dt0):
        """adaptive_gimp.py:323-352."""
        self._push()
        _lib.check(self._L.agimp_substep(self._h, float(dt0)), self._h, "agimp_substep")
        self._invalidate()

    def substep_fixed_topology(self, dt0):
        """One substep on the current octree (no rebuild)."""
        self._push()
        _lib.check(self._L.agimp_substep_fixed_topology(self._h, float(dt0)), self._h, "agimp_substep_fixed_topology")
        self._invalidate()

    def advance(self, dt0, n):
        self._push()
        _lib.check(self._L.agimp_advance(self._h, float(dt0), int(n)), self._h, "agimp_advance")
        self._invalidate()

    def set_particles_per_chunk(self, ch):
        """P2G work chunk size (the autotuned parameter)."""
        _lib.check(self._L.agimp_set_particles_per_chunk(self._h, int(ch)), self._h, "set_particles_per_chunk")

    def sync(self):
        _lib.check(self._L.agimp_sync(self._h), self._h, "agimp_sync")

    def set_timing(self, on=True):
        """Record HIP events around every phase of each substep (on the library's stream)."""
        _lib.check(self._L.agimp_set_timing(self._h, int(bool(on))), self._h, "agimp_set_timing")

    def timings(self):
        """{phase: (total_ms, calls)} since the last call (synchronises)."""
        ms = np.zeros(len(_lib.PHASES))
        calls = np.zeros(len(_lib.PHASES), np.int64)
        _lib.check(self._L.agimp_get_timings(self._h, _lib.ptr(ms, ctypes.c_double), _lib.ptr(calls, ctypes.c_int64),
                                             len(_lib.PHASES)), self._h, "agimp_get_timings")
        return {p: (float(ms[i]), int(calls[i])) for i, p in enumerate(_lib.PHASES)}

    def reset_counters(self):
        _lib.check(self._L.agimp_reset_counters(self._h), self._h, "agimp_reset_counters")

    def counters(self):
        out = np.zeros(8, np.int64)
        _lib.check(self._L.agimp_get_counters(self._h, _lib.ptr(out, ctypes.c_int64), 8), self._h, "get_counters")
        return {"oob": int(out[0]), "frozen": int(out[1]), "active": int(out[2]), "chunks": int(out[3]),
                "nonfinite": int(out[4]), "criterion_failures": int(out[5])}

    def kernel_timings(self):
        """{kernel: (total_ms, launches, min_ms, max_ms)} since the last call (synchronises);
        recorded while ``set_timing(True)`` is on, eager substeps only."""
        cap, nl = 128, 64
        names = ctypes.create_string_buffer(cap * nl)
        tot, mn, mx = np.zeros(cap), np.zeros(cap), np.zeros(cap)
        calls = np.zeros(cap, np.int64)
        cnt = ctypes.c_int32(0)
        dp = lambda a: _lib.ptr(a, ctypes.c_double)
        _lib.check(self._L.agimp_get_kernel_timings(self._h, names, nl, dp(tot), _lib.ptr(calls, ctypes.c_int64), dp(mn),
                                                    dp(mx), cap, ctypes.byref(cnt)), self._h, "agimp_get_kernel_timings")
        raw = names.raw
        out = {}
        for k in range(min(cnt.value, cap)):
            nm = raw[k * nl:(k + 1) * nl].split(b"\0", 1)[0].decode()
            out[nm] = (float(tot[k]), int(calls[k]), float(mn[k]), float(mx[k]))
        return out

    def profile_report(self):
        """GPU time per kernel and per phase since the last call, as text: the counterpart of
        the Taichi kernel profiler report that run_3d.py:12 and :43 print (share of the total,
        total, launches, min / avg / max per launch, kernel name).  Kernels and phases are timed
        with HIP events on the library's stream while ``set_timing(True)`` is on."""
        kt = self.kernel_timings()
        t = self.timings()
        ktot = sum(v[0] for v in kt.values())
        lines = [f"{'%':>6} {'total ms':>10} {'launches':>8} {'min ms':>9} {'avg ms':>9} {'max ms':>9}  kernel"]
        for name, (ms, calls, mn, mx) in sorted(kt.items(), key=lambda kv: -kv[1][0]):
            pct = 100.0 * ms / ktot if ktot > 0 else 0.0
            lines.append(f"{pct:6.2f} {ms:10.3f} {calls:8d} {mn:9.4f} {ms / calls:9.4f} {mx:9.4f}  {name}")
        lines.append(f"{100.0 if ktot > 0 else 0.0:6.2f} {ktot:10.3f} {'':>8} {'':>9} {'':>9} {'':>9}  [all kernels]")
        lines.append("")
        lines.append(f"{'phase':<10} {'calls':>7} {'total ms':>11} {'avg ms':>9}")
        for name, (ms, calls) in t.items():
            if calls:
                lines.append(f"{name:<10} {calls:>7d} {ms:>11.3f} {ms / calls:>9.4f}")
        tot = sum(ms for ms, _ in t.values())
        lines.append(f"{'total':<10} {'':>7} {tot:>11.3f}")
        return "\n".join(lines)

    # ------------------------------------------------------------------ snapshots (not in the reference)
    def save(self, path, **meta):
        """Checkpoint to an ``.npz`` file: the particle state (x, v, F, C, m, colour, g, J_p, in
        the caller's particle order) and the constructor settings, as plain arrays and JSON (no
        pickled objects).  ``meta`` (JSON-serialisable) is stored alongside, e.g. the time."""
        import json

        self._pull()
        H = self._host
        cfg = dict(self._kwargs, dim=self.dim, level=self.level, coarsest_size=self.coarsest_size,
                   n_particles=self.n_particles, static_adaptivity=bool(self.static_adaptivity))
        np.savez(path, x=H["x_p"], v=H["v_p"], F=H["F_p"], C=H["C_p"], m=H["m_p"], c=H["c_p"], g=H["g_p"],
                 Jp=H["Jp_p"], config=np.array(json.dumps(cfg)), meta=np.array(json.dumps(meta)))

    @classmethod
    def load(cls, path, grid_initializer=None, **overrides):
        """A simulator resumed from a checkpoint written by ``save`` (read with
        ``allow_pickle=False``).  A static-adaptivity run passes its ``grid_initializer`` again
        (a callable is not stored); ``overrides`` replace stored constructor settings.
        Returns ``(sim, meta)``."""
        import json

        with np.load(path, allow_pickle=False) as z:
            cfg = json.loads(str(z["config"]))
            meta = json.loads(str(z["meta"]))
            st = {k: z[k] for k in ("x", "v", "F", "C", "m", "c", "g", "Jp")}
        if cfg.pop("static_adaptivity") and grid_initializer is None:
            raise ValueError("checkpoint of a static-adaptivity run: pass its grid_initializer")
        dim, level, c, n = (cfg.pop(k) for k in ("dim", "level", "coarsest_size", "n_particles"))
        cfg.update(overrides)

        def init(sim):
            sim.x_p.from_numpy(st["x"])
            sim.v_p.from_numpy(st["v"])
            sim.F_p.from_numpy(st["F"])
            sim.C_p.from_numpy(st["C"])
            sim.m_p.from_numpy(st["m"])
            sim.c_p.from_numpy(st["c"])
            sim.g_p.from_numpy(st["g"])
            sim.Jp_p.from_numpy(st["Jp"])

        return cls(dim, level, c, n, init, grid_initializer, **cfg), meta

    # ------------------------------------------------------------------ learned prediction
    def learned_enable(self, window=20, tau=0.3):
        """Allocate the refinement-prediction state; switches the criterion check on."""
        _lib.check(self._L.agimp_learned_enable(self._h, int(window), float(tau)), self._h, "agimp_learned_enable")
        self._learn_window = int(window)

    def learned_set_model(self, params):
        p = np.ascontiguousarray(params, np.float32).ravel()
        if p.size != _lib.MLP_NPARAMS:
            raise ValueError(f"expected {_lib.MLP_NPARAMS} parameters, got {p.size}")
        _lib.check(self._L.agimp_learned_set_model(self._h, _lib.ptr(p, ctypes.c_float), p.size), self._h,
                   "agimp_learned_set_model")

    def learned_features(self, dt0):
        """(features [k, 16] float32, unit ids [k] int32) of the candidate units now."""
        self._push()
        cap = self._learn_cap()
        feat = np.empty((cap, 16), np.float32)
        units = np.empty(cap, np.int32)
        cnt = ctypes.c_int64(0)
        _lib.check(self._L.agimp_learned_features(self._h, float(dt0), _lib.ptr(feat, ctypes.c_float),
                                                  _lib.ptr(units, ctypes.c_int32), cap, ctypes.byref(cnt)),
                   self._h, "agimp_learned_features")
        k = min(cnt.value, cap)
        self._last_candidates = cnt.value
        return feat[:k].copy(), units[:k].copy()

    def learned_labels_reset(self):
        _lib.check(self._L.agimp_learned_labels_reset(self._h), self._h, "agimp_learned_labels_reset")

    def learned_labels_accumulate(self):
        self._push()
        _lib.check(self._L.agimp_learned_labels_accumulate(self._h), self._h, "agimp_learned_labels_accumulate")

    def learned_labels(self):
        """Window classes (0 finest, 1, 2, 3 none) of the candidates of the last features call."""
        k = self._last_candidates
        out = np.empty(max(k, 1), np.int32)
        _lib.check(self._L.agimp_learned_labels(self._h, _lib.ptr(out, ctypes.c_int32), k), self._h,
                   "agimp_learned_labels")
        return out[:k]

    def learned_predict(self, dt0):
        """(classes of the candidates, predicted level of every unit) on the current particles."""
        self._push()
        cap = self._learn_cap()
        cls = np.empty(cap, np.int32)
        lev = np.empty(cap, np.int8)
        cnt = ctypes.c_int64(0)
        _lib.check(self._L.agimp_learned_predict(self._h, float(dt0), _lib.ptr(cls, ctypes.c_int32),
                                                 _lib.ptr(lev, ctypes.c_int8), cap, ctypes.byref(cnt)),
                   self._h, "agimp_learned_predict")
        return cls[:min(cnt.value, cap)].copy(), lev

    def advance_learned(self, dt0, n):
        """n substeps reusing predicted octrees, with the plain criterion checked every substep."""
        self._push()
        _lib.check(self._L.agimp_learned_advance(self._h, float(dt0), int(n)), self._h, "agimp_learned_advance")
        self._invalidate()

    def learned_set_parts(self, network=True, sweep=True):
        """Parts of a predicted octree besides the current requests: the network's unit levels
        and/or each particle's requests swept along its velocity over the window."""
        _lib.check(self._L.agimp_learned_set_parts(self._h, int(bool(network)) | (2 * int(bool(sweep)))), self._h,
                   "agimp_learned_set_parts")

    def learned_stats(self):
        out = np.zeros(len(_lib.LSTATS), np.int64)
        _lib.check(self._L.agimp_learned_stats(self._h, _lib.ptr(out, ctypes.c_int64), len(out)), self._h,
                   "agimp_learned_stats")
        return {k: int(out[i]) for i, k in enumerate(_lib.LSTATS)}

    def learned_reset_stats(self):
        _lib.check(self._L.agimp_learned_reset_stats(self._h), self._h, "agimp_learned_reset_stats")

    def _learn_cap(self):
        ul = max(0, self.level - 3)
        nb = (self.extents[ul] + 8) // 4
        return nb ** self.dim * (16 if self.dim == 2 else 64)

    # ------------------------------------------------------------------ reference helpers
    def activate_cell(self, l, I):
        raise RuntimeError("activate_cell is only meaningful inside a grid_initializer call")

    def _map(self, l, I):  # adaptive_gimp.py:126-128
        return np.asarray(I) * (2 ** (self.level - 1 - l))

    def _nm(self, l, I):
        M = self._read_node_mask()
        f = self._map(l, I)
        return M[tuple(np.moveaxis(np.asarray(f), -1, 0))]

    def is_valid(self, l, I):        # :130-131
        return self._nm(l, I) > 0

    def is_activated(self, l, I):    # :133-134
        return (self._nm(l, I) & (ACTIVATED | T_JUNCTION)) == ACTIVATED

    def is_T_junction(self, l, I):   # :136-137
        return (self._nm(l, I) & T_JUNCTION) == T_JUNCTION

    def is_ghost(self, l, I):        # :139-140
        return (self._nm(l, I) & (ACTIVATED | GHOST | T_JUNCTION)) == GHOST


class Batch:
    """Many independent simulators of one dimension advanced together: each substep is one
    launch per phase for all scenes, captured once in a hipGraph and replayed.  The member
    simulators keep their state and API (read fields after ``advance``)."""

    def __init__(self, sims):
        self.sims = list(sims)
        if not self.sims:
            raise ValueError("empty batch")
        self._L = _lib.load()
        for s in self.sims:
            s._push()
        arr = (ctypes.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        b = ctypes.c_void_p()
        rc = self._L.agimp_batch_create(arr, len(self.sims), ctypes.byref(b))
        if rc != 0:
            raise _lib.AgimpError(f"agimp_batch_create failed with code {rc}")
        self._b = b

    def _push_members(self):
        # host-side edits of a member (from_numpy, item assignment) made after the batch was
        # created must reach the device before the batch advances the same buffers
        for s in self.sims:
            s._push()

    def set_particles_per_chunk(self, ch):
        self._push_members()
        ch = np.ascontiguousarray(np.broadcast_to(np.asarray(ch, np.int32), (len(self.sims),)))
        rc = self._L.agimp_batch_set_particles_per_chunk(self._b, _lib.ptr(ch, ctypes.c_int32))
        if rc != 0:
            raise _lib.AgimpError(f"agimp_batch_set_particles_per_chunk failed with code {rc}")

    def advance(self, dt, steps):
        self._push_members()
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (len(self.sims),)))
        rc = self._L.agimp_batch_advance(self._b, _lib.ptr(dt, ctypes.c_float), int(steps))
        if rc != 0:
            raw = self._L.agimp_batch_last_error(self._b)
            raise _lib.AgimpError(f"agimp_batch_advance failed with code {rc}: {raw.decode() if raw else ''}")
        for s in self.sims:
            s._invalidate()

    def close(self):
        b = getattr(self, "_b", None)
        if b:
            self._L.agimp_batch_destroy(b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
