"""TEST INFRASTRUCTURE ONLY -- ctypes wrapper of the oracle (oracle/agimp_oracle.cpp).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import this
module, and only as the checker (or as the timed CPU restatement).  The product path never
touches it.

The oracle restates one substep of g1n0st/adaptive-gimp (``adaptive_gimp.py:323-352``) on
the CPU, in float32 (the reference's field type) or float64 (the error oracle).  It keeps
the reference's state layout: particle arrays in original order, dense per-level
``cell_mask``/``ad_grid_m``/``ad_grid_v`` over the reference extents ``e_l``, and the
finest-resolution ``node_mask``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_agimp.so")
_lib = None


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("E", ctypes.c_double), ("nu", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("p_vol", ctypes.c_double), ("radius", ctypes.c_double),
        ("boundary_gap", ctypes.c_int), ("wall_mode", ctypes.c_int), ("plastic", ctypes.c_int),
        ("theta_c", ctypes.c_double), ("theta_s", ctypes.c_double), ("hardening", ctypes.c_double),
        ("slab_enable", ctypes.c_int), ("slab_axis", ctypes.c_int),
        ("slab_lo", ctypes.c_double), ("slab_hi", ctypes.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with the committed Makefile (g++ + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH) or (
            os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "agimp_oracle.cpp"))):
        args = ["make", "-s"] + (["-B"] if force else []) + ["liboracle_agimp.so"]
        subprocess.check_call(args, cwd=_HERE)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, dp, ip, i8p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32), \
            ctypes.POINTER(ctypes.c_int8)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                 ctypes.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_particles.argtypes = [vp, dp, dp, dp, dp, dp, dp, ip]
        L.orc_get_particles.argtypes = [vp, dp, dp, dp, dp, dp, ip]
        L.orc_set_static_map.argtypes = [vp, i8p, i8p]
        L.orc_refine.argtypes = [vp]
        L.orc_physics.argtypes = [vp, ctypes.c_double]
        L.orc_substep.argtypes = [vp, ctypes.c_double]
        L.orc_substep_fixed_topology.argtypes = [vp, ctypes.c_double]
        L.orc_get_level.argtypes = [vp, ctypes.c_int, ip, dp, dp]
        L.orc_get_node_mask.argtypes = [vp, ip]
        L.orc_get_oob.restype = ctypes.c_int64
        L.orc_get_oob.argtypes = [vp]
        L.orc_level_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.orc_weight_stencil.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp]
        L.orc_svd.argtypes = [ctypes.c_int, dp, dp, dp, dp]
        L.orc_stress.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, dp]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def align_size(x, a):  # utils.py:44-45
    return (x + (a - 1)) & ~(a - 1)


def extents(level, coarsest_size):
    """Reference per-level extents e_l (adaptive_gimp.py:36-37)."""
    return [4 * align_size((coarsest_size * 2 ** l) // 4 + 1, 4) for l in range(level)]


PAD = 4


class OracleSim:
    """CPU restatement with the same state layout as the reference simulator."""

    def __init__(self, dim, level, coarsest_size, n_particles, fp64=True, *, E=5e3, nu=0.2,
                 gravity=None, p_vol=1.0, radius=0.5, boundary_gap=4, wall_mode=0, plastic=0,
                 theta_c=2.5e-2, theta_s=7.5e-3, hardening=0.0, slab=None, threads=None):
        self.dim, self.level, self.coarsest_size, self.n = dim, level, coarsest_size, int(n_particles)
        self.fp64 = bool(fp64)
        g = [0.0, -100.0, 0.0] if gravity is None else list(gravity) + [0.0] * (3 - len(gravity))
        prm = OrcParams()
        prm.E, prm.nu = E, nu
        for i in range(3):
            prm.gravity[i] = g[i]
        prm.p_vol, prm.radius, prm.boundary_gap = p_vol, radius, boundary_gap
        prm.wall_mode, prm.plastic = wall_mode, plastic
        prm.theta_c, prm.theta_s, prm.hardening = theta_c, theta_s, hardening
        if slab is not None:
            prm.slab_enable, prm.slab_axis, prm.slab_lo, prm.slab_hi = 1, slab[0], slab[1], slab[2]
        self._lib = lib()
        if threads is not None:
            self._lib.orc_set_threads(int(threads))
        self._h = self._lib.orc_create(dim, level, coarsest_size, self.n, int(self.fp64), ctypes.byref(prm))
        self.e = extents(level, coarsest_size)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.orc_destroy(h)
            self._h = None

    def set_particles(self, x, v, F, C, m, g, Jp=None):
        n, d = self.n, self.dim
        f64 = lambda a, shape: np.ascontiguousarray(np.asarray(a, np.float64).reshape(shape))
        x, v = f64(x, (n, d)), f64(v, (n, d))
        F, C = f64(F, (n, d, d)), f64(C, (n, d, d))
        m = f64(m, (n,))
        Jp = np.ones(n) if Jp is None else f64(Jp, (n,))
        g = np.ascontiguousarray(np.asarray(g, np.int32).reshape(n))
        self._lib.orc_set_particles(self._h, _p(x, ctypes.c_double), _p(v, ctypes.c_double), _p(F, ctypes.c_double),
                                    _p(C, ctypes.c_double), _p(m, ctypes.c_double), _p(Jp, ctypes.c_double),
                                    _p(g, ctypes.c_int32))

    def set_static_map(self, ret_level, act_level):
        E = self.e[-1] + 2 * PAD
        r = np.ascontiguousarray(np.asarray(ret_level, np.int8).reshape((E,) * self.dim))
        a = np.ascontiguousarray(np.asarray(act_level, np.int8).reshape((E,) * self.dim))
        self._keep = (r, a)
        self._lib.orc_set_static_map(self._h, _p(r, ctypes.c_int8), _p(a, ctypes.c_int8))

    def refine(self):
        self._lib.orc_refine(self._h)

    def physics(self, dt):
        self._lib.orc_physics(self._h, float(dt))

    def substep(self, dt):
        self._lib.orc_substep(self._h, float(dt))

    def substep_fixed_topology(self, dt):
        self._lib.orc_substep_fixed_topology(self._h, float(dt))

    def get_particles(self):
        n, d = self.n, self.dim
        x, v = np.empty((n, d)), np.empty((n, d))
        F, C = np.empty((n, d, d)), np.empty((n, d, d))
        Jp, fl = np.empty(n), np.empty(n, np.int32)
        self._lib.orc_get_particles(self._h, _p(x, ctypes.c_double), _p(v, ctypes.c_double), _p(F, ctypes.c_double),
                                    _p(C, ctypes.c_double), _p(Jp, ctypes.c_double), _p(fl, ctypes.c_int32))
        return dict(x=x, v=v, F=F, C=C, Jp=Jp, fl=fl)

    def get_level(self, l):
        el, d = self.e[l], self.dim
        cell = np.empty((el,) * d, np.int32)
        m = np.empty((el,) * d)
        v = np.empty((el,) * d + (d,))
        self._lib.orc_get_level(self._h, l, _p(cell, ctypes.c_int32), _p(m, ctypes.c_double), _p(v, ctypes.c_double))
        return cell, m, v

    def get_node_mask(self):
        el = self.e[-1]
        out = np.empty((el,) * self.dim, np.int32)
        self._lib.orc_get_node_mask(self._h, _p(out, ctypes.c_int32))
        return out

    def oob(self):
        return int(self._lib.orc_get_oob(self._h))

    def level_counts(self):
        out = np.zeros(self.level, np.int64)
        self._lib.orc_level_counts(self._h, _p(out, ctypes.c_int64))
        return out


def weight_stencil(dim, t, radius=0.5):
    """utils.py:18-42 + adaptive_gimp.py:218-228: returns W[3^d, d+1] (weight, grad)."""
    t = np.ascontiguousarray(np.asarray(t, np.float64).reshape(dim))
    W = np.empty((3 ** dim, dim + 1))
    lib().orc_weight_stencil(dim, radius, _p(t, ctypes.c_double), _p(W, ctypes.c_double))
    return W


def svd(F):
    F = np.ascontiguousarray(np.asarray(F, np.float64))
    d = F.shape[0]
    U, s, V = np.empty((d, d)), np.empty(d), np.empty((d, d))
    lib().orc_svd(d, _p(F, ctypes.c_double), _p(U, ctypes.c_double), _p(s, ctypes.c_double), _p(V, ctypes.c_double))
    return U, s, V


def stress(F, mu, la):
    F = np.ascontiguousarray(np.asarray(F, np.float64))
    d = F.shape[0]
    tau = np.empty((d, d))
    lib().orc_stress(d, _p(F, ctypes.c_double), mu, la, _p(tau, ctypes.c_double))
    return tau


def canonical_node_mask(M):
    """Drop G where A or T is set: the reference's raw bits can be T or T|G for the same node
    (the ghost check at adaptive_gimp.py:182-183 races the T OR at :174); every predicate
    (adaptive_gimp.py:130-140) treats them alike."""
    M = np.asarray(M).astype(np.int32)
    return np.where((M & 5) != 0, M & ~2, M)
