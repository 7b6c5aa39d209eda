"""CPU restatement of the learned prediction's unit features, window labels and candidate
set (TEST INFRASTRUCTURE: only tests/ use it, as the checker of the library's kernels).

These quantities have no counterpart in the reference: they are this library's own inputs
to its refinement-prediction network (include/agimp.h, agimp_learned_*).  The one piece taken
from the reference is the plain criterion that the labels record: a particle requests its
2^d cells b + {0,1}^d at level g_p, b = int(x/dx + 0.5) - 1 (adaptive_gimp.py:87-96).

Units are the cells of level ulev = max(0, L-3) in the padded block-major layout of that
level (padding of 4 cells, 4^d-cell leaf blocks).
"""
from __future__ import annotations

import itertools

import numpy as np

PAD = 4


def extents(level, coarsest):
    """e_l = 4*align_size(n_l/4+1, 4) (adaptive_gimp.py:36-37)."""
    out = []
    for l in range(level):
        n = coarsest << l
        a = n // 4 + 1
        out.append(4 * ((a + 3) & ~3))
    return out


def unit_level(L):
    return max(0, L - 3)


def bidx(J, NB, d):
    """Block-major index of padded coordinates J [..., d]."""
    J = np.asarray(J, np.int64)
    B = J >> 2
    w = J & 3
    if d == 2:
        b = B[..., 0] * NB + B[..., 1]
        return (b << 4) | (w[..., 0] << 2) | w[..., 1]
    b = (B[..., 0] * NB + B[..., 1]) * NB + B[..., 2]
    return (b << 6) | (w[..., 0] << 4) | (w[..., 1] << 2) | w[..., 2]


def coords(q, NB, d):
    """Padded coordinates of block-major indices q."""
    q = np.asarray(q, np.int64)
    p4 = 16 if d == 2 else 64
    b, w = q // p4, q % p4
    if d == 2:
        B = np.stack([b // NB, b % NB], -1)
        W = np.stack([w >> 2, w & 3], -1)
    else:
        B = np.stack([b // (NB * NB), (b // NB) % NB, b % NB], -1)
        W = np.stack([w >> 4, (w >> 2) & 3, w & 3], -1)
    return B * 4 + W


def lev_code(lp1, L):
    lp1 = np.asarray(lp1)
    return np.where(lp1 <= 0, 0.0, 0.25 * (1 + np.maximum(0, 3 - (L - lp1))))


def level_dx(c, l):
    return np.float32(np.float32(1.0 / c) / np.float32(1 << l))


def base_cells(x, c, l):
    """int(x/dx + 0.5) - 1 in float32 (power-of-two coarsest size: x * (c 2^l) is exact)."""
    q = np.asarray(x, np.float32) * np.float32(c << l)
    return np.trunc(q + np.float32(0.5)).astype(np.int64) - 1


class Units:
    def __init__(self, dim, L, c):
        self.d, self.L, self.c = dim, L, c
        self.ulev = unit_level(L)
        self.E = extents(L, c)[self.ulev] + 2 * PAD
        self.NB = self.E // 4
        self.nu = self.NB ** dim * (16 if dim == 2 else 64)

    def unit_of(self, x):
        I = np.floor(np.asarray(x, np.float32) * np.float32(self.c << self.ulev)).astype(np.int64)
        J = I + PAD
        ok = np.all((J >= 0) & (J < self.E), axis=1)
        q = np.full(len(x), -1, np.int64)
        q[ok] = bidx(J[ok], self.NB, self.d)
        return q

    def requests(self, x, g, only=None, near=None):
        """Max (level + 1) requested per unit (dense [nu] int), levels >= ulev; with `near`
        (unit coordinates per particle) only units within one unit of it are recorded."""
        req = np.zeros(self.nu, np.int64)
        sel = (g >= self.ulev) & (g < self.L)
        if only is not None:
            sel &= only
        for lev in range(self.ulev, self.L):
            s = sel & (g == lev)
            if not s.any():
                continue
            b = base_cells(x[s], self.c, lev)
            for dl in itertools.product((0, 1), repeat=self.d):
                Iu = (b + np.asarray(dl)) >> (lev - self.ulev)
                J = Iu + PAD
                ok = np.all((J >= 0) & (J < self.E), axis=1)
                if near is not None:
                    ok &= np.all(np.abs(Iu - near[s]) <= 1, axis=1)
                np.maximum.at(req, bidx(J[ok], self.NB, self.d), lev + 1)
        return req

    def unit_coords_of(self, x):
        return np.floor(np.asarray(x, np.float32) * np.float32(self.c << self.ulev)).astype(np.int64)

    @staticmethod
    def fma32(a, b, c):
        """float32 fma(a, b, c) with a single rounding (extended precision holds the exact
        value for the magnitudes here)."""
        r = np.asarray(a, np.float32).astype(np.longdouble) * np.asarray(b, np.float32).astype(np.longdouble) \
            + np.asarray(c, np.float32).astype(np.longdouble)
        return r.astype(np.float32)

    def candidates(self, occupied):
        """Units within one unit (3^d neighbourhood, in the box) of an occupied unit."""
        J = coords(np.flatnonzero(occupied), self.NB, self.d)
        cand = np.zeros(self.nu, bool)
        for dl in itertools.product((-1, 0, 1), repeat=self.d):
            K = J + np.asarray(dl)
            ok = np.all((K >= 0) & (K < self.E), axis=1)
            cand[bidx(K[ok], self.NB, self.d)] = True
        return np.flatnonzero(cand)

    def features(self, x, v, F, C, m, g, p_vol, window, dt):
        """{unit id: 16 features} of the candidate units (float64), as the library defines them."""
        d, L = self.d, self.L
        x = np.asarray(x, np.float32)
        v = np.asarray(v, np.float32)
        q = self.unit_of(x)
        ok = q >= 0
        nu = self.nu
        F = np.asarray(F, np.float64)
        J = np.linalg.det(F)
        cnt = np.bincount(q[ok], minlength=nu).astype(np.float64)
        mass = np.bincount(q[ok], weights=np.asarray(m, np.float64)[ok], minlength=nu)
        vn = np.linalg.norm(v.astype(np.float64), axis=1)
        cn = np.sqrt(np.sum(np.asarray(C, np.float64).reshape(len(x), -1) ** 2, axis=1))
        vsum = np.bincount(q[ok], weights=vn[ok], minlength=nu)
        csum = np.bincount(q[ok], weights=cn[ok], minlength=nu)
        jsum = np.bincount(q[ok], weights=(J - 1)[ok], minlength=nu)
        gsum = np.bincount(q[ok], weights=g[ok].astype(np.float64), minlength=nu)
        vb = np.stack([np.bincount(q[ok], weights=v[ok, a].astype(np.float64), minlength=nu) for a in range(d)], 1)
        vmax = np.zeros(nu)
        np.maximum.at(vmax, q[ok], vn[ok])
        jmin = np.full(nu, np.inf)
        np.minimum.at(jmin, q[ok], J[ok])
        gmax = np.zeros(nu, np.int64)
        np.maximum.at(gmax, q[ok], g[ok] + 1)
        req = self.requests(x, g, only=ok)
        win32 = np.float32(np.float32(window) * np.float32(dt))
        xf = self.fma32(v, win32, x)
        reqf = self.requests(xf, g, only=ok, near=self.unit_coords_of(x))
        occ = cnt > 0
        cand = self.candidates(occ)
        dxu = np.float64(level_dx(self.c, self.ulev))
        mean_mass = float(np.sum(np.asarray(m, np.float64))) / len(m)
        inv_mref = p_vol / (mean_mass * dxu ** d)
        win = float(np.float32(window) * np.float32(dt))
        disp = win * float(self.c << self.ulev)
        out = {}
        Jc = coords(cand, self.NB, d)
        for k, u in enumerate(cand):
            o = occ[u]
            ic = 1.0 / cnt[u] if o else 0.0
            f = np.zeros(16)
            f[0] = mass[u] * inv_mref
            f[1] = vsum[u] * ic * disp
            f[2] = vmax[u] * disp
            f[3] = csum[u] * ic * win
            f[4] = jsum[u] * ic * 10
            f[5] = (jmin[u] - 1) * 10 if o else 0.0
            f[6] = lev_code(gmax[u], L)
            f[7] = lev_code(reqf[u], L)
            f[8] = lev_code(req[u], L)
            f[12:12 + d] = vb[u] * ic * disp
            nm, ng, nr, up = 0.0, 0.0, 0.0, 0.0
            for dl in itertools.product((-1, 0, 1), repeat=d):
                if not any(dl):
                    continue
                K = Jc[k] + np.asarray(dl)
                if np.any(K < 0) or np.any(K >= self.E):
                    continue
                w = int(bidx(K, self.NB, d))
                nm += mass[w]
                gc = float(lev_code(gmax[w], L))
                ng = max(ng, gc)
                nr = max(nr, float(lev_code(req[w], L)))
                if cnt[w] > 0 and -np.dot(vb[w], np.asarray(dl, np.float64)) > 0:
                    up = max(up, gc)
            f[9] = nm * inv_mref / (3 ** d - 1)
            f[10], f[11], f[15] = ng, nr, up
            out[int(u)] = f
        return out

    def window_classes(self, xs, g, cand):
        """Class of each candidate from the particle positions at every substep of a window."""
        lab = np.zeros(self.nu, np.int64)
        for x in xs:
            lab = np.maximum(lab, self.requests(np.asarray(x, np.float32), np.asarray(g)))
        r = lab[np.asarray(cand)]
        return np.where(r <= 0, 3, np.minimum(3, self.L - r))
