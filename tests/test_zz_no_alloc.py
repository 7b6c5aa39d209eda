"""agimp_get_particles_noalloc (the allocation-free variant of agimp_get_particles, staging
through the library's file-scope scratch array) against agimp_get_particles, through the
C-ABI: bit-identical outputs on two calls in a row, the second on a smaller scene (the array
keeps the first call's contents), and a clear error for a scene above the array's size."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _staging_bytes(n, d):
    return (2 * n * d + 2 * n * d * d + n) * 4 + 2 * n * 4


def _read(sim, fn_name):
    L = sim._L
    n, d = sim.n_particles, sim.dim
    out = dict(x=np.full((n, d), np.nan, np.float32), v=np.full((n, d), np.nan, np.float32),
               F=np.full((n, d, d), np.nan, np.float32), C=np.full((n, d, d), np.nan, np.float32),
               Jp=np.full(n, np.nan, np.float32), fl=np.full(n, -7, np.int32), g=np.full(n, -7, np.int32))
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    rc = getattr(L, fn_name)(sim._h, fp(out["x"]), fp(out["v"]), fp(out["F"]), fp(out["C"]), fp(out["Jp"]),
                             ip(out["fl"]), ip(out["g"]))
    return rc, out


def _scene(pkg, dim, n, seed):
    if dim == 3:
        wl = pkg.scenes.small_3d(seed=seed, n=n, L=3, c=16)
    else:
        wl = pkg.scenes.c1_two_squares(seed=seed, scale=n / 21000.0, rho=1000.0)
    sim = pkg.AdaptiveGIMP(dim, wl.level, wl.coarsest_size, wl.n, wl.initializer(), **wl.params)
    return sim, wl


def test_noalloc_download_equals_allocating_download(pkg):
    first, wl1 = _scene(pkg, 3, 20000, 1)
    second, wl2 = _scene(pkg, 2, 5000, 2)
    assert second.n_particles < first.n_particles
    for sim, wl in ((first, wl1), (second, wl2)):
        sim.advance(wl.dt, 3)  # particles re-sorted on the device: the download un-permutes them
        rc_ref, ref = _read(sim, "agimp_get_particles")
        rc, got = _read(sim, "agimp_get_particles_noalloc")
        assert rc_ref == 0 and rc == 0, sim._L.agimp_last_error(sim._h)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), k
        assert np.isfinite(got["x"]).all() and (got["fl"] >= 0).all()
    # a second call on the first scene after the smaller one: nothing of the smaller call remains
    rc, again = _read(first, "agimp_get_particles_noalloc")
    _, ref = _read(first, "agimp_get_particles")
    assert rc == 0 and all(np.array_equal(again[k], ref[k]) for k in ref)
    first.close()
    second.close()


def test_noalloc_rejects_scene_above_scratch(pkg):
    n = SCRATCH_BYTES // _staging_bytes(1, 3) + 1
    x = np.random.default_rng(0).uniform(0.3, 0.7, (n, 3))

    def init(sim):
        sim.x_p.from_numpy(x)
        sim.F_p.from_numpy(np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)))
        sim.m_p.from_numpy(np.ones(n))

    sim = pkg.AdaptiveGIMP(3, 2, 16, n, init)
    assert _staging_bytes(n, 3) > SCRATCH_BYTES >= _staging_bytes(n - 1, 3)
    rc, out = _read(sim, "agimp_get_particles_noalloc")
    assert rc == -1  # AGIMP_EINVAL
    assert b"scratch" in sim._L.agimp_last_error(sim._h)
    assert np.isnan(out["x"]).all()  # nothing was written
    sim.close()
