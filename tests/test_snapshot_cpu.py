"""Checkpoint file format (host logic only): what ``AdaptiveGIMP.load`` rejects before any
device work."""
import json

import numpy as np
import pytest


def _write(path, static):
    cfg = dict(E=5e3, nu=0.2, gravity=None, p_mass=1.0, p_vol=1.0, boundary_gap=4, wall_mode="slip",
               plastic=False, theta_c=2.5e-2, theta_s=7.5e-3, hardening=0.0, slab=None, particles_per_chunk=1024,
               dim=2, level=2, coarsest_size=16, n_particles=4, static_adaptivity=static)
    z = np.zeros((4, 2), np.float32)
    np.savez(path, x=z, v=z, F=np.zeros((4, 2, 2), np.float32), C=np.zeros((4, 2, 2), np.float32),
             m=np.ones(4, np.float32), c=np.zeros((4, 3), np.float32), g=np.zeros(4, np.int32),
             Jp=np.ones(4, np.float32), config=np.array(json.dumps(cfg)), meta=np.array("{}"))


def test_static_checkpoint_needs_grid_initializer(pkg, tmp_path):
    p = tmp_path / "s.npz"
    _write(p, True)
    with pytest.raises(ValueError, match="grid_initializer"):
        pkg.AdaptiveGIMP.load(p)


def test_checkpoint_holds_no_pickles(tmp_path):
    p = tmp_path / "s.npz"
    _write(p, False)
    with np.load(p, allow_pickle=False) as z:  # every member loads without unpickling
        assert json.loads(str(z["config"]))["level"] == 2
        assert {k for k in z.files} == {"x", "v", "F", "C", "m", "c", "g", "Jp", "config", "meta"}
