"""CPU checks of the C-ABI library: it loads, exports every symbol include/agimp.h declares,
and its FP64 host maths (the same functions the kernels run) matches numpy / the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest
from scipy.linalg import polar

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "agimp.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(agimp_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol(pkg):
    L = pkg._lib.load()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(pkg._lib.EXPORTS)


def test_create_without_gpu_fails_cleanly(pkg):
    """Bad configurations are rejected before any device call; nothing aborts."""
    L = pkg._lib.load()
    cfg = pkg._lib.AgimpConfig()
    assert L.agimp_config_init(ctypes.byref(cfg), 3, 3, 64, 1000) == 0
    assert cfg.E == pytest.approx(5e3) and cfg.gravity[1] == pytest.approx(-100.0) and cfg.boundary_gap == 4
    h = ctypes.c_void_p()
    bad = pkg._lib.AgimpConfig()
    L.agimp_config_init(ctypes.byref(bad), 4, 3, 64, 1000)
    assert L.agimp_create(ctypes.byref(bad), ctypes.byref(h)) == pkg._lib.AGIMP_EINVAL
    L.agimp_config_init(ctypes.byref(bad), 3, 7, 64, 1000)
    assert L.agimp_create(ctypes.byref(bad), ctypes.byref(h)) == pkg._lib.AGIMP_EINVAL
    L.agimp_config_init(ctypes.byref(bad), 3, 3, 64, 1000)
    bad.radius = 0.7
    assert L.agimp_create(ctypes.byref(bad), ctypes.byref(h)) == pkg._lib.AGIMP_EINVAL
    assert L.agimp_refine(None) == pkg._lib.AGIMP_EINVAL


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("dim", [2, 3])
def test_kernel_svd_against_numpy(pkg, dim):
    L = pkg._lib.load()
    rng = np.random.default_rng(10)
    for k in range(2000):
        F = np.ascontiguousarray(np.eye(dim) + (0.01 if k % 2 else 0.6) * rng.standard_normal((dim, dim)))
        if k % 5 == 0:
            F[0] = -F[0]
        if k % 97 == 0:
            F = np.eye(dim)  # repeated singular values
        U, s, V = np.zeros((dim, dim)), np.zeros(dim), np.zeros((dim, dim))
        assert L.agimp_host_svd(dim, _p(F), _p(U), _p(s), _p(V)) == 0
        # 3D: the Jacobi sweeps stop once the off-diagonal part of F^T F is 1e-10 of its
        # diagonal (the kernel's outputs are FP32), so F is reconstructed to ~3e-10; U and V
        # stay orthogonal and the singular values exact to FP64 rounding
        np.testing.assert_allclose(U @ np.diag(s) @ V.T, F, atol=2e-14 if dim == 2 else 1e-9)
        np.testing.assert_allclose(U @ U.T, np.eye(dim), atol=2e-14)
        np.testing.assert_allclose(V @ V.T, np.eye(dim), atol=2e-14)
        assert np.linalg.det(U) == pytest.approx(1, abs=1e-13) and np.linalg.det(V) == pytest.approx(1, abs=1e-13)
        np.testing.assert_allclose(np.sort(np.abs(s))[::-1], np.linalg.svd(F)[1], atol=2e-14)


def _cfg(pkg, dim, **kw):
    cfg = pkg._lib.AgimpConfig()
    pkg._lib.load().agimp_config_init(ctypes.byref(cfg), dim, 2, 16, 1)
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


@pytest.mark.parametrize("dim", [2, 3])
def test_kernel_stress_matches_oracle_and_polar(pkg, O, dim):
    L = pkg._lib.load()
    cfg = _cfg(pkg, dim, E=1.4e5, nu=0.2)
    E, nu = float(np.float32(1.4e5)), float(np.float32(0.2))  # the config holds float32
    mu = E / (2 * (1 + nu))
    la = E * nu / ((1 + nu) * (1 - 2 * nu))
    rng = np.random.default_rng(11)
    for _ in range(300):
        F = np.ascontiguousarray(np.eye(dim) + 0.1 * rng.standard_normal((dim, dim)))
        Jp = np.array([1.0])
        tau = np.zeros((dim, dim))
        F2 = F.copy()
        assert L.agimp_host_constitutive(ctypes.byref(cfg), _p(F2), _p(Jp), _p(tau), 0) == 0
        np.testing.assert_array_equal(F2, F)
        # 3D: the kernel's Jacobi sweeps stop at an off-diagonal residual of 1e-10 (its outputs
        # are FP32, rounded at 6e-8), which the polar rotation carries into tau at up to 5e-10
        # of tau's largest entry; 2D (closed form) stays at FP64 rounding
        ref = O.stress(F, mu, la)
        tol = (1e-13 if dim == 2 else 2e-9) * np.abs(ref).max() + 1e-7
        np.testing.assert_allclose(tau, ref, rtol=0, atol=tol)
        R, _ = polar(F)
        J = np.linalg.det(F)
        np.testing.assert_allclose(tau, 2 * mu * (F - R) @ F.T + la * J * (J - 1) * np.eye(dim), rtol=0, atol=tol)


def test_plasticity_clamp_closed_form(pkg):
    """Stomakhin et al. 2013: singular values clamped to [1-theta_c, 1+theta_s], Jp scaled by
    sigma/sigma_clamped, mu and lambda by exp(xi (1 - Jp))."""
    L = pkg._lib.load()
    tc, ts, xi = 2.5e-2, 7.5e-3, 10.0
    cfg = _cfg(pkg, 3, E=1.4e5, nu=0.2, plastic=1, theta_c=tc, theta_s=ts, hardening=xi)
    c, s_ = np.cos(0.3), np.sin(0.3)
    Rz = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]])
    sig = np.array([1.1, 0.9, 1.001])
    F = np.ascontiguousarray(Rz @ np.diag(sig))
    Jp = np.array([0.98])
    tau = np.zeros((3, 3))
    assert L.agimp_host_constitutive(ctypes.byref(cfg), _p(F), _p(Jp), _p(tau), 1) == 0
    sc = np.clip(sig, 1 - float(np.float32(tc)), 1 + float(np.float32(ts)))  # bounds in double
    np.testing.assert_allclose(F, Rz @ np.diag(sc), atol=1e-12)
    jp = 0.98 * np.prod(sig / sc)
    assert Jp[0] == pytest.approx(jp, rel=1e-12)
    h = np.exp(float(np.float32(xi)) * (1 - jp))
    E, nu = float(np.float32(1.4e5)), float(np.float32(0.2))
    mu = E / (2 * (1 + nu)) * h
    la = E * nu / ((1 + nu) * (1 - 2 * nu)) * h
    J = np.prod(sc)
    ref = 2 * mu * (F - Rz) @ F.T + la * J * (J - 1) * np.eye(3)
    np.testing.assert_allclose(tau, ref, rtol=1e-6, atol=1e-6)
