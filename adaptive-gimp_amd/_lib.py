"""ctypes binding of libagimp.so (include/agimp.h).

The library is the only compute path: if it is missing this module raises, there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libagimp.so")

PHASES = ["refine", "bucket", "p2g", "grid", "g2p"]  # AGIMP_PHASE_* order

AGIMP_OK, AGIMP_EINVAL, AGIMP_EHIP, AGIMP_ESTATE, AGIMP_ENOMEM = 0, -1, -2, -3, -4


class AgimpConfig(ctypes.Structure):
    """Mirror of ``agimp_config`` (include/agimp.h)."""
    _fields_ = [
        ("dim", ctypes.c_int32), ("level", ctypes.c_int32), ("coarsest_size", ctypes.c_int32),
        ("boundary_gap", ctypes.c_int32),
        ("n_particles", ctypes.c_int64),
        ("E", ctypes.c_float), ("nu", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("p_vol", ctypes.c_float), ("radius", ctypes.c_float),
        ("wall_mode", ctypes.c_int32), ("plastic", ctypes.c_int32),
        ("theta_c", ctypes.c_float), ("theta_s", ctypes.c_float), ("hardening", ctypes.c_float),
        ("slab_enable", ctypes.c_int32), ("slab_axis", ctypes.c_int32),
        ("slab_lo", ctypes.c_float), ("slab_hi", ctypes.c_float),
        ("particles_per_chunk", ctypes.c_int32), ("device", ctypes.c_int32),
    ]


EXPORTS = [
    "agimp_config_init", "agimp_create", "agimp_destroy", "agimp_set_particles", "agimp_set_static_levels",
    "agimp_refine", "agimp_substep", "agimp_substep_fixed_topology", "agimp_advance", "agimp_get_particles",
    "agimp_get_particles_noalloc",
    "agimp_get_level", "agimp_get_node_mask", "agimp_get_counters", "agimp_reset_counters",
    "agimp_level_extent", "agimp_set_timing", "agimp_get_timings", "agimp_get_kernel_timings", "agimp_sync",
    "agimp_last_error",
    "agimp_set_particles_per_chunk", "agimp_batch_create", "agimp_batch_destroy",
    "agimp_batch_set_particles_per_chunk", "agimp_batch_advance", "agimp_batch_last_error",
    "agimp_host_svd", "agimp_host_constitutive",
    "agimp_learned_enable", "agimp_learned_set_model", "agimp_learned_features", "agimp_learned_labels_reset",
    "agimp_learned_labels_accumulate", "agimp_learned_labels", "agimp_learned_predict", "agimp_learned_advance",
    "agimp_learned_stats", "agimp_learned_reset_stats", "agimp_learned_set_parts", "agimp_mlp_forward", "agimp_mlp_loss_grad", "agimp_mlp_adam",
    "agimp_trainer_create", "agimp_trainer_destroy", "agimp_trainer_set_data", "agimp_trainer_set_params",
    "agimp_trainer_get_params", "agimp_trainer_epoch", "agimp_trainer_logits", "agimp_trainer_grad",
    "agimp_trainer_apply",
]
MLP_NPARAMS = 5508
LSTATS = ["rebuilds", "forced", "discarded", "replayed", "substeps", "violations", "candidates", "refined"]
MIN_CHUNK, MAX_CHUNK = 64, 2048

_lib = None


class AgimpError(RuntimeError):
    pass


def load():
    """Load libagimp.so (built by ``adaptive-gimp_amd/build.py``); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python adaptive-gimp_amd/build.py` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int32)
    i8p = ctypes.POINTER(ctypes.c_int8)
    dp = ctypes.POINTER(ctypes.c_double)
    i64p = ctypes.POINTER(ctypes.c_int64)
    cfgp = ctypes.POINTER(AgimpConfig)
    sig = {
        "agimp_config_init": [cfgp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64],
        "agimp_create": [cfgp, ctypes.POINTER(vp)],
        "agimp_destroy": [vp],
        "agimp_set_particles": [vp, fp, fp, fp, fp, fp, ip, fp],
        "agimp_set_static_levels": [vp, i8p, i8p],
        "agimp_refine": [vp],
        "agimp_substep": [vp, ctypes.c_float],
        "agimp_substep_fixed_topology": [vp, ctypes.c_float],
        "agimp_advance": [vp, ctypes.c_float, ctypes.c_int32],
        "agimp_get_particles": [vp, fp, fp, fp, fp, fp, ip, ip],
        "agimp_get_particles_noalloc": [vp, fp, fp, fp, fp, fp, ip, ip],
        "agimp_get_level": [vp, ctypes.c_int32, ip, fp, fp],
        "agimp_get_node_mask": [vp, ip],
        "agimp_get_counters": [vp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32],
        "agimp_reset_counters": [vp],
        "agimp_level_extent": [vp, ctypes.c_int32],
        "agimp_set_timing": [vp, ctypes.c_int32],
        "agimp_get_timings": [vp, dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32],
        "agimp_get_kernel_timings": [vp, ctypes.c_char_p, ctypes.c_int32, dp, ctypes.POINTER(ctypes.c_int64), dp, dp,
                                     ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
        "agimp_sync": [vp],
        "agimp_last_error": [vp],
        "agimp_set_particles_per_chunk": [vp, ctypes.c_int32],
        "agimp_batch_create": [ctypes.POINTER(vp), ctypes.c_int32, ctypes.POINTER(vp)],
        "agimp_batch_destroy": [vp],
        "agimp_batch_set_particles_per_chunk": [vp, ip],
        "agimp_batch_advance": [vp, fp, ctypes.c_int32],
        "agimp_batch_last_error": [vp],
        "agimp_host_svd": [ctypes.c_int32, dp, dp, dp, dp],
        "agimp_host_constitutive": [cfgp, dp, dp, dp, ctypes.c_int32],
        "agimp_learned_enable": [vp, ctypes.c_int32, ctypes.c_float],
        "agimp_learned_set_model": [vp, fp, ctypes.c_int32],
        "agimp_learned_features": [vp, ctypes.c_float, fp, ip, ctypes.c_int64, i64p],
        "agimp_learned_labels_reset": [vp],
        "agimp_learned_labels_accumulate": [vp],
        "agimp_learned_labels": [vp, ip, ctypes.c_int64],
        "agimp_learned_predict": [vp, ctypes.c_float, ip, i8p, ctypes.c_int64, i64p],
        "agimp_learned_advance": [vp, ctypes.c_float, ctypes.c_int32],
        "agimp_learned_stats": [vp, i64p, ctypes.c_int32],
        "agimp_learned_reset_stats": [vp],
        "agimp_learned_set_parts": [vp, ctypes.c_int32],
        "agimp_mlp_forward": [vp, ctypes.c_int64, vp, vp, vp],
        "agimp_mlp_loss_grad": [vp, vp, ctypes.c_int64, vp, vp, ctypes.c_float, vp, vp, vp],
        "agimp_mlp_adam": [vp, vp, vp, vp, ctypes.c_int32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                           ctypes.c_float, ctypes.c_float, vp],
        "agimp_trainer_create": [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(vp)],
        "agimp_trainer_destroy": [vp],
        "agimp_trainer_set_data": [vp, fp, ip, ctypes.c_int64],
        "agimp_trainer_set_params": [vp, fp],
        "agimp_trainer_get_params": [vp, fp],
        "agimp_trainer_epoch": [vp, ip, ctypes.c_int64, fp, ctypes.c_float, ctypes.c_float, ctypes.c_float, dp],
        "agimp_trainer_logits": [vp, fp, ctypes.c_int64, fp],
        "agimp_trainer_grad": [vp, fp, ip, ctypes.c_int64, fp, ctypes.c_float, fp, fp],
        "agimp_trainer_apply": [vp, fp, ctypes.c_float, ctypes.c_float],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_char_p if name.endswith("last_error") else ctypes.c_int
    _lib = L
    return L


def check(rc, handle=None, what=""):
    if rc != 0:
        msg = ""
        if handle:
            raw = load().agimp_last_error(handle)
            msg = raw.decode() if raw else ""
        raise AgimpError(f"{what} failed with code {rc}: {msg}")
    return rc


def ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))
