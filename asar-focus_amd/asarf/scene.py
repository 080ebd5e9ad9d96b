"""ctypes binding of the seeded synthetic raw-scene generator (lib/libasarf_scene.so)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .configs import SceneConfig

_LIB = None
_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libasarf_scene.so")


class _SceneParams(ctypes.Structure):
    _fields_ = [
        ("na", ctypes.c_int), ("nr", ctypes.c_int), ("nc", ctypes.c_int), ("format", ctypes.c_int),
        ("fs", ctypes.c_double), ("kr", ctypes.c_double), ("prf", ctypes.c_double), ("wavelength", ctypes.c_double),
        ("vr", ctypes.c_double), ("r_first", ctypes.c_double), ("dc_near", ctypes.c_double),
        ("dc_far", ctypes.c_double), ("az_band_frac", ctypes.c_double),
        ("n_targets_az", ctypes.c_int), ("n_targets_rg", ctypes.c_int), ("n_targets_random", ctypes.c_int),
        ("tgt_rg_lo", ctypes.c_double), ("tgt_rg_hi", ctypes.c_double), ("tgt_az_lo", ctypes.c_double),
        ("tgt_az_hi", ctypes.c_double), ("snr_db", ctypes.c_double), ("clutter_power", ctypes.c_double),
        ("clutter_shape", ctypes.c_double), ("clutter_rho", ctypes.c_double), ("quant_std_lsb", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        lib.asarf_scene_generate.argtypes = [ctypes.POINTER(_SceneParams), ctypes.c_void_p, ctypes.c_void_p]
        lib.asarf_scene_generate.restype = ctypes.c_int
        lib.asarf_scene_targets.argtypes = [ctypes.POINTER(_SceneParams), ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int]
        lib.asarf_scene_targets.restype = ctypes.c_int
        lib.asarf_philox4x32_10.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.asarf_philox4x32_10.restype = None
        _LIB = lib
    return _LIB


def _params(cfg: SceneConfig) -> _SceneParams:
    p = _SceneParams()
    p.na, p.nr, p.nc = cfg.na, cfg.nr, cfg.nc
    p.format = 0 if cfg.raw_format == "u8_ers" else 1
    p.fs, p.kr, p.prf, p.wavelength = cfg.fs, cfg.kr, cfg.prf, cfg.wavelength
    p.vr, p.r_first, p.dc_near, p.dc_far = cfg.vr, cfg.r_first, cfg.dc_near, cfg.dc_far
    p.az_band_frac = cfg.az_band_frac
    p.n_targets_az, p.n_targets_rg, p.n_targets_random = cfg.n_targets_az, cfg.n_targets_rg, cfg.n_targets_random
    p.tgt_rg_lo, p.tgt_rg_hi = cfg.tgt_rg
    p.tgt_az_lo, p.tgt_az_hi = cfg.tgt_az
    p.snr_db, p.clutter_power, p.clutter_shape = cfg.snr_db, cfg.clutter_power, cfg.clutter_shape
    p.clutter_rho = cfg.clutter_rho
    p.quant_std_lsb = cfg.quant_std_lsb
    p.seed = cfg.seed
    return p


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    _lib().asarf_philox4x32_10(c, k, o)
    return list(o)


def targets(cfg: SceneConfig):
    p = _params(cfg)
    n = _lib().asarf_scene_targets(ctypes.byref(p), None, None, 0)
    tx = np.zeros(max(n, 1), np.float64)
    ty = np.zeros(max(n, 1), np.float64)
    _lib().asarf_scene_targets(ctypes.byref(p), tx.ctypes.data, ty.ctypes.data, n)
    return tx[:n], ty[:n]


def generate(cfg: SceneConfig, want_float: bool = False):
    """Returns (raw bytes array (na, nr, 2) uint8/int8, float array (na, nr) complex64 or None)."""
    p = _params(cfg)
    dt = np.uint8 if cfg.raw_format == "u8_ers" else np.int8
    raw = np.empty((cfg.na, cfg.nr, 2), dt)
    flt = np.empty((cfg.na, cfg.nr), np.complex64) if want_float else None
    rc = _lib().asarf_scene_generate(ctypes.byref(p), raw.ctypes.data, flt.ctypes.data if flt is not None else None)
    if rc < 0:
        raise RuntimeError(f"scene generation failed ({rc})")
    return raw, flt


def raw_to_complex(raw: np.ndarray) -> np.ndarray:
    """Ingest: I/Q bytes -> complex64 (float of each byte, envisat_format_kernels.cu:116-137)."""
    f = raw.astype(np.float32)
    return (f[..., 0] + 1j * f[..., 1]).astype(np.complex64)
