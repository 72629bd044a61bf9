"""Seeded synthetic images of the benchmark workloads (ctypes over libdcf_synth.so).

The generators are deterministic functions of (shape, seed, counts); see synth/dcf_synth.c.
Shapes follow numpy order: 2D images are (H, W), volumes (D, H, W).
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os

import numpy as np

from . import SYNTH_PATH

_lib = None

# per-workload seeds
SEED = {"disks": 0xD15C0000, "vnoise": 0xD15C0001, "spheres": 0xD15C0002, "blobs": 0xD15C0003, "batch": 0xD15C0004}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_PATH):
            from .build import build_synth
            build_synth()
        L = C.CDLL(SYNTH_PATH)
        vp, i64, u64, i = C.c_void_p, C.c_int64, C.c_uint64, C.c_int
        L.dcf_synth_disks.argtypes = [i64, i64, i, u64, i, vp]
        L.dcf_synth_vnoise.argtypes = [i64, i64, u64, i, vp]
        L.dcf_synth_spheres.argtypes = [i64, i64, i64, u64, i, vp]
        L.dcf_synth_blobs.argtypes = [i64, i64, u64, i, vp, vp]
        L.dcf_synth_batch_shapes.argtypes = [i, i64, i64, u64, vp, vp, vp, vp, vp, C.c_double, C.c_double]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def disks(H: int, W: int, is16: bool = False, seed: int = SEED["disks"], ndisks: int | None = None) -> np.ndarray:
    """uniform noise -> 5x5 box blur -> disks r in [3,40] with offsets +-60 (x257 for u16), clipped."""
    if ndisks is None:
        ndisks = max(1, int(round(200 * (H * W) / (1024 * 1024))))
    out = np.empty((H, W), dtype=np.uint16 if is16 else np.uint8)
    if lib().dcf_synth_disks(W, H, 1 if is16 else 0, seed, ndisks, _p(out)) != 0:
        raise RuntimeError("dcf_synth_disks failed")
    return out


def vnoise(H: int, W: int, seed: int = SEED["vnoise"], nell: int | None = None) -> np.ndarray:
    """6-octave value noise on [0,65535] + rotated ellipses with offsets +-4000, clipped (u16)."""
    if nell is None:
        nell = max(1, int(round(20000 * (H * W) / (8192 * 8192))))
    out = np.empty((H, W), dtype=np.uint16)
    if lib().dcf_synth_vnoise(W, H, seed, nell, _p(out)) != 0:
        raise RuntimeError("dcf_synth_vnoise failed")
    return out


def spheres(D: int, H: int, W: int, seed: int = SEED["spheres"], nsph: int | None = None) -> np.ndarray:
    """uniform noise -> Gaussian blur sigma 2 -> spheres r in [2,30] with offsets +-50, clipped (u8)."""
    if nsph is None:
        nsph = max(1, int(round(5000 * (D * H * W) / (512 ** 3))))
    out = np.empty((D, H, W), dtype=np.uint8)
    if lib().dcf_synth_spheres(W, H, D, seed, nsph, _p(out)) != 0:
        raise RuntimeError("dcf_synth_spheres failed")
    return out


def blobs(H: int, W: int, seed: int = SEED["blobs"], nblobs: int | None = None, raw: bool = False):
    """lognormal background + Gaussian blobs (f32) and its quantisation to 65536 levels (u16)."""
    if nblobs is None:
        nblobs = max(1, int(round(50000 * (H * W) / (4096 * 4096))))
    q = np.empty((H, W), dtype=np.uint16)
    r = np.empty((H, W), dtype=np.float32) if raw else None
    if lib().dcf_synth_blobs(W, H, seed, nblobs, _p(r) if raw else None, _p(q)) != 0:
        raise RuntimeError("dcf_synth_blobs failed")
    return (q, r) if raw else q


def batch_shapes(count: int = 256, lo: int = 64, hi: int = 2048, seed: int = SEED["batch"], lam_lo: float = 10.0,
                 lam_hi: float = 5000.0):
    """(W, H, is16, seed, lambda) of every image of the uneven batch."""
    W = np.empty(count, np.int64)
    H = np.empty(count, np.int64)
    is16 = np.empty(count, np.int32)
    seeds = np.empty(count, np.uint64)
    lam = np.empty(count, np.float64)
    if lib().dcf_synth_batch_shapes(count, lo, hi, seed, _p(W), _p(H), _p(is16), _p(seeds), _p(lam), lam_lo, lam_hi) != 0:
        raise RuntimeError("dcf_synth_batch_shapes failed")
    return [(int(W[i]), int(H[i]), bool(is16[i]), int(seeds[i]), float(lam[i])) for i in range(count)]


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def log_lambdas(n_bins: int, lo: float, hi: float) -> np.ndarray:
    """lambda vector with lambda_0 = 0 and n_bins-1 log-spaced float32 thresholds lo..hi."""
    v = np.zeros(n_bins, dtype=np.float32)
    v[1:] = np.geomspace(lo, hi, n_bins - 1).astype(np.float32)
    return v
