"""Seeded synthetic workload inputs: device generators (C ABI) and their numpy twins.

The device kernels (lue_amd/csrc/synthetic.hpp) and the numpy functions below produce identical bytes:
both use the same counter-based hash and evaluate the DEM in FP64 without contraction. The numpy side
exists so that tests can check the device generator and build small inputs on the host.

Workloads (the benchmark configurations):
  - DEM e(r, c) = (R - r) + 50 |sin(pi c / 256)| + 0.9 u, ldd = sqrt(2)-weighted steepest descent over
    in-array neighbours, ties to the first of N, NE, E, SE, S, SW, W, NW, bottom row outlets (5);
  - uniform fields lo + (hi - lo) u.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
STREAM_MUL = 0xD1B54A32D192ED03

# streams of the seeded fields; stream 0 is the DEM noise
STREAM_DEM = 0
STREAM_MATERIAL = 1
STREAM_THRESHOLD = 2
STREAM_CURRENT = 3
STREAM_LATERAL = 4
STREAM_ALPHA = 5


def _finalise(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def hash3(seed, stream, idx):
    key = _finalise(np.uint64((seed + GOLDEN * (stream + 1)) & MASK64))
    with np.errstate(over="ignore"):
        mixed = np.asarray(idx, dtype=np.uint64) * np.uint64(STREAM_MUL) + np.uint64(GOLDEN)
    return _finalise(key ^ mixed)


def unit_f64(h):
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform_host(seed, stream, lo, hi, shape, dtype=np.float32):
    n = int(np.prod(shape))
    u = unit_f64(hash3(seed, stream, np.arange(n, dtype=np.uint64)))
    width = hi - lo
    return (lo + width * u).astype(dtype).reshape(shape)


def sin_table():
    return np.array([abs(math.sin(math.pi * k / 256.0)) for k in range(256)], dtype=np.float64)


def dem_host(seed, rows, cols):
    r = np.arange(rows, dtype=np.int64)[:, None]
    c = np.arange(cols, dtype=np.int64)[None, :]
    idx = (r * cols + c).astype(np.uint64)
    u = unit_f64(hash3(seed, STREAM_DEM, idx))
    base = (rows - r).astype(np.float64)
    ridge = 50.0 * sin_table()[c & 255]
    return (base + ridge) + 0.9 * u


CODES = [8, 9, 6, 3, 2, 1, 4, 7]  # N, NE, E, SE, S, SW, W, NW
OFFSETS = [(-1, 0), (-1, 1), (0, 1), (1, 1), (1, 0), (1, -1), (0, -1), (-1, -1)]


def ldd_host(seed, rows, cols):
    e = dem_host(seed, rows, cols)
    best = np.zeros((rows, cols))
    code = np.full((rows, cols), 5, dtype=np.uint8)
    inv_sqrt2 = 0.70710678118654752440
    padded = np.full((rows + 2, cols + 2), np.nan)
    padded[1:-1, 1:-1] = e
    for k, (dr, dc) in enumerate(OFFSETS):
        e1 = padded[1 + dr:1 + dr + rows, 1 + dc:1 + dc + cols]
        diff = e - e1
        drop = diff * inv_sqrt2 if k & 1 else diff
        better = drop > best  # NaN (outside the array) never wins
        best = np.where(better, drop, best)
        code = np.where(better, np.uint8(CODES[k]), code)
    code[rows - 1, :] = 5
    return code


# ---- device generators ------------------------------------------------------------------------------

def _stream():
    return torch.cuda.current_stream().cuda_stream


def ldd_device(seed, rows, cols, device=None):
    lib = _lib.load()
    out = torch.empty((rows, cols), dtype=torch.uint8, device=device or "cuda")
    _lib.check(lib.lue_amd_synthetic_ldd(seed, rows, cols, out.data_ptr(), _stream()))
    return out


def dem_device(seed, rows, cols, dtype=torch.float32, device=None):
    """dem_host(seed, rows, cols) on the device, rounded to dtype (the same values)."""
    lib = _lib.load()
    out = torch.empty((rows, cols), dtype=dtype, device=device or "cuda")
    fn = lib.lue_amd_synthetic_dem_f32 if dtype == torch.float32 else lib.lue_amd_synthetic_dem_f64
    _lib.check(fn(seed, rows, cols, out.data_ptr(), _stream()))
    return out


def uniform_device(seed, stream, lo, hi, shape, dtype=torch.float32, device=None):
    lib = _lib.load()
    out = torch.empty(tuple(shape), dtype=dtype, device=device or "cuda")
    fn = lib.lue_amd_synthetic_uniform_f32 if dtype == torch.float32 else lib.lue_amd_synthetic_uniform_f64
    _lib.check(fn(seed, stream, lo, hi, out.numel(), out.data_ptr(), _stream()))
    return out


def uniform_into(seed, stream, lo, hi, out):
    """uniform_device into an existing contiguous float tensor (stream order)."""
    lib = _lib.load()
    fn = lib.lue_amd_synthetic_uniform_f32 if out.dtype == torch.float32 else lib.lue_amd_synthetic_uniform_f64
    _lib.check(fn(seed, stream, lo, hi, out.numel(), out.data_ptr(), _stream()))
    return out


# ---- batches of catchments --------------------------------------------------------------------------

def batch_shapes(seed, count, side_lo, side_hi, aspect_lo=0.5, aspect_hi=2.0):
    """Shapes (rows, cols) of `count` catchments: side log-uniform on [side_lo, side_hi] (the geometric mean
    of rows and cols), aspect ratio cols / rows uniform on [aspect_lo, aspect_hi]; each side at least 2.
    Catchment i uses seed + 1 + i for its own fields."""
    u = unit_f64(hash3(seed, 100, np.arange(count, dtype=np.uint64)))
    v = unit_f64(hash3(seed, 101, np.arange(count, dtype=np.uint64)))
    side = np.exp(math.log(side_lo) + (math.log(side_hi) - math.log(side_lo)) * u)
    aspect = aspect_lo + (aspect_hi - aspect_lo) * v
    rows = np.maximum(2, np.rint(side / np.sqrt(aspect))).astype(np.int64)
    cols = np.maximum(2, np.rint(side * np.sqrt(aspect))).astype(np.int64)
    return [(int(r), int(c)) for r, c in zip(rows, cols)]
