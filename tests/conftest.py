import json
import os
import sys

import numpy as np
import pytest

# beta as the FP32 path receives it (0.6f): FP64 references for FP32 results take this value
BETA32 = float(np.float32(0.6))

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_golden(op):
    with open(os.path.join(GOLDEN, f"{op}.json")) as f:
        return json.load(f)["cases"]


def raster(values, dtype):
    """JSON raster (nested lists, 'nan' strings) -> numpy array."""
    return np.array([[np.nan if v == "nan" else v for v in row] for row in values], dtype=dtype)


def material_raster(values, dtype):
    """JSON raster -> numpy array of a material element type; 'nan' is that type's no-data (NaN, or the
    minimum of a signed / maximum of an unsigned integer type)."""
    dtype = np.dtype(dtype)
    if dtype.kind == "f":
        return raster(values, dtype)
    info = np.iinfo(dtype)
    nd = info.min if dtype.kind == "i" else info.max
    return np.array([[nd if v == "nan" else v for v in row] for row in values], dtype=dtype)


MATERIAL_DTYPES = [np.uint8, np.uint32, np.uint64, np.int8, np.int32, np.int64, np.float32, np.float64]


def golden_arg(arg, dtype):
    if "raster" in arg:
        return raster(arg["raster"], dtype)
    v = arg["scalar"]
    return float("nan") if v == "nan" else float(v)


def assert_equal_nan(got, want):
    """Exact comparison with NaN == NaN (lue::test::check_arrays_are_equal, compare.hpp:53-100)."""
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape
    both_nan = np.isnan(got) & np.isnan(want) if got.dtype.kind == "f" else np.zeros(got.shape, bool)
    ok = both_nan | (got == want)
    assert ok.all(), f"mismatch at {np.argwhere(~ok)[:5].tolist()}: got {got[~ok][:5]} want {want[~ok][:5]}"


def assert_close_boost(got, want, tol=1e-6):
    """Boost.Test tolerance comparison as used by check_arrays_are_close (compare.hpp:103-118): relative
    tolerance on both operands; when either side is exactly zero the other must be smaller than tol."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape
    diff = np.abs(got - want)
    zero = (got == 0) | (want == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.maximum(diff / np.abs(got), diff / np.abs(want))
    ok = np.where(zero, np.maximum(np.abs(got), np.abs(want)) < tol, rel <= tol)
    assert ok.all(), f"mismatch at {np.argwhere(~ok)[:5].tolist()}: got {got[~ok][:5]} want {want[~ok][:5]}"


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle

    oracle.build()
    return oracle
