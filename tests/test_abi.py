"""CPU checks of the C ABI library: it loads, and exports every function include/lue_amd/routing.h declares."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lue_amd", "routing.h")
LIB = os.path.join(ROOT, "lue_amd", "liblue_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lue_amd_\w+)\s*\(", text)))


def test_header_declares_the_routing_ops():
    names = declared_functions()
    for op in ("inflow_count_u8", "accu_f32", "accu_f64", "accu_threshold_f32", "accu_threshold_f64",
               "accu_fraction_f32", "accu_fraction_f64", "kinematic_wave_f32", "kinematic_wave_f64"):
        assert f"lue_amd_{op}" in names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from lue_amd import build

        build.build()
    return ctypes.CDLL(LIB)


def test_library_exports_every_declared_symbol(lib):
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing


def test_status_strings(lib):
    lib.lue_amd_status_string.restype = ctypes.c_char_p
    lib.lue_amd_version.restype = ctypes.c_char_p
    assert lib.lue_amd_status_string(0) == b"success"
    assert lib.lue_amd_status_string(5) == b"invalid flow direction network"
    assert lib.lue_amd_version().count(b".") == 2


def test_argument_errors_without_device(lib):
    """Argument checks run before any HIP call."""
    lib.lue_amd_accu_f32.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                     ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
    lib.lue_amd_last_error.restype = ctypes.c_char_p
    assert lib.lue_amd_accu_f32(None, 4, 4, None, 1.0, None, None) == 1
    assert b"ldd" in lib.lue_amd_last_error()
    assert lib.lue_amd_accu_f32(ctypes.c_void_p(16), 0, 4, None, 1.0, ctypes.c_void_p(16), None) == 1


def test_python_front_end_fails_loudly_without_gpu():
    import torch

    from lue_amd import framework

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        framework.create_partitioned_array((4, 4), (2, 2), torch.float32, 1.0)
