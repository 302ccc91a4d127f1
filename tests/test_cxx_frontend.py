"""The C++ front end (lue_amd/include/lue/framework/routing.hpp): compiles against the C ABI (CPU) and runs
LUE's golden cases through lue::value_policies::* (GPU)."""
import math
import os
import subprocess

import pytest

from conftest import ROOT, load_golden

CXX_DIR = os.path.join(ROOT, "tests", "cxx")


def _num(v):
    if v == "nan" or (isinstance(v, float) and math.isnan(v)):
        return "nan"
    return repr(float(v))


def _vec(raster):
    return "{" + ", ".join(_num(v) for row in raster for v in row) + "}"


def write_cases(path):
    lines = ["std::vector<Case> const& cases() {", "static std::vector<Case> const all{"]
    for op in ("inflow_count", "accu", "accu_threshold", "accu_fraction", "accu_capacity", "accu_trigger",
               "kinematic_wave"):
        for c in load_golden(op):
            if c["kind"] == "iterate_to_new_discharge":
                continue
            rows, cols = c["array_shape"]
            prows, pcols = c["partition_shape"]
            args = []
            for a in c.get("args", {}).values():
                if "raster" in a:
                    args.append(f"Arg{{true, 0.0, {_vec(a['raster'])}}}")
                else:
                    args.append(f"Arg{{false, {_num(a['scalar'])}, {{}}}}")
            keys = ["count"] if c["kind"] == "inflow_count" else ["outflow", "remainder"]
            exp = [c["expected"][k] for k in keys if k in c.get("expected", {})]
            close = any(e["compare"] == "close" for e in exp)
            lines.append(
                f'Case{{"{c["name"]}", "{c["kind"]}", {{{rows}, {cols}}}, {{{prows}, {pcols}}}, {_vec(c["ldd"])}, '
                f'{{{", ".join(args)}}}, {{{", ".join(_vec(e["raster"]) for e in exp)}}}, '
                f'{"true" if close else "false"}, {"true" if c.get("property") == "all_positive" else "false"}}},')
    for op in ("upstream", "downstream"):
        for c in load_golden(op):
            rows, cols = c["array_shape"]
            prows, pcols = c["partition_shape"]
            lines.append(
                f'Case{{"{c["name"]}", "{op}", {{{rows}, {cols}}}, {{{prows}, {pcols}}}, {_vec(c["ldd"])}, '
                f'{{Arg{{true, 0.0, {_vec(c["material"])}}}}}, {{{_vec(c["expected"]["material"]["raster"])}}}, '
                f'false, false}},')
    for c in load_golden("downstream_distance"):
        rows, cols = c["array_shape"]
        prows, pcols = c["partition_shape"]
        lines.append(
            f'Case{{"{c["name"]}", "downstream_distance", {{{rows}, {cols}}}, {{{prows}, {pcols}}}, '
            f'{_vec(c["ldd"])}, {{Arg{{false, {_num(c["cell_size"])}, {{}}}}}}, '
            f'{{{_vec(c["expected"]["distance"]["raster"])}}}, false, false}},')
    for c in load_golden("accu_info"):
        rows, cols = c["array_shape"]
        prows, pcols = c["partition_shape"]
        lines.append(
            f'Case{{"{c["name"]}", "accu_info", {{{rows}, {cols}}}, {{{prows}, {pcols}}}, {_vec(c["ldd"])}, {{}}, '
            f'{{{_vec(c["expected"]["cell_class"]["raster"])}}}, false, false}},')
    for c in load_golden("d8_flow_direction"):
        rows, cols = c["array_shape"]
        prows, pcols = c["partition_shape"]
        lines.append(
            f'Case{{"{c["name"]}", "d8_flow_direction", {{{rows}, {cols}}}, {{{prows}, {pcols}}}, '
            f'{_vec(c["elevation"])}, {{}}, {{{_vec(c["expected"]["flow_direction"]["raster"])}}}, false, false}},')
    lines += ["};", "return all;", "}"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def build_driver(tmp_path):
    inc = tmp_path / "golden_cases.inc"
    write_cases(inc)
    exe = tmp_path / "frontend_test"
    lib_dir = os.path.join(ROOT, "lue_amd")
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-I", str(tmp_path), "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "lue_amd", "include"), "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
           os.path.join(CXX_DIR, "frontend_test.cpp"), "-L", lib_dir, "-llue_amd", f"-Wl,-rpath,{lib_dir}",
           "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-4000:]
    return exe


def test_frontend_compiles(tmp_path):
    if not os.path.exists(os.path.join(ROOT, "lue_amd", "liblue_amd.so")):
        from lue_amd import build

        build.build()
    build_driver(tmp_path)


@pytest.mark.gpu
def test_frontend_golden_cases(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    exe = build_driver(tmp_path)
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(proc.stdout[-3000:])
    assert proc.returncode == 0, proc.stdout[-3000:] + proc.stderr[-2000:]
