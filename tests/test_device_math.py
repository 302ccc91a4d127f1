"""Device math helpers checked on the host: the FP64 power of the block Newton (kinematic_wave.hpp,
pow_positive, full and short) against long-double powl."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_pow_positive_accuracy(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "pow_positive_check"
    cmd = [hipcc, "-O2", "-std=c++20", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "lue_amd", "csrc"),
           "-x", "hip", os.path.join(ROOT, "tests", "cxx", "pow_positive_check.cpp"), "-o", str(exe)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-4000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    values = dict(line.split() for line in run.stdout.strip().splitlines())
    worst = float(values["worst_relative_error"])
    worst_short = float(values["worst_relative_error_short"])
    # |y ln x| <= 1.5 * 103 here: full bound 2^-52 * 155 + 1e-15 ~ 3.5e-14; short bound
    # 4.4e-13 |y| + 2.2e-13 + 2^-52 |y ln x| plus the roundings ~ 1e-12
    assert worst < 4e-14, worst
    assert worst_short < 1.2e-12, worst_short
