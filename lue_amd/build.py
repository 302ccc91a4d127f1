"""Build liblue_amd.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

    python -m lue_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "routing.hip")]
HEADERS_DIR = os.path.join(HERE, "csrc")
OUTPUT = os.path.join(HERE, "liblue_amd.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-shared",
    # packed-FP32 instructions stay off (they are generated from plain float code otherwise)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def build(verbose=False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", HEADERS_DIR, "-o", OUTPUT + ".tmp",
           *SOURCES]
    if verbose:
        print(" ".join(cmd))
    proc = subprocess.run(cmd, capture_output=True, text=True)
    errors = [line for line in proc.stderr.splitlines() if "packed-fp32-ops' is not a recognized" not in line]
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    os.replace(OUTPUT + ".tmp", OUTPUT)
    if verbose and errors:
        print("\n".join(errors), file=sys.stderr)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
