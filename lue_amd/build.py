"""Build liblue_amd.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

    python -m lue_amd.build

The translation units (one per routing operation, plus the rest of the C ABI) compile in parallel and are
linked into one shared library.
"""
from __future__ import annotations

import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUTPUT = os.path.join(HERE, "liblue_amd.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    # packed-FP32 instructions stay off (they are generated from plain float code otherwise)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _errors(stderr):
    return [line for line in stderr.splitlines() if "packed-fp32-ops' is not a recognized" not in line]


def _compile(src, obj, verbose):
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    return src, proc.returncode, _errors(proc.stderr)


def build(verbose=False, jobs=None) -> str:
    jobs = jobs or min(len(sources()), max(1, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory(prefix="lue_amd_build_") as tmp:
        objects = []
        with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
            futures = []
            for src in sources():
                obj = os.path.join(tmp, os.path.basename(src) + ".o")
                objects.append(obj)
                futures.append(pool.submit(_compile, src, obj, verbose))
            failures = []
            for f in futures:
                src, rc, errors = f.result()
                if rc != 0:
                    failures.append(f"{os.path.basename(src)}:\n" + "\n".join(errors))
                elif verbose and errors:
                    print("\n".join(errors), file=sys.stderr)
        if failures:
            raise RuntimeError("hipcc failed:\n" + "\n".join(failures))
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUTPUT + ".tmp", *objects]
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError("link failed:\n" + "\n".join(_errors(proc.stderr)))
    os.replace(OUTPUT + ".tmp", OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
