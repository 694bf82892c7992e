"""Build ``libolhip.so`` (HIP kernels + C ABI) for gfx950 with hipcc.

Usage: ``python -m ordinarylight_amd.build_native [--jobs N]``.  The library
is written next to this file so it travels with the source tree.
"""

from __future__ import annotations

import argparse
import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIBRARY = HERE / "libolhip.so"
SOURCES = ("bvh_build.hip", "scan.hip", "gi.hip", "engine.hip", "e0.hip", "output.hip", "api.hip")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OLHIP_ARCH", "gfx950")

# -ffp-contract=off: every float expression rounds as written so the oracle
# reproduces it; packed-fp32 instructions are disabled: some packed FP32 forms
# have produced wrong results on this hardware now and then.
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{CSRC}",
]


def _compile(source: str) -> pathlib.Path:
    obj = BUILD / (source + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(CSRC / source), "-o", str(obj)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed for {source}:\n{result.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> pathlib.Path:
    """Compile every HIP source and link the shared library; returns its path."""
    if not shutil.which(HIPCC) and not pathlib.Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(_compile, SOURCES))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{result.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--jobs", type=int, default=None)
    args = parser.parse_args(argv)
    build(args.jobs, verbose=True)


if __name__ == "__main__":
    sys.exit(main())
