"""Build the HIP library in-tree (hipcc, gfx950) and the test oracle.

    python -m tensordem_amd.build

`-ffp-contract=off` keeps every multiply and add separately rounded so the
FP64 kernels reproduce the oracle's arithmetic bit for bit;
`-packed-fp32-ops` is disabled for all HIP code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SOURCES = sorted((PKG / "csrc").glob("*.hip"))
TARGET = PKG / "libtensordem_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function"]


def build_library(verbose: bool = False) -> Path:
    tmp = TARGET.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, "-I", str(ROOT / "include"), "-o", str(tmp), *map(str, SOURCES)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, TARGET)
    return TARGET


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


def build(verbose: bool = False) -> None:
    build_library(verbose)
    build_oracle()


if __name__ == "__main__":
    build(verbose=True)
    print(TARGET)
