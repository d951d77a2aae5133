"""Build recipes for the native libraries of the package.

* ``libtea_engine.so`` -- the HIP kernel layer (C-ABI of include/tea_leaf_engine.h)
  and the warm-start network and trainer (include/tea_warmstart.h), compiled for
  gfx950 with hipcc;
* ``libtea_driver.so`` -- the host driver (deck reader, time loop, solver control).

Both are built in-tree, next to this file, so they travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
ENGINE_SO = PKG_DIR / "libtea_engine.so"
DRIVER_SO = PKG_DIR / "libtea_driver.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OFFLOAD_ARCH = os.environ.get("TEA_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no implicit fused multiply-add, so every kernel evaluates
# its expressions exactly as written (bitwise equal to the CPU evaluation).
# -packed-fp32-ops: the compiler must not emit packed FP32 instructions.
HIP_FLAGS = [
    f"--offload-arch={OFFLOAD_ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fvisibility=hidden",
    "-Wl,-soname,libtea_engine.so",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def build_engine() -> Path:
    sources = [CSRC / "engine.hip", CSRC / "warmstart.hip"]
    tmp = ENGINE_SO.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-o", str(tmp), *map(str, sources)])
    os.replace(tmp, ENGINE_SO)
    return ENGINE_SO


def build_driver() -> Path:
    tmp = DRIVER_SO.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=default", "-Wall", "-o", str(tmp),
          str(CSRC / "tea_driver.cpp"), "-ldl"])
    os.replace(tmp, DRIVER_SO)
    return DRIVER_SO


def build_all() -> None:
    build_engine()
    build_driver()


if __name__ == "__main__":
    build_all()
    print(ENGINE_SO)
    print(DRIVER_SO)
