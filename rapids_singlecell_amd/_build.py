"""Build libharmony_hip.so in-tree with hipcc for gfx950."""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB_NAME = "libharmony_hip.so"
LIB_PATH = PKG / "_lib" / LIB_NAME
SOURCES = ["hz_capi.hip"]
ARCH = os.environ.get("HZ_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def compile_command(out: Path = LIB_PATH) -> list[str]:
    return [
        hipcc(),
        "-O3",
        "-std=c++17",
        f"--offload-arch={ARCH}",
        "-fPIC",
        "-shared",
        # no packed FP32 instructions in device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
        *[str(CSRC / s) for s in SOURCES],
        "-o",
        str(out),
    ]


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source into ``_lib/libharmony_hip.so`` (always rebuilds)."""
    LIB_PATH.parent.mkdir(parents=True, exist_ok=True)
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = compile_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(
            "hipcc failed:\n" + res.stdout[-4000:] + "\n" + res.stderr[-8000:]
        )
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
