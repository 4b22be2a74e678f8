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
# (source, preprocessor defines): the templated matrix-core kernels build once
# per precision, so the units compile in parallel
UNITS = [
    ("hz_capi.hip", {}),
    ("hz_knn.hip", {}),
    ("hz_pca.hip", {}),
    *[(f"hz_k_{k}.hip", {"HZ_UNIT_PREC": p}) for k in ("sweep", "row", "cent") for p in (0, 1)],
]
ARCH = os.environ.get("HZ_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def compile_command(src: Path, obj: Path, defines: dict | None = None) -> list[str]:
    defs = [f"-D{k}={v}" for k, v in (defines or {}).items()]
    return [hipcc(), *FLAGS, *defs, "-c", str(src), "-o", str(obj)]


def link_command(objs: list[Path], out: Path) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", *map(str, objs), "-o", str(out)]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stdout[-4000:] +
                           "\n" + res.stderr[-8000:])


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source (one process per translation unit, in
    parallel) and link ``_lib/libharmony_hip.so``; always rebuilds."""
    LIB_PATH.parent.mkdir(parents=True, exist_ok=True)
    objdir = LIB_PATH.parent / "obj"
    objdir.mkdir(exist_ok=True)
    jobs = []
    for src, defines in UNITS:
        tag = "".join(f"_{v}" for v in defines.values())
        obj = objdir / (Path(src).stem + tag + ".o")
        cmd = compile_command(CSRC / src, obj, defines)
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((cmd, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.PIPE, text=True)))
    errors = []
    for cmd, obj, proc in jobs:
        out, err = proc.communicate()
        if proc.returncode != 0:
            errors.append(" ".join(cmd) + "\n" + out[-4000:] + "\n" + err[-8000:])
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = link_command([o for _, o, _ in jobs], tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
