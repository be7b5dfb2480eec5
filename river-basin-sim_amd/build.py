"""Build librbs.so (the gfx950 HIP kernels + C-ABI) in-tree with hipcc.

``python build.py`` compiles every ``csrc/*.hip`` for gfx950 and links
``rbs/librbs.so``.  Flags: no FMA contraction (-ffp-contract=off: the float32
operation order of the reference is part of the contract), no fast-math (Kahan
compensation must survive), packed-FP32 instructions disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "rbs" / "librbs.so"
TERRAIN_OUT = HERE / "rbs" / "libterrain.so"
INCLUDE = HERE.parent / "include"
ARCH = os.environ.get("RBS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; librbs needs ROCm")
    return exe


def build(verbose: bool = False) -> Path:
    """Compile all HIP sources and link librbs.so; returns its path."""
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    cc = hipcc()

    def compile_one(src: Path) -> Path:
        obj = objdir / (src.stem + ".o")
        cmd = [cc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
        return obj

    jobs = min(len(sources), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, sources))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    tmp.replace(OUT)
    build_terrain()
    return OUT


def build_terrain() -> Path:
    """Host-only helper of the synthetic-terrain generator (gcc, no GPU code)."""
    src = HERE / "csrc_host" / "terrain.c"
    tmp = TERRAIN_OUT.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-o", str(tmp), str(src), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"terrain helper build failed:\n{r.stderr}")
    tmp.replace(TERRAIN_OUT)
    return TERRAIN_OUT


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
