"""Compile the engine for gfx950: libglazier_hip.so beside this file.

hipcc cross-compiles without a GPU. Every source is built with
-ffp-contract=off so each f32/f64 operation rounds once (the CPU emulation
that checks the kernels bit for bit is built the same way), and with packed
f32 instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libglazier_hip.so"
CLI = HERE / "glazier"
CLI_SOURCES = ["cli.cpp"]
SOURCES = ["sweep.hip", "sweep_plain.hip", "fields.hip", "cells.hip", "batch.hip", "adjoint.hip",
           "engine.cpp", "batch_host.cpp", "fit_host.cpp", "blueprint.cpp"]
OBJ_DIR = HERE / "_build"
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    return "hipcc"


def _compile(src: str) -> Path:
    obj = OBJ_DIR / (src.replace(".", "_") + ".o")
    flags = [f for f in FLAGS if f != "-shared"]
    cmd = [hipcc(), *flags, "-c", "-o", str(obj), str(CSRC / src)]
    done = subprocess.run(cmd, capture_output=True, text=True)
    if done.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n" + done.stderr[-4000:])
    return obj


def build(verbose: bool = False) -> Path:
    OBJ_DIR.mkdir(exist_ok=True)
    # one translation unit per process: the kernel files dominate the time
    workers = max(1, min(len(SOURCES), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        objects = list(pool.map(_compile, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects)]
    done = subprocess.run(cmd, capture_output=True, text=True)
    if done.returncode != 0:
        raise RuntimeError("linking failed:\n" + done.stderr[-4000:])
    if verbose and done.stderr:
        print(done.stderr)
    os.replace(tmp, OUT)
    # the command line, linked against the library beside it
    tmp_cli = CLI.with_name("glazier.tmp")
    cmd = [hipcc(), "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unused-command-line-argument",
           "-o", str(tmp_cli), *[str(CSRC / s) for s in CLI_SOURCES], "-L", str(HERE),
           "-lglazier_hip", "-Wl,-rpath,$ORIGIN"]
    done = subprocess.run(cmd, capture_output=True, text=True)
    if done.returncode != 0:
        raise RuntimeError("hipcc failed on the command line:\n" + done.stderr[-4000:])
    os.replace(tmp_cli, CLI)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
