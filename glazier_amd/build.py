"""Compile the engine for gfx950: libglazier_hip.so beside this file.

hipcc cross-compiles without a GPU. Every source is built with
-ffp-contract=off so each f32/f64 operation rounds once (the CPU emulation
that checks the kernels bit for bit is built the same way), and with packed
f32 instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libglazier_hip.so"
CLI = HERE / "glazier"
CLI_SOURCES = ["cli.cpp", "blueprint.cpp"]
SOURCES = ["sweep.hip", "sweep_plain.hip", "fields.hip", "cells.hip", "engine.cpp"]
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


def build(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, "-o", str(tmp), *[str(CSRC / s) for s in SOURCES]]
    done = subprocess.run(cmd, capture_output=True, text=True)
    if done.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + done.stderr[-4000:])
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
