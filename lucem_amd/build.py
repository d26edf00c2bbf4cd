"""Builds ``liblucem_hip.so`` (HIP kernels + C ABI) for gfx950 in place with hipcc."""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "liblucem_hip.so"
SOURCES = ["rl_kernels.hip", "rl_grad.hip", "rl_direct.hip", "wiener.hip", "fft_axis.hip", "noise.hip", "correlate.hip",
           "phantom.hip", "phase.hip", "spectral.hip", "lucem_hip.cpp", "lucem_batch.cpp", "lucem_grad.cpp", "lucem_tune.cpp", "lucem_wiener.cpp",
           "lucem_fft.cpp", "lucem_noise.cpp", "lucem_correlate.cpp", "lucem_phantom.cpp",
           "lucem_phase.cpp", "lucem_spectral.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def sources() -> list[Path]:
    return [CSRC / s for s in SOURCES if (CSRC / s).exists()]


def build(verbose: bool = False) -> Path:
    """Compiles every HIP/C++ source of the library into ``lucem_amd/liblucem_hip.so``."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", str(tmp), *map(str, sources())]
    res = subprocess.run(cmd, cwd=str(CSRC), capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-6000:]}")
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
