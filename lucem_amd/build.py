"""Builds ``liblucem_hip.so`` (HIP kernels + C ABI) for gfx950 in place with hipcc.

Each source compiles to its own object under ``lucem_amd/_build/`` (in parallel), then the
objects link into the shared library. Every source carries its own kernels and launchers, so no
relocatable device code is needed.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "liblucem_hip.so"
SOURCES = ["lucem_budget.cpp", "rl_tiled.hip", "lucem_tiled.cpp", "rl_kernels.hip", "rl_grad.hip", "rl_direct.hip", "wiener.hip", "fft_axis.hip", "noise.hip", "correlate.hip",
           "phantom.hip", "phase.hip", "spectral.hip", "lucem_hip.cpp", "lucem_batch.cpp", "lucem_grad.cpp", "lucem_tune.cpp", "lucem_wiener.cpp",
           "lucem_fft.cpp", "lucem_noise.cpp", "lucem_correlate.cpp", "lucem_phantom.cpp",
           "lucem_phase.cpp", "lucem_spectral.cpp", "lucem_noalloc.hip"]

COMPILE_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
LINK_FLAGS = ["--offload-arch=gfx950", "-shared", "-fPIC"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def sources() -> list[Path]:
    return [CSRC / s for s in SOURCES if (CSRC / s).exists()]


def _compile(src: Path) -> tuple[Path, subprocess.CompletedProcess]:
    obj = OBJ / (src.name + ".o")
    # .cpp sources hold host code only but include the HIP runtime headers: compile them as HIP too
    lang = [] if src.suffix == ".hip" else ["-x", "hip"]
    cmd = [hipcc(), *COMPILE_FLAGS, *lang, "-c", str(src), "-o", str(obj)]
    return obj, subprocess.run(cmd, cwd=str(CSRC), capture_output=True, text=True)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compiles every HIP/C++ source of the library into ``lucem_amd/liblucem_hip.so``."""
    OBJ.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, sources()))
    errs = [(o, r) for o, r in results if r.returncode != 0]
    if errs:
        o, r = errs[0]
        raise RuntimeError(f"hipcc failed on {o.name} ({r.returncode}):\n{r.stderr[-6000:]}")
    if verbose:
        for o, r in results:
            if r.stderr:
                print(r.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    res = subprocess.run([hipcc(), *LINK_FLAGS, "-o", str(tmp), *[str(o) for o, _ in results]], cwd=str(CSRC),
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc link failed ({res.returncode}):\n{res.stderr[-6000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
