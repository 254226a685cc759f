"""Build recipe for librfmf.so (HIP kernels for gfx950 + C++ host + C ABI)
and the rfmf_main CLI, in-tree under rfmf_amd/lib and rfmf_amd/bin.

    python -m rfmf_amd.build            # build everything
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "lib" / "librfmf.so"
CLI = PKG / "bin" / "rfmf_main"

ARCH = os.environ.get("RFMF_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
          f"-I{ROOT / 'include'}"]
# Packed FP32 instructions are kept out of every device code object.
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

HOST_SOURCES = ["host/graph.cpp", "host/read_gset.cpp", "host/max_cut.cpp", "host/host_common.cpp"]
# The slice kernel's variants compile in one unit per word width and score
# mode (slice_w*.hip), in parallel.
SLICE_SOURCES = sorted(f"hip/{p.name}" for p in (CSRC / "hip").glob("slice_w*.hip"))
HIP_SOURCES = ["hip/anneal_kernels.hip", *SLICE_SOURCES, "hip/engine.cpp", "hip/capi.cpp", "hip/server.cpp",
               "hip/tune_flips.cpp", "hip/noalloc.hip"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _compile(src: str) -> Path:
    obj = OBJ / (src.replace("/", "_") + ".o")
    path = CSRC / src
    if src.startswith("hip/"):
        cmd = [HIPCC, *HIP_FLAGS, *COMMON, "-x", "hip", "-c", str(path), "-o", str(obj)]
    else:
        cmd = [CXX, *COMMON, "-c", str(path), "-o", str(obj)]
    _run(cmd)
    return obj


def build(jobs: int | None = None) -> Path:
    """Compile every source (always, no staleness check) and link the library and CLI."""
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    CLI.parent.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES + HOST_SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    _run([CXX, *COMMON, str(CSRC / "cli" / "rfmf_main.cpp"), "-o", str(CLI),
          f"-L{LIB.parent}", "-lrfmf", "-Wl,-rpath,$ORIGIN/../lib"])
    return LIB


def build_oracle() -> None:
    """Compile the CPU oracle (test infrastructure) and, when the reference
    checkout is present, the reference host sources it is checked against."""
    _run(["make", "-B", "-C", str(ROOT / "oracle"), "oracle"])
    if Path("/root/reference/src/host_common.cpp").exists():
        _run(["make", "-B", "-C", str(ROOT / "oracle"), "ref"])


if __name__ == "__main__":
    build()
    build_oracle()
    print(LIB)
    sys.exit(0)
