"""Builds libctlproj.so (HIP kernels for gfx950 + host runtime) in-tree with hipcc/clang++.

Run ``python -m ctl_amd.build`` or call :func:`build`. Outputs go to ``ctl_amd/lib/`` (object
files to ``ctl_amd/_build/``). Every call recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "lib"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
CLANG = str(ROCM / "llvm" / "bin" / "clang++")
ARCH = "gfx950"

DEVICE_SOURCES = ["kernels.hip", "processing.hip"]
HOST_SOURCES = ["api.cpp", "batch.cpp", "geometry.cpp", "phantom.cpp", "serve.cpp", "autotune.cpp",
                "fileio.cpp", "processing.cpp", "extensions.cpp", "errors.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
# Packed FP32 instructions are kept out of every device code object.
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", "-pthread"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        line for line in (res.stdout + res.stderr).splitlines()
        if "packed-fp32-ops" not in line and "--hip-link" not in line
    )
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _compile(src: str) -> Path:
    obj = OBJ / (src + ".o")
    path = CSRC / src
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE_FLAGS, "-c", str(path), "-o", str(obj)]
    else:
        cmd = [CLANG, *COMMON, *HOST_FLAGS, "-c", str(path), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> Path:
    """Compiles every source and links ``ctl_amd/lib/libctlproj.so``; returns its path."""
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    sources = [s for s in DEVICE_SOURCES + HOST_SOURCES if (CSRC / s).exists()]
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    lib = LIB / "libctlproj.so"
    tmp = LIB / "libctlproj.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-pthread"])
    os.replace(tmp, lib)
    # the C++ mirror's demo/test programs, linked against the library
    for name in ("projector_demo", "processing_demo", "extensions_demo"):
        demo = ROOT / "tests" / "cpp" / f"{name}.cpp"
        if demo.exists():
            _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(demo), "-o",
                  str(LIB / name), f"-L{LIB}", "-lctlproj", "-Wl,-rpath,$ORIGIN",
                  f"-L{ROCM / 'lib'}", f"-Wl,-rpath,{ROCM / 'lib'}", "-lamdhip64"])
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build(verbose=True)
