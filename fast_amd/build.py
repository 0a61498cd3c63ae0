"""Build libfasttau.so (HIP kernels for gfx950 + C++ host driver) and the `fast` CLI.

Device code goes through hipcc with -ffp-contract=off (the operator must round exactly like
the reference CPU operator) and with packed FP32 instructions disabled.  Host-only C++ is
compiled with g++.  Outputs land next to this file: fast_amd/libfasttau.so, fast_amd/fast.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
ROOT = HERE.parent
INCLUDE = ROOT / "include"
OBJ = HERE / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("FAST_TAU_ARCH", "gfx950")

HIP_SOURCES = ["tau_kernels.hip", "rb_kernels.hip", "ccl_kernels.hip", "alpha_kernels.hip"]
CXX_SOURCES = ["tau_api.cpp", "ccl_api.cpp", "alpha_api.cpp", "server.cpp", "volume_io.cpp"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
CXX_FLAGS = [
    "-O2",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Wall",
    "-Wextra",
    "-Wno-unused-parameter",
    "-D__HIP_PLATFORM_AMD__",
    f"-I{ROCM / 'include'}",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    # surface compiler warnings
    if res.stderr.strip():
        sys.stderr.write(res.stderr)


def _compile(src: str) -> Path:
    out = OBJ / (src + ".o")
    path = CSRC / src
    if src.endswith(".hip"):
        cmd = [str(ROCM / "bin" / "hipcc"), *HIPCC_FLAGS, "-c", str(path), "-o", str(out)]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", str(path), "-o", str(out)]
    _run(cmd)
    return out


def build(verbose: bool = False) -> Path:
    """Compile every source and link libfasttau.so and the `fast` CLI.  Always rebuilds."""
    OBJ.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES + CXX_SOURCES))
    lib = HERE / "libfasttau.so"
    tmp = HERE / "libfasttau.so.tmp"
    _run([str(ROCM / "bin" / "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objs], f"-L{ROCM / 'lib'}", "-lamdhip64", "-ldl", "-Wl,-soname,libfasttau.so"])
    os.replace(tmp, lib)
    cli_tmp = HERE / "fast.tmp"
    _run(["g++", *CXX_FLAGS, str(CSRC / "fast_cli.cpp"), "-o", str(cli_tmp), f"-L{HERE}", "-lfasttau",
          f"-Wl,-rpath,$ORIGIN", f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(cli_tmp, HERE / "fast")
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build(verbose=True)
