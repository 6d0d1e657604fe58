"""In-tree build of libzombie_amd.so (HIP for gfx950) and of the CPU oracle.

`build_library()` compiles every translation unit under zombie_amd/csrc with hipcc for
gfx950 and links the shared library next to this file, so it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
BUILD_DIR = PKG_DIR / "_build"
LIB_PATH = PKG_DIR / "libzombie_amd.so"
REPO = PKG_DIR.parent

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=gfx950",
    "-munsafe-fp-atomics",
    # packed-FP32 instructions stay off (known miscompare risk on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]

SOURCES = ["bvh_build.cpp", "abi.cpp", "wos_kernel.hip", "wost_kernel.hip", "mlp_kernels.hip", "probe_kernel.hip",
           "bvc_kernel.hip", "sdf_kernel.hip", "spatial_sort.cpp", "noalloc.hip"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build zombie_amd")


def _compile(src: Path, obj: Path, extra: list[str]) -> None:
    cmd = [_hipcc(), *HIP_FLAGS, *extra, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build_library(verbose: bool = False) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    objs = []
    jobs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for src in srcs:
            obj = BUILD_DIR / (src.stem + ".o")
            extra = ["-x", "hip"] if src.suffix == ".hip" else []
            objs.append(obj)
            jobs.append(ex.submit(_compile, src, obj, extra))
        for j in jobs:
            j.result()
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}", file=sys.stderr)
    return LIB_PATH


def build_oracle(verbose: bool = False) -> Path:
    """Builds the CPU oracle (test infrastructure) with its own Makefile."""
    oracle = REPO / "oracle"
    res = subprocess.run(["make", "-C", str(oracle), "-j4"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{res.stdout}\n{res.stderr}")
    lib = oracle / "_build" / "libzombie_oracle.so"
    if verbose:
        print(f"built {lib}", file=sys.stderr)
    return lib


ADAPTER_CHECK = BUILD_DIR / "adapter_check"


def build_adapter_check(verbose: bool = False, out: Path | None = None) -> Path:
    """Compiles tests/adapter_check.cpp, a C++ caller of the header-only adapter
    include/zombie_amd.hpp, against libzombie_amd.so (plain g++, no HIP headers)."""
    out = Path(out) if out is not None else ADAPTER_CHECK
    out.parent.mkdir(parents=True, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", str(REPO / "include"),
           str(REPO / "tests" / "adapter_check.cpp"), "-L", str(PKG_DIR), "-lzombie_amd",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + str(PKG_DIR), "-Wl,-rpath,/opt/rocm/lib", "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"adapter check build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {out}", file=sys.stderr)
    return out


if __name__ == "__main__":
    build_library(verbose=True)
    build_oracle(verbose=True)
    build_adapter_check(verbose=True)
