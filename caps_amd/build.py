"""Build libcaps_amd.so (HIP kernels + C++ host engine) for gfx950 with hipcc.

    python -m caps_amd.build          # or caps_amd.build.build()

Objects go to caps_amd/_build/, the library to caps_amd/libcaps_amd.so (in
tree, so it travels with the source snapshot).  Every call recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcaps_amd.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources() -> list[str]:
    return sorted(os.path.join(SRC, f) for f in os.listdir(SRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_drivers()
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources, drivers in {BIN}")
    return LIB


BIN = os.path.join(HERE, "bin")
DRIVERS = os.path.join(HERE, "drivers")


def build_drivers() -> None:
    """The C++ drivers use only include/caps/*.h and the C ABI: plain g++."""
    os.makedirs(BIN, exist_ok=True)
    for f in sorted(os.listdir(DRIVERS)):
        if not f.endswith(".cpp"):
            continue
        out = os.path.join(BIN, f[:-4])
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(DRIVERS, f), "-o", out, "-L", HERE, "-lcaps_amd",
               "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed for {f}:\n{r.stderr}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
