"""Build the mdmc shared library for gfx950 with hipcc (in-tree).

Output: framework_amd/lib/libmdmc.so.  Flags that matter for results:
  -ffp-contract=off                    no FMA contraction: every float + - * /
                                       is one correctly rounded operation, the
                                       sequence the CPU oracle also follows
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE float division
  -fno-gpu-flush-denormals-to-zero     keep denormal inputs as they are
  -target-feature -packed-fp32-ops     no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmdmc.so")
ARCH = os.environ.get("MDMC_ARCH", "gfx950")

SOURCES = [os.path.join(CSRC, "mdmc.hip"), os.path.join(CSRC, "mdmc_server.cpp"), os.path.join(CSRC, "mdmc_io.cpp")]
HEADERS = [os.path.join(CSRC, f) for f in ("mdmc_kernels.hpp", "mdmc_internal.h", "dmc_luts.h",
                                           "mdmc_fields.h", "mdmc_pool.h", "mdmc_io.h", "mdmc_tilemap.h", "mdmc_fmt.h")] + [os.path.join(ROOT, "include", "mdmc.h")]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile libmdmc.so (always recompiles)."""
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *SOURCES,
           "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    noise = [ln for ln in (res.stderr or "").splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "is not a recognized feature" not in ln]
    if verbose and noise:
        sys.stderr.write("\n".join(noise) + "\n")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
