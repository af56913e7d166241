"""Build the MI355X (gfx950) shared library libfip.so and the sharedlibrary_gpu CLI.

Compiles every HIP/C++ source under csrc/ with hipcc into csrc/../libfip.so
(in-tree, so it travels with the repository snapshot) and links the CLI
against it.  ``python -m fastimagingpipe_amd.build`` or ``build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfip.so")
CLI = os.path.join(ROOT, "sharedlibrary_gpu")

ARCH = os.environ.get("FIP_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for all device code (the compiler
# would otherwise form v_pk_* FP32 ops from float2 arithmetic).
COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-result",
          "-I", CSRC, "-I", INCLUDE]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-ffp-contract=fast"]

SOURCES = [
    "fip_kernels.hip",
    "fip_preprocess.hip",
    "fip_noalloc.hip",
    "fip_plan.cpp",
    "fip_context.cpp",
    "fip_api.cpp",
    "fip_synth.cpp",
    "fip_fits.cpp",
    "fip_serve.cpp",
    "fip_ingest.cpp",
]


def _sources():
    return [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "fip_cli.cpp")
    if os.path.exists(cli_src):
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, cli_src, "-o", CLI,
               "-L", HERE, "-lfip", "-Wl,-rpath,$ORIGIN/fastimagingpipe_amd"]
        subprocess.run(cmd, check=True)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
